// interpolate_01.cpp -- the reference's tests/interpolate/interpolate_01.cpp
// (3-D runs) restated on the ibtk_amd C++ layer: 100 points drawn with
// std::mt19937(42) and std::uniform_real_distribution over the fine patch
// (:160-177), the trilinear field of interpolate_01_3d.*.input at cell centres
// (ghosts included: SAMRAI's LINEAR refine is exact for it), then
// LEInteractor::interpolate over the patch box (:181) and the output format
// of :186-230.  Also a raw-array spread into side data and an LDataManager /
// IBMethod round trip on a periodic level, printed for the oracle comparison.
//
//   interpolate_01 interp <IB_4|IB_6|BSPLINE_3> <output>
//   interpolate_01 side   <kernel> <output>     raw-array side interp + spread
//   interpolate_01 level  <kernel> <output>     LDataManager spread/interp, forwardEulerStep
//   interpolate_01 regrid <kernel> <output>     IBStandardForceGen (springs under force functions 0,
//                                               1 and 2: Hookean, linear strain, polynomial), then
//                                               LDataManager::redistribute
#include "ibtk_amd.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <random>
#include <string>
#include <vector>

using namespace ibtk_amd;

namespace
{
const int NDIM = 3;

double field(double x, double y, double z) { return 1 + 2 * x + 3 * y - z + 4 * x * y + 2 * x * z + 3 * x * y * z; }

std::shared_ptr<Patch> fine_patch()
{
    // level 1 of interpolate_01_3d: coarse 8^3 on [0,1]^3, ratio 4, refine box (2..3)^3
    const double dx[3] = {1.0 / 32, 1.0 / 32, 1.0 / 32};
    const double xlo[3] = {0.25, 0.25, 0.25}, xhi[3] = {0.5, 0.5, 0.5};
    return std::make_shared<Patch>(Box(NDIM, Index(8, 8, 8), Index(15, 15, 15)),
                                   CartesianPatchGeometry(NDIM, dx, xlo, xhi));
}

std::vector<double> random_points(const Patch& patch, int n)
{
    std::mt19937 std_seq(42u);
    std::uniform_real_distribution<double> distribution(patch.getPatchGeometry().getXLower()[0],
                                                        patch.getPatchGeometry().getXUpper()[0]);
    std::vector<double> X(NDIM * n);
    for (double& v : X) v = distribution(std_seq);
    return X;
}

int run_interp(const std::string& k, const std::string& out_name)
{
    auto patch = fine_patch();
    const int gw = LEInteractor::getMinimumGhostWidth(k);
    auto q_data = std::make_shared<CellData<double>>(patch->getBox(), NDIM, IntVector(gw));
    const Box g = q_data->getGhostBox();
    const double* dx = patch->getPatchGeometry().getDx();
    const double* xl = patch->getPatchGeometry().getXLower();
    std::vector<double> h(q_data->components() * NDIM);
    size_t e = 0;
    for (int c = 0; c < NDIM; ++c)
        for (int k2 = g.lower()(2); k2 <= g.upper()(2); ++k2)
            for (int k1 = g.lower()(1); k1 <= g.upper()(1); ++k1)
                for (int k0 = g.lower()(0); k0 <= g.upper()(0); ++k0)
                    h[e++] = field(xl[0] + (k0 - 8 + 0.5) * dx[0], xl[1] + (k1 - 8 + 0.5) * dx[1],
                                   xl[2] + (k2 - 8 + 0.5) * dx[2]);
    q_data->copyFromHost(h);

    const std::size_t n_points = 100;
    const int Q_depth = NDIM;
    std::vector<double> Q_data(Q_depth * n_points, std::numeric_limits<double>::max());
    const int X_depth = NDIM;
    std::vector<double> X_data = random_points(*patch, n_points);
    LEInteractor::interpolate(Q_data, Q_depth, X_data, X_depth, q_data, patch, patch->getBox(), k);

    std::ofstream out(out_name);
    out.precision(16);
    out << "x, y, z, u0, u1, u2, e0, e1, e2\n";
    for (std::size_t p = 0; p < n_points; ++p)
        out << X_data[p * NDIM + 0] << ", " << X_data[p * NDIM + 1] << ", " << X_data[p * NDIM + 2] << ", "
            << Q_data[p * NDIM + 0] << ", " << Q_data[p * NDIM + 1] << ", " << Q_data[p * NDIM + 2] << '\n';
    return 0;
}

// side data on the same patch: interp of an analytic face field over the
// patch box grown by one cell, then spread of per-point forces over the patch box
int run_side(const std::string& k, const std::string& out_name)
{
    auto patch = fine_patch();
    const int gw = LEInteractor::getMinimumGhostWidth(k);
    auto u = std::make_shared<SideData<double>>(patch->getBox(), 1, IntVector(gw));
    auto f = std::make_shared<SideData<double>>(patch->getBox(), 1, IntVector(gw));
    const double* dx = patch->getPatchGeometry().getDx();
    const double* xl = patch->getPatchGeometry().getXLower();
    for (int a = 0; a < NDIM; ++a)
    {
        std::vector<double> h(u->components(a));
        const int lo = 8 - gw, n0 = 8 + 2 * gw + (a == 0), n1 = 8 + 2 * gw + (a == 1), n2 = 8 + 2 * gw + (a == 2);
        size_t e = 0;
        for (int k2 = 0; k2 < n2; ++k2)
            for (int k1 = 0; k1 < n1; ++k1)
                for (int k0 = 0; k0 < n0; ++k0)
                {
                    const double x = xl[0] + (lo + k0 - 8 + (a == 0 ? 0.0 : 0.5)) * dx[0];
                    const double y = xl[1] + (lo + k1 - 8 + (a == 1 ? 0.0 : 0.5)) * dx[1];
                    const double z = xl[2] + (lo + k2 - 8 + (a == 2 ? 0.0 : 0.5)) * dx[2];
                    h[e++] = std::sin(7 * x + a) * std::cos(5 * y - 2 * z) + 0.25 * a;
                }
        u->copyFromHost(a, h);
    }
    f->fillAll(0.0);
    const int n = 200;
    std::mt19937 seq(7u);
    std::uniform_real_distribution<double> dist(0.18, 0.57), fdist(-1.0, 1.0);
    std::vector<double> X(NDIM * n), F(NDIM * n), U(NDIM * n, -7.0);
    for (double& v : X) v = dist(seq);
    for (double& v : F) v = fdist(seq);
    LEInteractor::interpolate(U, NDIM, X, NDIM, u, patch, patch->getBox().grown(IntVector(1)), k);
    LEInteractor::spread(f, F, NDIM, X, NDIM, patch, patch->getBox(), k);
    std::ofstream out(out_name);
    out.precision(17);
    out << "# X, F, U (grown-box interp), then f and u per axis (ghost box, x fastest)\n";
    for (int p = 0; p < n; ++p)
        out << X[3 * p] << ' ' << X[3 * p + 1] << ' ' << X[3 * p + 2] << ' ' << F[3 * p] << ' ' << F[3 * p + 1] << ' '
            << F[3 * p + 2] << ' ' << U[3 * p] << ' ' << U[3 * p + 1] << ' ' << U[3 * p + 2] << '\n';
    for (int a = 0; a < NDIM; ++a)
    {
        std::vector<double> h;
        f->copyToHost(a, h);
        out << "f" << a;
        for (double v : h) out << ' ' << v;
        out << '\n';
        u->copyToHost(a, h);
        out << "u" << a;
        for (double v : h) out << ' ' << v;
        out << '\n';
    }
    return 0;
}

// a periodic 16^3 single-patch level: IBMethod::spreadForce into f (which
// holds 1 beforehand), interpolateVelocity from a face field, forwardEulerStep
int run_level(const std::string& k, const std::string& out_name)
{
    const int N = 16;
    Level L;
    L.ndim = NDIM;
    L.dx = {{1.0 / N, 1.0 / N, 1.0 / N}};
    L.domain = Box(NDIM, Index(0, 0, 0), Index(N - 1, N - 1, N - 1));
    L.x_lower = {{0, 0, 0}};
    L.x_upper = {{1, 1, 1}};
    L.periodic_shift = IntVector(N, N, N);
    const double dx[3] = {1.0 / N, 1.0 / N, 1.0 / N}, xlo[3] = {0, 0, 0}, xhi[3] = {1, 1, 1};
    auto patch = std::make_shared<Patch>(L.domain, CartesianPatchGeometry(NDIM, dx, xlo, xhi));
    L.patches.push_back(patch);
    const int gw = LEInteractor::getMinimumGhostWidth(k);
    const int n = 300;
    auto ldm = std::make_shared<LDataManager>(L, k, k, IntVector(gw), n);
    IBMethod ib(ldm);
    std::vector<std::shared_ptr<SideData<double>>> f{std::make_shared<SideData<double>>(L.domain, 1, IntVector(gw))};
    std::vector<std::shared_ptr<SideData<double>>> u{std::make_shared<SideData<double>>(L.domain, 1, IntVector(gw))};
    f[0]->fillAll(1.0);
    for (int a = 0; a < NDIM; ++a)
    {
        std::vector<double> h(u[0]->components(a));
        for (size_t e = 0; e < h.size(); ++e) h[e] = std::cos(0.37 * (double)e + a);
        u[0]->copyFromHost(a, h);
    }
    std::mt19937 seq(11u);
    std::uniform_real_distribution<double> dist(0.0, 1.0), fdist(-1.0, 1.0);
    std::vector<double> Xh(NDIM * n), Fh(NDIM * n);
    for (double& v : Xh) v = dist(seq);
    for (double& v : Fh) v = fdist(seq);
    LData<double> X(n, NDIM), F(n, NDIM), U(n, NDIM), X_new(n, NDIM);
    X.copyFromHost(Xh);
    F.copyFromHost(Fh);
    ldm->endDataRedistribution(X);
    ib.spreadForce(f, F, X);
    ib.interpolateVelocity(u, U, X);  // same-level ghost fill (periodic wrap), then interp
    IBMethod::forwardEulerStep(X_new, X, U, 0.01);
    std::vector<double> Uh, Xn;
    U.copyToHost(Uh);
    X_new.copyToHost(Xn);
    std::ofstream out(out_name);
    out.precision(17);
    out << "# X, F, U, X_new; then f and u per axis (ghost box, x fastest)\n";
    for (int p = 0; p < n; ++p)
    {
        for (int d = 0; d < 3; ++d) out << Xh[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Fh[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Uh[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Xn[3 * p + d] << (d == 2 ? '\n' : ' ');
    }
    for (int a = 0; a < NDIM; ++a)
    {
        std::vector<double> h;
        f[0]->copyToHost(a, h);
        out << "f" << a;
        for (double v : h) out << ' ' << v;
        out << '\n';
    }
    for (int a = 0; a < NDIM; ++a)
    {
        std::vector<double> h;
        u[0]->copyToHost(a, h);
        out << "u" << a;
        for (double v : h) out << ' ' << v;
        out << '\n';
    }
    return 0;
}
int run_regrid(const std::string& k, const std::string& out_name)
{
    // periodic 16^3 level in two patches (x halves); 400 points, some outside the box
    const int N = 16, n = 400;
    Level L;
    L.ndim = NDIM;
    L.dx = {{1.0 / N, 1.0 / N, 1.0 / N}};
    L.domain = Box(NDIM, Index(0, 0, 0), Index(N - 1, N - 1, N - 1));
    L.x_lower = {{0, 0, 0}};
    L.x_upper = {{1, 1, 1}};
    L.periodic_shift = IntVector(N, N, N);
    for (int q = 0; q < 2; ++q)
    {
        const double dx[3] = {1.0 / N, 1.0 / N, 1.0 / N}, xlo[3] = {0.5 * q, 0, 0}, xhi[3] = {0.5 * (q + 1), 1, 1};
        L.patches.push_back(std::make_shared<Patch>(Box(NDIM, Index(8 * q, 0, 0), Index(8 * q + 7, N - 1, N - 1)),
                                                    CartesianPatchGeometry(NDIM, dx, xlo, xhi)));
    }
    const int gw = LEInteractor::getMinimumGhostWidth(k);
    auto ldm = std::make_shared<LDataManager>(L, k, k, IntVector(gw), n);
    std::mt19937 seq(5u);
    std::uniform_real_distribution<double> dist(-0.3, 1.3), small(-0.01, 0.01);
    std::vector<double> Xh(NDIM * n), Uh(NDIM * n);
    for (double& v : Xh) v = dist(seq);
    for (double& v : Uh) v = small(seq);
    std::vector<Spring> springs;
    std::vector<Beam> beams;
    std::vector<TargetPoint> targets;
    for (int i = 0; i < n; ++i)
    {
        Spring s{i, (i + 1) % n, 1.0 + 0.01 * i, 0.05};
        s.force_fcn_idx = i % 3;
        if (i % 3 == 2) s.extra_parameters = {30.0, -40.0};
        springs.push_back(s);
        beams.push_back({(i + n - 1) % n, i, (i + 1) % n, 0.5, {{0.001, -0.002, 0.0}}});
        if (i % 7 == 0) targets.push_back({i, 3.0, 0.5, {{0.5, 0.5, 0.5}}});
    }
    IBStandardForceGen gen(NDIM, n, springs, beams, targets);
    gen.registerSpringForceFunction(1, IBAMR_SPRING_LINEAR_STRAIN);
    gen.registerSpringForceFunction(2, IBAMR_SPRING_POLYNOMIAL);
    LData<double> X(n, NDIM), F(n, NDIM), U(n, NDIM);
    LData<int> lag(n, 1);
    std::vector<int> lh(n);
    for (int i = 0; i < n; ++i) lh[i] = n - 1 - i;
    X.copyFromHost(Xh);
    U.copyFromHost(Uh);
    F.copyFromHost(std::vector<double>(NDIM * n, 0.0));
    lag.copyFromHost(lh);
    gen.computeLagrangianForce(F, X, U);
    const std::vector<int> offsets = ldm->redistribute(X, {&F, &U}, &lag);
    std::vector<double> Xa, Fa, Ua;
    std::vector<int> la;
    X.copyToHost(Xa);
    F.copyToHost(Fa);
    U.copyToHost(Ua);
    lag.copyToHost(la);
    std::ofstream out(out_name);
    out.precision(17);
    out << "# X_before U_before, then after redistribution: X F U lag; offsets\n";
    for (int p = 0; p < n; ++p)
    {
        for (int d = 0; d < 3; ++d) out << Xh[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Uh[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Xa[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Fa[3 * p + d] << ' ';
        for (int d = 0; d < 3; ++d) out << Ua[3 * p + d] << ' ';
        out << la[p] << '\n';
    }
    out << "offsets";
    for (int v : offsets) out << ' ' << v;
    out << '\n';
    return 0;
}
} // namespace

int main(int argc, char** argv)
{
    if (argc != 4)
    {
        std::cerr << "usage: " << argv[0] << " interp|side|level|regrid <kernel> <output>\n";
        return 2;
    }
    const std::string mode = argv[1], k = argv[2], out = argv[3];
    try
    {
        if (mode == "interp") return run_interp(k, out);
        if (mode == "side") return run_side(k, out);
        if (mode == "level") return run_level(k, out);
        if (mode == "regrid") return run_regrid(k, out);
        std::cerr << "unknown mode " << mode << "\n";
        return 2;
    }
    catch (const Error& e)
    {
        std::cerr << e.what() << "\n";
        return 1;
    }
}
