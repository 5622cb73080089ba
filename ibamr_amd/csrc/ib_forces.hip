// ib_forces.hip -- Lagrangian force generation on the GPU: springs, beams
// (three-point bending) and target points, the forces IBStandardForceGen
// computes before the spread (src/IB/IBStandardForceGen.cpp:254-305; springs
// :814-930 with default_spring_force, include/ibamr/IBSpringForceFunctions.h:104,
// or another registered force function; beams :1038-1147; target points
// :1202-1299).
//
// The reference keeps a map force_fcn_idx -> function pointer
// (registerSpringForceFunction, :133-140) and resolves every spring's index to
// its function when the level's data is set up (:728-790).  Host function
// pointers cannot run on the GPU, so the map here holds built-in device laws
// (spring_tension below), resolved per spring the same way, when the spring
// parameters are set.
//
// The reference loops over springs, then beams, then target points, adding each
// element's force into its nodes in element order (a scatter into F_ghost,
// which starts at zero), and finally adds F_ghost to F.  Here every element's
// force is computed in parallel with the reference's arithmetic, then one
// thread per node gathers its entries from a node -> (element, role) list kept
// in exactly that order, so every node's force is the same sequence of
// floating-point operations as the reference loop: bitwise identical, and
// deterministic (no atomics).
#include "ibamr_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <map>
#include <string>
#include <vector>

namespace ibamr
{
int set_error(int code, const std::string& msg);
}

namespace
{
using ibamr::set_error;

enum Role : int
{
    R_SPRING_MASTER = 0,
    R_SPRING_SLAVE = 1,
    R_BEAM_MASTER = 2,
    R_BEAM_NEXT = 3,
    R_BEAM_PREV = 4,
    R_TARGET = 5
};

#define F_HIP(expr)                                                                               \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// Tension T(R) of spring law `law` with parameters p[0 .. np):
//   IBAMR_SPRING_HOOKEAN: p[0] (R - p[1]) (default_spring_force,
//     IBSpringForceFunctions.h:104-107);
//   IBAMR_SPRING_LINEAR_STRAIN: p[0] (R / p[1] - 1) (the linear_spring_force
//     examples/IB/explicit/ex3/example.cpp:58-63 registers);
//   IBAMR_SPRING_POLYNOMIAL: x (p[0] + x (p[2] + ... + x p[np-1])), x = R - p[1].
__device__ inline double spring_tension(int law, const double* __restrict__ p, int np, double R)
{
    if (law == IBAMR_SPRING_LINEAR_STRAIN) return p[0] * (R / p[1] - 1.0);
    if (law == IBAMR_SPRING_POLYNOMIAL)
    {
        const double x = R - p[1];
        double h = np > 2 ? p[np - 1] : p[0];
        for (int m = np - 2; m >= 2; --m) h = p[m] + x * h;
        if (np > 2) h = p[0] + x * h;
        return x * h;
    }
    return p[0] * (R - p[1]);
}

// Spring k: D = X_slave - X_master, R = |D|, T/R = T(R) / R, F = (T/R) D;
// skipped when R < DBL_EPSILON (IBStandardForceGen.cpp:903-911).
__global__ void k_spring_elem(int ns, int ndim, const int* __restrict__ m, const int* __restrict__ sl,
                              const int* __restrict__ law, const int* __restrict__ poff,
                              const double* __restrict__ par, const double* __restrict__ X, double* __restrict__ Fs,
                              unsigned char* __restrict__ skip)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const int a = m[k], b = sl[k];
    double D[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < ndim; ++d) D[d] = X[(size_t)ndim * b + d] - X[(size_t)ndim * a + d];
    double R2 = D[0] * D[0] + D[1] * D[1];
    if (ndim == 3) R2 = R2 + D[2] * D[2];
    const double R = sqrt(R2);
    const bool sk = R < DBL_EPSILON;
    skip[k] = sk ? 1 : 0;
    const int o = poff[k];
    const double T_over_R = sk ? 0.0 : spring_tension(law[k], par + o, poff[k + 1] - o, R) / R;
    for (int d = 0; d < ndim; ++d) Fs[(size_t)ndim * k + d] = T_over_R * D[d];
}

// Beam k: F = K (X_next + X_prev - 2 X_master - curvature) (IBStandardForceGen.cpp:1129-1131)
__global__ void k_beam_elem(int nb, int ndim, const int* __restrict__ m, const int* __restrict__ nx,
                            const int* __restrict__ pv, const double* __restrict__ K, const double* __restrict__ curv,
                            const double* __restrict__ X, double* __restrict__ Fb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const size_t im = (size_t)ndim * m[k], in = (size_t)ndim * nx[k], ip = (size_t)ndim * pv[k];
    for (int d = 0; d < ndim; ++d)
        Fb[(size_t)ndim * k + d] = K[k] * (((X[in + d] + X[ip + d]) - 2.0 * X[im + d]) - curv[(size_t)ndim * k + d]);
}

template <typename TF, typename TU>
__global__ void k_node_gather(int n, int ndim, const int* __restrict__ start, const int2* __restrict__ ent,
                              const double* __restrict__ Fs, const unsigned char* __restrict__ skip,
                              const double* __restrict__ Fb, const int* __restrict__ t_node,
                              const double* __restrict__ t_K, const double* __restrict__ t_E,
                              const double* __restrict__ t_X0, const double* __restrict__ X,
                              const TU* __restrict__ U, TF* __restrict__ F, int overwrite)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int e = start[v]; e < start[v + 1]; ++e)
    {
        const int2 en = ent[e];
        const int k = en.y;
        switch (en.x)
        {
        case R_SPRING_MASTER:
            if (!skip[k])
                for (int d = 0; d < ndim; ++d) acc[d] = acc[d] + Fs[(size_t)ndim * k + d];
            break;
        case R_SPRING_SLAVE:
            if (!skip[k])
                for (int d = 0; d < ndim; ++d) acc[d] = acc[d] - Fs[(size_t)ndim * k + d];
            break;
        case R_BEAM_MASTER:
            for (int d = 0; d < ndim; ++d) acc[d] = acc[d] + 2.0 * Fb[(size_t)ndim * k + d];
            break;
        case R_BEAM_NEXT:
        case R_BEAM_PREV:
            for (int d = 0; d < ndim; ++d) acc[d] = acc[d] - Fb[(size_t)ndim * k + d];
            break;
        default:  // target point (IBStandardForceGen.cpp:1285-1288)
            for (int d = 0; d < ndim; ++d)
            {
                const double u = U ? (double)U[(size_t)ndim * v + d] : 0.0;
                acc[d] = acc[d] + (t_K[k] * (t_X0[(size_t)ndim * k + d] - X[(size_t)ndim * v + d]) - t_E[k] * u);
            }
            break;
        }
    }
    // F += F_ghost (VecAXPY with alpha 1, :300), or F = F_ghost
    for (int d = 0; d < ndim; ++d)
    {
        const size_t i = (size_t)ndim * v + d;
        F[i] = overwrite ? (TF)acc[d] : (TF)((double)F[i] + acc[d]);
    }
}

template <typename T>
int dev_copy(T** dst, const T* src, size_t n)
{
    *dst = nullptr;
    if (n == 0) return IBAMR_OK;
    F_HIP(hipMalloc(dst, sizeof(T) * n));
    F_HIP(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return IBAMR_OK;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}
} // namespace

struct ibamr_forces
{
    int ndim = 3, n = 0, ns = 0, nb = 0, nt = 0;
    int *s_m = nullptr, *s_s = nullptr;
    int *s_law = nullptr, *s_off = nullptr;  // per spring: law, first parameter ([ns + 1])
    double* s_par = nullptr;
    std::map<int, int> registry{{0, IBAMR_SPRING_HOOKEAN}};  // force function index -> law
    int *b_m = nullptr, *b_n = nullptr, *b_p = nullptr;
    double *b_K = nullptr, *b_c = nullptr;
    int* t_i = nullptr;
    double *t_K = nullptr, *t_E = nullptr, *t_X0 = nullptr;
    int* start = nullptr;
    int2* ent = nullptr;
    double *Fs = nullptr, *Fb = nullptr;
    unsigned char* skip = nullptr;
};

extern "C" {

int ibamr_forces_destroy(ibamr_forces* f)
{
    if (!f) return IBAMR_OK;
    for (void* p : {(void*)f->s_m, (void*)f->s_s, (void*)f->s_law, (void*)f->s_off, (void*)f->s_par, (void*)f->b_m, (void*)f->b_n,
                    (void*)f->b_p, (void*)f->b_K, (void*)f->b_c, (void*)f->t_i, (void*)f->t_K, (void*)f->t_E,
                    (void*)f->t_X0, (void*)f->start, (void*)f->ent, (void*)f->Fs, (void*)f->Fb, (void*)f->skip})
        dfree(p);
    delete f;
    return IBAMR_OK;
}

int ibamr_forces_create(ibamr_forces** out, int ndim, int npoints, int nsprings, const int* spring_master,
                        const int* spring_slave, const double* spring_kappa, const double* spring_rest, int nbeams,
                        const int* beam_master, const int* beam_next, const int* beam_prev,
                        const double* beam_rigidity, const double* beam_curvature, int ntargets,
                        const int* target_node, const double* target_kappa, const double* target_eta,
                        const double* target_X0)
{
    if (!out) return set_error(IBAMR_ERR_INVALID, "null argument");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return set_error(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (npoints < 0 || nsprings < 0 || nbeams < 0 || ntargets < 0) return set_error(IBAMR_ERR_INVALID, "negative count");
    if ((nsprings > 0 && (!spring_master || !spring_slave || !spring_kappa || !spring_rest)) ||
        (nbeams > 0 && (!beam_master || !beam_next || !beam_prev || !beam_rigidity || !beam_curvature)) ||
        (ntargets > 0 && (!target_node || !target_kappa || !target_eta || !target_X0)))
        return set_error(IBAMR_ERR_INVALID, "null element array");
    auto bad = [&](int i) { return i < 0 || i >= npoints; };
    // node -> (role, element) lists in the reference's order: springs, beams,
    // target points, each in element order
    std::vector<std::vector<int2>> lists((size_t)npoints);
    for (int k = 0; k < nsprings; ++k)
    {
        if (bad(spring_master[k]) || bad(spring_slave[k]) || spring_master[k] == spring_slave[k])
            return set_error(IBAMR_ERR_INVALID, "spring " + std::to_string(k) + ": bad node index");
        lists[spring_master[k]].push_back(make_int2(R_SPRING_MASTER, k));
        lists[spring_slave[k]].push_back(make_int2(R_SPRING_SLAVE, k));
    }
    for (int k = 0; k < nbeams; ++k)
    {
        if (bad(beam_master[k]) || bad(beam_next[k]) || bad(beam_prev[k]) || beam_master[k] == beam_next[k] ||
            beam_master[k] == beam_prev[k])
            return set_error(IBAMR_ERR_INVALID, "beam " + std::to_string(k) + ": bad node index");
        lists[beam_master[k]].push_back(make_int2(R_BEAM_MASTER, k));
        lists[beam_next[k]].push_back(make_int2(R_BEAM_NEXT, k));
        lists[beam_prev[k]].push_back(make_int2(R_BEAM_PREV, k));
    }
    for (int k = 0; k < ntargets; ++k)
    {
        if (bad(target_node[k])) return set_error(IBAMR_ERR_INVALID, "target point " + std::to_string(k) + ": bad node index");
        lists[target_node[k]].push_back(make_int2(R_TARGET, k));
    }
    std::vector<int> start((size_t)npoints + 1, 0);
    std::vector<int2> ent;
    for (int v = 0; v < npoints; ++v)
    {
        start[v] = (int)ent.size();
        ent.insert(ent.end(), lists[v].begin(), lists[v].end());
    }
    start[npoints] = (int)ent.size();

    ibamr_forces* f = new ibamr_forces();
    f->ndim = ndim;
    f->n = npoints;
    f->ns = nsprings;
    f->nb = nbeams;
    f->nt = ntargets;
    int rc = IBAMR_OK;
    auto chk = [&](int r) {
        if (rc == IBAMR_OK) rc = r;
    };
    chk(dev_copy(&f->s_m, spring_master, (size_t)nsprings));
    chk(dev_copy(&f->s_s, spring_slave, (size_t)nsprings));
    {
        // the default force function: law IBAMR_SPRING_HOOKEAN with {kappa, rest}
        std::vector<int> law((size_t)nsprings, IBAMR_SPRING_HOOKEAN), off((size_t)nsprings + 1);
        std::vector<double> par((size_t)nsprings * 2);
        for (int k = 0; k <= nsprings; ++k) off[k] = 2 * k;
        for (int k = 0; k < nsprings; ++k)
        {
            par[2 * (size_t)k] = spring_kappa[k];
            par[2 * (size_t)k + 1] = spring_rest[k];
        }
        chk(dev_copy(&f->s_law, law.data(), law.size()));
        chk(dev_copy(&f->s_off, off.data(), nsprings > 0 ? off.size() : 0));
        chk(dev_copy(&f->s_par, par.data(), par.size()));
    }
    chk(dev_copy(&f->b_m, beam_master, (size_t)nbeams));
    chk(dev_copy(&f->b_n, beam_next, (size_t)nbeams));
    chk(dev_copy(&f->b_p, beam_prev, (size_t)nbeams));
    chk(dev_copy(&f->b_K, beam_rigidity, (size_t)nbeams));
    chk(dev_copy(&f->b_c, beam_curvature, (size_t)nbeams * ndim));
    chk(dev_copy(&f->t_i, target_node, (size_t)ntargets));
    chk(dev_copy(&f->t_K, target_kappa, (size_t)ntargets));
    chk(dev_copy(&f->t_E, target_eta, (size_t)ntargets));
    chk(dev_copy(&f->t_X0, target_X0, (size_t)ntargets * ndim));
    chk(dev_copy(&f->start, start.data(), start.size()));
    chk(dev_copy(&f->ent, ent.data(), ent.size()));
    if (rc == IBAMR_OK && nsprings > 0)
    {
        if (hipMalloc(&f->Fs, sizeof(double) * (size_t)nsprings * ndim) != hipSuccess ||
            hipMalloc(&f->skip, (size_t)nsprings) != hipSuccess)
            rc = set_error(IBAMR_ERR_DEVICE, "hipMalloc of spring scratch failed");
    }
    if (rc == IBAMR_OK && nbeams > 0 && hipMalloc(&f->Fb, sizeof(double) * (size_t)nbeams * ndim) != hipSuccess)
        rc = set_error(IBAMR_ERR_DEVICE, "hipMalloc of beam scratch failed");
    if (rc != IBAMR_OK)
    {
        ibamr_forces_destroy(f);
        return rc;
    }
    *out = f;
    return IBAMR_OK;
}

int ibamr_forces_register_spring_function(ibamr_forces* f, int force_fcn_index, int law)
{
    if (!f) return set_error(IBAMR_ERR_INVALID, "null force object");
    if (law != IBAMR_SPRING_HOOKEAN && law != IBAMR_SPRING_LINEAR_STRAIN && law != IBAMR_SPRING_POLYNOMIAL)
        return set_error(IBAMR_ERR_INVALID, "unknown spring law " + std::to_string(law));
    f->registry[force_fcn_index] = law;
    return IBAMR_OK;
}

int ibamr_forces_set_spring_parameters(ibamr_forces* f, const int* force_fcn_idx, const int* param_offsets,
                                       const double* params)
{
    if (!f) return set_error(IBAMR_ERR_INVALID, "null force object");
    const int ns = f->ns;
    if (ns == 0) return IBAMR_OK;
    if (!force_fcn_idx || !param_offsets || (!params && param_offsets[ns] > 0))
        return set_error(IBAMR_ERR_INVALID, "null spring parameter array");
    if (param_offsets[0] != 0) return set_error(IBAMR_ERR_INVALID, "spring parameter offsets must start at 0");
    std::vector<int> law((size_t)ns);
    for (int k = 0; k < ns; ++k)
    {
        const auto it = f->registry.find(force_fcn_idx[k]);
        if (it == f->registry.end())
            return set_error(IBAMR_ERR_INVALID, "spring " + std::to_string(k) + ": no spring force function is "
                                                "registered under index " + std::to_string(force_fcn_idx[k]));
        const int np = param_offsets[k + 1] - param_offsets[k];
        if (np < 2)
            return set_error(IBAMR_ERR_INVALID, "spring " + std::to_string(k) + ": " + std::to_string(np) +
                                                    " parameters (at least stiffness and rest length)");
        law[k] = it->second;
    }
    int* d_law = nullptr;
    int* d_off = nullptr;
    double* d_par = nullptr;
    int rc = dev_copy(&d_law, law.data(), law.size());
    if (rc == IBAMR_OK) rc = dev_copy(&d_off, param_offsets, (size_t)ns + 1);
    if (rc == IBAMR_OK) rc = dev_copy(&d_par, params, (size_t)param_offsets[ns]);
    if (rc != IBAMR_OK)
    {
        dfree(d_law);
        dfree(d_off);
        dfree(d_par);
        return rc;
    }
    dfree(f->s_law);
    dfree(f->s_off);
    dfree(f->s_par);
    f->s_law = d_law;
    f->s_off = d_off;
    f->s_par = d_par;
    return IBAMR_OK;
}

int ibamr_forces_compute(ibamr_forces* f, void* F, int f_precision, const double* X, const void* U,
                         int u_precision, int mode, void* stream)
{
    if (!f) return set_error(IBAMR_ERR_INVALID, "null force object");
    if (f->n == 0) return IBAMR_OK;
    if (!F || !X) return set_error(IBAMR_ERR_INVALID, "null F or X");
    if (f_precision != IBAMR_FP64 && f_precision != IBAMR_FP32C) return set_error(IBAMR_ERR_INVALID, "bad precision");
    if (U && u_precision != IBAMR_FP64 && u_precision != IBAMR_FP32C)
        return set_error(IBAMR_ERR_INVALID, "bad velocity precision");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE) return set_error(IBAMR_ERR_INVALID, "bad mode");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    if (f->ns > 0)
    {
        hipLaunchKernelGGL(k_spring_elem, dim3((f->ns + B - 1) / B), dim3(B), 0, s, f->ns, f->ndim, f->s_m, f->s_s,
                           f->s_law, f->s_off, f->s_par, X, f->Fs, f->skip);
        F_HIP(hipGetLastError());
    }
    if (f->nb > 0)
    {
        hipLaunchKernelGGL(k_beam_elem, dim3((f->nb + B - 1) / B), dim3(B), 0, s, f->nb, f->ndim, f->b_m, f->b_n,
                           f->b_p, f->b_K, f->b_c, X, f->Fb);
        F_HIP(hipGetLastError());
    }
    const int ow = mode == IBAMR_SPREAD_OVERWRITE ? 1 : 0;
    const dim3 grid((f->n + B - 1) / B);
    const bool u32 = U && u_precision == IBAMR_FP32C;
#define GATHER(TF, TU)                                                                                          \
    hipLaunchKernelGGL((k_node_gather<TF, TU>), grid, dim3(B), 0, s, f->n, f->ndim, f->start, f->ent, f->Fs,     \
                       f->skip, f->Fb, f->t_i, f->t_K, f->t_E, f->t_X0, X, static_cast<const TU*>(U),             \
                       static_cast<TF*>(F), ow)
    if (f_precision == IBAMR_FP64)
    {
        if (u32) GATHER(double, float);
        else GATHER(double, double);
    }
    else
    {
        if (u32) GATHER(float, float);
        else GATHER(float, double);
    }
#undef GATHER
    F_HIP(hipGetLastError());
    return IBAMR_OK;
}

} // extern "C"
