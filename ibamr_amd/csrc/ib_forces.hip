// ib_forces.hip -- Lagrangian force generation on the GPU: Hookean springs,
// beams (three-point bending) and target points, the forces IBStandardForceGen
// computes before the spread (src/IB/IBStandardForceGen.cpp:254-305; springs
// :814-930 with default_spring_force, include/ibamr/IBSpringForceFunctions.h:104;
// beams :1038-1147; target points :1202-1299).
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

// Spring k: D = X_slave - X_master, R = |D|, T/R = kappa (R - rest) / R,
// F = (T/R) D; skipped when R < DBL_EPSILON (IBStandardForceGen.cpp:903-911).
__global__ void k_spring_elem(int ns, int ndim, const int* __restrict__ m, const int* __restrict__ sl,
                              const double* __restrict__ kappa, const double* __restrict__ rest,
                              const double* __restrict__ X, double* __restrict__ Fs, unsigned char* __restrict__ skip)
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
    const double T_over_R = sk ? 0.0 : (kappa[k] * (R - rest[k])) / R;
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
    double *s_k = nullptr, *s_r = nullptr;
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
    for (void* p : {(void*)f->s_m, (void*)f->s_s, (void*)f->s_k, (void*)f->s_r, (void*)f->b_m, (void*)f->b_n,
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
    chk(dev_copy(&f->s_k, spring_kappa, (size_t)nsprings));
    chk(dev_copy(&f->s_r, spring_rest, (size_t)nsprings));
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
                           f->s_k, f->s_r, X, f->Fs, f->skip);
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
