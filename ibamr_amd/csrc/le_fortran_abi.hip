// le_fortran_abi.hip -- link-level drop-in for IBTK's Fortran spreading and
// interpolation kernels (IB_4, IB_6, BSPLINE_3; 2-D and 3-D).  See
// include/ibamr_amd_fortran.h for the contract.
//
// interp: one thread per list entry, the reference's per-entry arithmetic
//         (f.m4 IB_4 :1203-1334, IB_6 :2178-2375, BSPLINE_3 :2591-2693).  When
//         an index occurs more than once, the last entry wins, as in the
//         reference's sequential loop.
// spread: every (entry, depth, stencil point) contribution is formed with the
//         reference's operation order (f.m4 :1432-1465, :51-64), keyed by its
//         grid index, stably sorted by key (list order kept within a key), and
//         each grid value then adds its contributions sequentially: u is
//         updated exactly as the reference loop over l updates it.
#include "ibamr_amd.h"
#include "ibamr_amd_fortran.h"
#include "le_weights.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <map>
#include <mutex>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace ibamr
{
int set_error(int code, const std::string& msg);  // le_plan.hip
}

namespace
{
using namespace ibamr;

constexpr unsigned S0_SENTINEL = 0xFFFFFFFFu;
constexpr size_t S0_MAX_TUPLES = size_t(1) << 25;  // contributions per spread batch

struct S0Geom
{
    int ndim;
    int depth;
    double dx[3];
    double x_lower[3];
    int glo[3];   // ghost box lower (ilower - nugc)
    int ghi[3];   // ghost box upper (iupper + nugc)
    int ilo[3];   // ilower
    int n[3];     // ghost box extents
};

// Xs = X(d,s) + Xshift(d,l); Xu = X(d,s) (BSPLINE_4/6 choose their stencil side with it)
template <int NDIM, int KER>
__device__ __forceinline__ void s0_stencil(const S0Geom& g, const double* Xs, const double* Xu, int* bse,
                                           double (*w)[KernelTraits<KER>::W])
{
#pragma unroll
    for (int d = 0; d < NDIM; ++d) bse[d] = g.ilo[d] + weights_1d_u<KER>(Xs[d], Xu[d], g.x_lower[d], g.dx[d], w[d]);
}

// last[s] = largest list position with index s
// (entries whose index lies outside [0, npts) take no part)
__global__ void k_s0_last(const int* __restrict__ indices, int n, int npts, int* __restrict__ last)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int s = indices[l];
    if (s >= 0 && s < npts) atomicMax(&last[s], l);
}

// last[0..n) = -1 (a kernel, not a memset: the allocation-free form may be
// captured in a HIP graph)
__global__ void k_s0_clear(int* __restrict__ last, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) last[i] = -1;
}

template <int NDIM, int KER>
__global__ void k_s0_interp(S0Geom g, const double* __restrict__ u, const int* __restrict__ indices,
                            const double* __restrict__ Xshift, int n, int npts, const double* __restrict__ X,
                            double* __restrict__ V, const int* __restrict__ last)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int s = indices[l];
    if (s < 0 || s >= npts) return;
    if (last[s] != l) return;  // a later entry with the same index overwrites V(:, s)
    double Xs[3];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) Xs[d] = X[NDIM * (size_t)s + d] + Xshift[NDIM * (size_t)l + d];
    int bse[3];
    double w[3][W];
    s0_stencil<NDIM, KER>(g, Xs, X + NDIM * (size_t)s, bse, w);
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
    {
        lo[d] = max(bse[d], g.glo[d]) - bse[d];
        hi[d] = min(bse[d] + W - 1, g.ghi[d]) - bse[d];
    }
    const size_t ncomp = (size_t)g.n[0] * g.n[1] * (NDIM == 3 ? g.n[2] : 1);
    for (int c = 0; c < g.depth; ++c)
    {
        const double* uc = u + c * ncomp;
        double v = 0.0;
        if (NDIM == 3)
        {
            for (int i2 = lo[2]; i2 <= hi[2]; ++i2)
                for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                    for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
                    {
                        const double uv = uc[(size_t)(bse[0] + i0 - g.glo[0]) +
                                             (size_t)g.n[0] * ((size_t)(bse[1] + i1 - g.glo[1]) +
                                                               (size_t)g.n[1] * (size_t)(bse[2] + i2 - g.glo[2]))];
                        v = v + (KT::IB ? (w[0][i0] * (w[1][i1] * w[2][i2])) * uv
                                        : ((w[0][i0] * w[1][i1]) * w[2][i2]) * uv);
                    }
        }
        else
        {
            for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
                {
                    const double uv =
                        uc[(size_t)(bse[0] + i0 - g.glo[0]) + (size_t)g.n[0] * (size_t)(bse[1] + i1 - g.glo[1])];
                    v = v + (w[0][i0] * w[1][i1]) * uv;
                }
        }
        V[(size_t)g.depth * s + c] = v;
    }
}

// One thread per (entry, depth component): writes W^NDIM contributions in the
// reference loop order (i2, i1, i0 fastest) at t0 = ((l-l0)*depth + c) * W^NDIM.
template <int NDIM, int KER>
__global__ void k_s0_contrib(S0Geom g, const int* __restrict__ indices, const double* __restrict__ Xshift, int l0,
                             int l1, const double* __restrict__ X, const double* __restrict__ V,
                             unsigned* __restrict__ key, int* __restrict__ tid, double* __restrict__ cval)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr int WN = NDIM == 3 ? W * W * W : W * W;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int nl = l1 - l0;
    if (q >= nl * g.depth) return;
    const int l = l0 + q / g.depth, c = q % g.depth;
    const int s = indices[l];
    double Xs[3];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) Xs[d] = X[NDIM * (size_t)s + d] + Xshift[NDIM * (size_t)l + d];
    int bse[3];
    double w[3][W];
    s0_stencil<NDIM, KER>(g, Xs, X + NDIM * (size_t)s, bse, w);
    const double vol = NDIM == 3 ? (g.dx[0] * g.dx[1]) * g.dx[2] : g.dx[0] * g.dx[1];
    const double Vv = V[(size_t)g.depth * s + c];
    const size_t ncomp = (size_t)g.n[0] * g.n[1] * (NDIM == 3 ? g.n[2] : 1);
    const size_t t0 = (size_t)q * WN;
    int e = 0;
#pragma unroll
    for (int i2 = 0; i2 < (NDIM == 3 ? W : 1); ++i2)
#pragma unroll
        for (int i1 = 0; i1 < W; ++i1)
#pragma unroll
            for (int i0 = 0; i0 < W; ++i0, ++e)
            {
                const int j0 = bse[0] + i0, j1 = bse[1] + i1, j2 = NDIM == 3 ? bse[2] + i2 : 0;
                bool in = j0 >= g.glo[0] && j0 <= g.ghi[0] && j1 >= g.glo[1] && j1 <= g.ghi[1];
                if (NDIM == 3) in = in && j2 >= g.glo[2] && j2 <= g.ghi[2];
                double cv = 0.0;
                unsigned k = S0_SENTINEL;
                if (in)
                {
                    if (NDIM == 3)
                        cv = KT::IB ? (w[0][i0] * (w[1][i1] * (w[2][i2] / vol))) * Vv
                                    : (((w[0][i0] * w[1][i1]) * w[2][i2]) * Vv) / vol;
                    else
                        cv = KT::IB ? (w[0][i0] * (w[1][i1] / vol)) * Vv : ((w[0][i0] * w[1][i1]) * Vv) / vol;
                    k = (unsigned)(c * ncomp + (size_t)(j0 - g.glo[0]) +
                                   (size_t)g.n[0] * ((size_t)(j1 - g.glo[1]) +
                                                     (NDIM == 3 ? (size_t)g.n[1] * (size_t)(j2 - g.glo[2]) : 0)));
                }
                key[t0 + e] = k;
                tid[t0 + e] = (int)(t0 + e);
                cval[t0 + e] = cv;
            }
}

// Each run of equal keys adds its contributions, in list order, to u.
__global__ void k_s0_sum(const unsigned* __restrict__ skey, const int* __restrict__ stid, size_t nt,
                         const double* __restrict__ cval, double* __restrict__ u)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const unsigned k = skey[i];
    if (k == S0_SENTINEL) return;
    if (i > 0 && skey[i - 1] == k) return;
    double v = u[k];
    for (size_t j = i; j < nt && skey[j] == k; ++j) v = v + cval[stid[j]];
    u[k] = v;
}

int kid(const char* k)
{
    if (!k) return -1;
    if (!std::strcmp(k, "IB_4")) return K_IB4;
    if (!std::strcmp(k, "IB_6")) return K_IB6;
    if (!std::strcmp(k, "BSPLINE_3")) return K_BS3;
    if (!std::strcmp(k, "PIECEWISE_LINEAR")) return K_PWL;
    if (!std::strcmp(k, "BSPLINE_4")) return K_BS4;
    if (!std::strcmp(k, "BSPLINE_5")) return K_BS5;
    if (!std::strcmp(k, "BSPLINE_6")) return K_BS6;
    return -1;
}

int width(int k)
{
    switch (k)
    {
    case K_IB4: return KernelTraits<K_IB4>::W;
    case K_IB6: return KernelTraits<K_IB6>::W;
    case K_BS3: return KernelTraits<K_BS3>::W;
    case K_BS4: return KernelTraits<K_BS4>::W;
    case K_BS5: return KernelTraits<K_BS5>::W;
    case K_BS6: return KernelTraits<K_BS6>::W;
    default: return KernelTraits<K_PWL>::W;
    }
}

struct DevBuf
{
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

#define S0_HIP(expr)                                                                              \
    do                                                                                            \
    {                                                                                             \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

int make_geom(int ndim, const double* dx, const double* x_lower, int depth, const int* ilower, const int* iupper,
              const int* nugc, S0Geom& g, size_t& ucount)
{
    if (ndim != 2 && ndim != 3) return set_error(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (depth < 1) return set_error(IBAMR_ERR_DEPTH, "depth must be >= 1");
    std::memset(&g, 0, sizeof(g));
    g.ndim = ndim;
    g.depth = depth;
    ucount = (size_t)depth;
    for (int d = 0; d < 3; ++d)
    {
        const bool on = d < ndim;
        g.dx[d] = on ? dx[d] : 1.0;
        g.x_lower[d] = on ? x_lower[d] : 0.0;
        g.ilo[d] = on ? ilower[d] : 0;
        g.glo[d] = on ? ilower[d] - nugc[d] : 0;
        g.ghi[d] = on ? iupper[d] + nugc[d] : 0;
        g.n[d] = g.ghi[d] - g.glo[d] + 1;
        if (on && (g.n[d] < 1 || nugc[d] < 0)) return set_error(IBAMR_ERR_INVALID, "empty ghost box");
        ucount *= (size_t)g.n[d];
    }
    if (ucount >= (size_t)S0_SENTINEL) return set_error(IBAMR_ERR_INVALID, "patch array too large");
    return IBAMR_OK;
}

template <typename F>
int by_kernel(int ndim, int k, F&& f)
{
    auto go = [&](auto nd) -> int {
        switch (k)
        {
        case K_IB4: return f(nd, std::integral_constant<int, K_IB4>{});
        case K_IB6: return f(nd, std::integral_constant<int, K_IB6>{});
        case K_BS3: return f(nd, std::integral_constant<int, K_BS3>{});
        case K_PWL: return f(nd, std::integral_constant<int, K_PWL>{});
        case K_BS4: return f(nd, std::integral_constant<int, K_BS4>{});
        case K_BS5: return f(nd, std::integral_constant<int, K_BS5>{});
        case K_BS6: return f(nd, std::integral_constant<int, K_BS6>{});
        default: return set_error(IBAMR_ERR_BAD_KERNEL, "unknown kernel function");
        }
    };
    return ndim == 3 ? go(std::integral_constant<int, 3>{}) : go(std::integral_constant<int, 2>{});
}

// The allocation-free interpolation's scratch: the last list entry of every
// point, one file-scope array per GPU (held until the process ends).  Host
// code uses the per-device address hipGetSymbolAddress returns, never the
// array's name.
__device__ __align__(256) int scratch[IBAMR_S0_NOALLOC_MAX_POINTS];

std::mutex g_scratch_mu;
std::map<int, int*> g_scratch_ptr;  // device -> address of its copy of scratch

// The current device's copy of scratch: looked up once per device, with that
// device current.
int scratch_ptr(int** out)
{
    int dev = 0;
    S0_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end())
    {
        int* ptr = nullptr;
        S0_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch)));
        if (!ptr) return set_error(IBAMR_ERR_DEVICE, "hipGetSymbolAddress returned a null scratch address");
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return IBAMR_OK;
}
} // namespace

extern "C" {

int ibamr_s0_interp_noalloc(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                            const int* ilower, const int* iupper, const int* nugc, const double* u,
                            const int* indices, const double* Xshift, int nindices, const double* X, int npoints,
                            double* V, void* stream)
{
    const int k = kid(kernel);
    if (k < 0) return set_error(IBAMR_ERR_BAD_KERNEL, std::string("unknown kernel function ") + (kernel ? kernel : ""));
    S0Geom g;
    size_t ucount = 0;
    int rc = make_geom(ndim, dx, x_lower, depth, ilower, iupper, nugc, g, ucount);
    if (rc != IBAMR_OK) return rc;
    if (nindices < 0 || npoints < 0) return set_error(IBAMR_ERR_INVALID, "negative count");
    if (npoints > IBAMR_S0_NOALLOC_MAX_POINTS)
        return set_error(IBAMR_ERR_CAPACITY, "npoints " + std::to_string(npoints) +
                                                 " exceeds the allocation-free scratch of " +
                                                 std::to_string(IBAMR_S0_NOALLOC_MAX_POINTS) + " points");
    if (nindices == 0 || npoints == 0) return IBAMR_OK;
    if (!u || !indices || !Xshift || !X || !V) return set_error(IBAMR_ERR_INVALID, "null device pointer");
    int* last = nullptr;
    rc = scratch_ptr(&last);
    if (rc != IBAMR_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256, G = (nindices + B - 1) / B;
    // the array keeps the previous call's contents: clear the part this call uses
    hipLaunchKernelGGL(k_s0_clear, dim3((npoints + B - 1) / B), dim3(B), 0, s, last, npoints);
    hipLaunchKernelGGL(k_s0_last, dim3(G), dim3(B), 0, s, indices, nindices, npoints, last);
    S0_HIP(hipGetLastError());
    rc = by_kernel(ndim, k, [&](auto nd, auto ker) -> int {
        hipLaunchKernelGGL((k_s0_interp<decltype(nd)::value, decltype(ker)::value>), dim3(G), dim3(B), 0, s, g, u,
                           indices, Xshift, nindices, npoints, X, V, last);
        return IBAMR_OK;
    });
    if (rc != IBAMR_OK) return rc;
    S0_HIP(hipGetLastError());
    return IBAMR_OK;
}

int ibamr_s0_interp(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                    const int* ilower, const int* iupper, const int* nugc, const double* u, const int* indices,
                    const double* Xshift, int nindices, const double* X, double* V)
{
    const int k = kid(kernel);
    if (k < 0) return set_error(IBAMR_ERR_BAD_KERNEL, std::string("unknown kernel function ") + (kernel ? kernel : ""));
    S0Geom g;
    size_t ucount = 0;
    int rc = make_geom(ndim, dx, x_lower, depth, ilower, iupper, nugc, g, ucount);
    if (rc != IBAMR_OK) return rc;
    if (nindices <= 0) return IBAMR_OK;
    int smax = 0;
    for (int l = 0; l < nindices; ++l)
    {
        if (indices[l] < 0) return set_error(IBAMR_ERR_INVALID, "negative index");
        smax = std::max(smax, indices[l]);
    }
    const size_t ns = (size_t)smax + 1;
    DevBuf du, di, dsh, dX, dV, dlast;
    S0_HIP(hipMalloc(&du.p, ucount * sizeof(double)));
    S0_HIP(hipMalloc(&di.p, (size_t)nindices * sizeof(int)));
    S0_HIP(hipMalloc(&dsh.p, (size_t)nindices * ndim * sizeof(double)));
    S0_HIP(hipMalloc(&dX.p, ns * ndim * sizeof(double)));
    S0_HIP(hipMalloc(&dV.p, ns * depth * sizeof(double)));
    S0_HIP(hipMalloc(&dlast.p, ns * sizeof(int)));
    S0_HIP(hipMemcpy(du.p, u, ucount * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(di.p, indices, (size_t)nindices * sizeof(int), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dsh.p, Xshift, (size_t)nindices * ndim * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dX.p, X, ns * ndim * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dV.p, V, ns * depth * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemset(dlast.p, 0xff, ns * sizeof(int)));
    const int B = 256, G = (nindices + B - 1) / B;
    hipLaunchKernelGGL(k_s0_last, dim3(G), dim3(B), 0, 0, static_cast<int*>(di.p), nindices, (int)ns,
                       static_cast<int*>(dlast.p));
    S0_HIP(hipGetLastError());
    rc = by_kernel(ndim, k, [&](auto nd, auto ker) -> int {
        hipLaunchKernelGGL((k_s0_interp<decltype(nd)::value, decltype(ker)::value>), dim3(G), dim3(B), 0, 0, g,
                           static_cast<double*>(du.p), static_cast<int*>(di.p), static_cast<double*>(dsh.p),
                           nindices, (int)ns, static_cast<double*>(dX.p), static_cast<double*>(dV.p),
                           static_cast<int*>(dlast.p));
        return IBAMR_OK;
    });
    if (rc != IBAMR_OK) return rc;
    S0_HIP(hipGetLastError());
    S0_HIP(hipMemcpy(V, dV.p, ns * depth * sizeof(double), hipMemcpyDeviceToHost));
    return IBAMR_OK;
}

int ibamr_s0_spread(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                    const int* indices, const double* Xshift, int nindices, const double* X, const double* V,
                    const int* ilower, const int* iupper, const int* nugc, double* u)
{
    const int k = kid(kernel);
    if (k < 0) return set_error(IBAMR_ERR_BAD_KERNEL, std::string("unknown kernel function ") + (kernel ? kernel : ""));
    S0Geom g;
    size_t ucount = 0;
    int rc = make_geom(ndim, dx, x_lower, depth, ilower, iupper, nugc, g, ucount);
    if (rc != IBAMR_OK) return rc;
    if (nindices <= 0) return IBAMR_OK;
    int smax = 0;
    for (int l = 0; l < nindices; ++l)
    {
        if (indices[l] < 0) return set_error(IBAMR_ERR_INVALID, "negative index");
        smax = std::max(smax, indices[l]);
    }
    const size_t ns = (size_t)smax + 1;
    const int W = width(k);
    const size_t WN = ndim == 3 ? (size_t)W * W * W : (size_t)W * W;
    const size_t per_entry = WN * depth;
    const int batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nindices, S0_MAX_TUPLES / per_entry));
    const size_t tmax = per_entry * batch;
    DevBuf du, di, dsh, dX, dV, dkey, dtid, dskey, dstid, dcv, dtmp;
    S0_HIP(hipMalloc(&du.p, ucount * sizeof(double)));
    S0_HIP(hipMalloc(&di.p, (size_t)nindices * sizeof(int)));
    S0_HIP(hipMalloc(&dsh.p, (size_t)nindices * ndim * sizeof(double)));
    S0_HIP(hipMalloc(&dX.p, ns * ndim * sizeof(double)));
    S0_HIP(hipMalloc(&dV.p, ns * depth * sizeof(double)));
    S0_HIP(hipMalloc(&dkey.p, tmax * sizeof(unsigned)));
    S0_HIP(hipMalloc(&dtid.p, tmax * sizeof(int)));
    S0_HIP(hipMalloc(&dskey.p, tmax * sizeof(unsigned)));
    S0_HIP(hipMalloc(&dstid.p, tmax * sizeof(int)));
    S0_HIP(hipMalloc(&dcv.p, tmax * sizeof(double)));
    size_t tb = 0;
    S0_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, static_cast<unsigned*>(dkey.p),
                                              static_cast<unsigned*>(dskey.p), static_cast<int*>(dtid.p),
                                              static_cast<int*>(dstid.p), (int)tmax, 0, 32));
    S0_HIP(hipMalloc(&dtmp.p, std::max<size_t>(tb, 16)));
    S0_HIP(hipMemcpy(du.p, u, ucount * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(di.p, indices, (size_t)nindices * sizeof(int), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dsh.p, Xshift, (size_t)nindices * ndim * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dX.p, X, ns * ndim * sizeof(double), hipMemcpyHostToDevice));
    S0_HIP(hipMemcpy(dV.p, V, ns * depth * sizeof(double), hipMemcpyHostToDevice));
    for (int l0 = 0; l0 < nindices; l0 += batch)
    {
        const int l1 = std::min(nindices, l0 + batch);
        const size_t nt = per_entry * (size_t)(l1 - l0);
        const int nq = (l1 - l0) * depth;
        const int B = 256;
        rc = by_kernel(ndim, k, [&](auto nd, auto ker) -> int {
            hipLaunchKernelGGL((k_s0_contrib<decltype(nd)::value, decltype(ker)::value>), dim3((nq + B - 1) / B),
                               dim3(B), 0, 0, g, static_cast<int*>(di.p), static_cast<double*>(dsh.p), l0, l1,
                               static_cast<double*>(dX.p), static_cast<double*>(dV.p), static_cast<unsigned*>(dkey.p),
                               static_cast<int*>(dtid.p), static_cast<double*>(dcv.p));
            return IBAMR_OK;
        });
        if (rc != IBAMR_OK) return rc;
        S0_HIP(hipGetLastError());
        size_t tb2 = tb;
        S0_HIP(hipcub::DeviceRadixSort::SortPairs(dtmp.p, tb2, static_cast<unsigned*>(dkey.p),
                                                  static_cast<unsigned*>(dskey.p), static_cast<int*>(dtid.p),
                                                  static_cast<int*>(dstid.p), (int)nt, 0, 32));
        hipLaunchKernelGGL(k_s0_sum, dim3((unsigned)((nt + B - 1) / B)), dim3(B), 0, 0,
                           static_cast<unsigned*>(dskey.p), static_cast<int*>(dstid.p), nt,
                           static_cast<double*>(dcv.p), static_cast<double*>(du.p));
        S0_HIP(hipGetLastError());
    }
    S0_HIP(hipMemcpy(u, du.p, ucount * sizeof(double), hipMemcpyDeviceToHost));
    return IBAMR_OK;
}

} // extern "C"

namespace
{
void s0_check(int rc, const char* name)
{
    if (rc != IBAMR_OK)
    {
        std::fprintf(stderr, "%s: %s\n", name, ibamr_last_error());
        std::abort();
    }
}
} // namespace

#define S0_INTERP3(NAME, KNAME)                                                                          \
    extern "C" void NAME(IBAMR_S0_INTERP3D_ARGS)                                                          \
    {                                                                                                     \
        (void)x_upper;                                                                                    \
        const int il[3] = {*ilower0, *ilower1, *ilower2}, iu[3] = {*iupper0, *iupper1, *iupper2};         \
        const int ng[3] = {*nugc0, *nugc1, *nugc2};                                                       \
        s0_check(ibamr_s0_interp(KNAME, 3, dx, x_lower, *depth, il, iu, ng, u, indices, Xshift, *nindices, X, V), \
                 #NAME);                                                                                  \
    }
#define S0_SPREAD3(NAME, KNAME)                                                                          \
    extern "C" void NAME(IBAMR_S0_SPREAD3D_ARGS)                                                          \
    {                                                                                                     \
        (void)x_upper;                                                                                    \
        const int il[3] = {*ilower0, *ilower1, *ilower2}, iu[3] = {*iupper0, *iupper1, *iupper2};         \
        const int ng[3] = {*nugc0, *nugc1, *nugc2};                                                       \
        s0_check(ibamr_s0_spread(KNAME, 3, dx, x_lower, *depth, indices, Xshift, *nindices, X, V, il, iu, ng, u), \
                 #NAME);                                                                                  \
    }
#define S0_INTERP2(NAME, KNAME)                                                                          \
    extern "C" void NAME(IBAMR_S0_INTERP2D_ARGS)                                                          \
    {                                                                                                     \
        (void)x_upper;                                                                                    \
        const int il[2] = {*ilower0, *ilower1}, iu[2] = {*iupper0, *iupper1}, ng[2] = {*nugc0, *nugc1};   \
        s0_check(ibamr_s0_interp(KNAME, 2, dx, x_lower, *depth, il, iu, ng, u, indices, Xshift, *nindices, X, V), \
                 #NAME);                                                                                  \
    }
#define S0_SPREAD2(NAME, KNAME)                                                                          \
    extern "C" void NAME(IBAMR_S0_SPREAD2D_ARGS)                                                          \
    {                                                                                                     \
        (void)x_upper;                                                                                    \
        const int il[2] = {*ilower0, *ilower1}, iu[2] = {*iupper0, *iupper1}, ng[2] = {*nugc0, *nugc1};   \
        s0_check(ibamr_s0_spread(KNAME, 2, dx, x_lower, *depth, indices, Xshift, *nindices, X, V, il, iu, ng, u), \
                 #NAME);                                                                                  \
    }

S0_INTERP3(lagrangian_ib_4_interp3d_, "IB_4")
S0_SPREAD3(lagrangian_ib_4_spread3d_, "IB_4")
S0_INTERP3(lagrangian_ib_6_interp3d_, "IB_6")
S0_SPREAD3(lagrangian_ib_6_spread3d_, "IB_6")
S0_INTERP3(lagrangian_bspline_3_interp3d_, "BSPLINE_3")
S0_SPREAD3(lagrangian_bspline_3_spread3d_, "BSPLINE_3")
S0_INTERP2(lagrangian_ib_4_interp2d_, "IB_4")
S0_SPREAD2(lagrangian_ib_4_spread2d_, "IB_4")
S0_INTERP2(lagrangian_ib_6_interp2d_, "IB_6")
S0_SPREAD2(lagrangian_ib_6_spread2d_, "IB_6")
S0_INTERP2(lagrangian_bspline_3_interp2d_, "BSPLINE_3")
S0_SPREAD2(lagrangian_bspline_3_spread2d_, "BSPLINE_3")
S0_INTERP3(lagrangian_piecewise_linear_interp3d_, "PIECEWISE_LINEAR")
S0_SPREAD3(lagrangian_piecewise_linear_spread3d_, "PIECEWISE_LINEAR")
S0_INTERP2(lagrangian_piecewise_linear_interp2d_, "PIECEWISE_LINEAR")
S0_SPREAD2(lagrangian_piecewise_linear_spread2d_, "PIECEWISE_LINEAR")
S0_INTERP3(lagrangian_bspline_4_interp3d_, "BSPLINE_4")
S0_SPREAD3(lagrangian_bspline_4_spread3d_, "BSPLINE_4")
S0_INTERP3(lagrangian_bspline_5_interp3d_, "BSPLINE_5")
S0_SPREAD3(lagrangian_bspline_5_spread3d_, "BSPLINE_5")
S0_INTERP3(lagrangian_bspline_6_interp3d_, "BSPLINE_6")
S0_SPREAD3(lagrangian_bspline_6_spread3d_, "BSPLINE_6")
S0_INTERP2(lagrangian_bspline_4_interp2d_, "BSPLINE_4")
S0_SPREAD2(lagrangian_bspline_4_spread2d_, "BSPLINE_4")
S0_INTERP2(lagrangian_bspline_5_interp2d_, "BSPLINE_5")
S0_SPREAD2(lagrangian_bspline_5_spread2d_, "BSPLINE_5")
S0_INTERP2(lagrangian_bspline_6_interp2d_, "BSPLINE_6")
S0_SPREAD2(lagrangian_bspline_6_spread2d_, "BSPLINE_6")
