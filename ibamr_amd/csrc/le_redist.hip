// le_redist.hip -- Lagrangian redistribution at a regrid
// (LDataManager::beginDataRedistribution / endDataRedistribution,
// LDataManager.cpp:1349-1960) and the row permutation of LData arrays.
#include "le_plan_internal.h"

namespace ibamr
{
namespace
{
struct RedistArgs
{
    double* X;
    const int* lag;
    int n;
    double xlo[3], xhi[3], dx[3];
    int dom_lo[3], dom_hi[3];
    int pshift[3];
    int nb[3], nsb[3];
    int bin, nmem, npatches;
    int error_if_leave;
    double tol;
    const PatchDev* patches;
    const int* bin_patch_start;
    const int* bin_patches;
    const int* pcell_off;
    int* key_lag;
    int* key_cell;
    int* offset;   // [n][ndim] periodic offsets, or null
    int* blockmax; // [gridDim.x] largest Lagrangian index of each block
    int* status;   // 0: escaped, 1: in no patch, 2: largest Lagrangian index, 3: not finite or beyond the
                   // wrap limit, 4: negative Lagrangian index
};

constexpr int REDIST_MAX_WRAPS = 64;  // periodic wraps per coordinate before a point counts as lost

// LDataManager::beginDataRedistribution (LDataManager.cpp:1393-1431): move the
// position into the domain (periodic wrap, or clamp / error), record the
// periodic offset round((X_old - X_new) / dx); then the owning patch of its
// cell (IndexUtilities::getCellIndex) and the keys of the new order.
template <int NDIM>
__global__ void k_redist_prepare(RedistArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    int lag = 0;
    if (p < a.n)
    {
        int c[3] = {0, 0, 0};
        bool ok = true, lost = false, escaped = false;
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
        {
            const double Xr = a.X[NDIM * (size_t)p + d];
            double X = Xr;
            if (!isfinite(Xr))
            {
                lost = true;
            }
            else if (a.pshift[d] > 0)
            {
                const double L = a.xhi[d] - a.xlo[d];
                int w = 0;
                while (X < a.xlo[d] && w < REDIST_MAX_WRAPS)
                {
                    X += L;
                    ++w;
                }
                while (X >= a.xhi[d] && w < REDIST_MAX_WRAPS)
                {
                    X -= L;
                    ++w;
                }
                if (!(X >= a.xlo[d] && X < a.xhi[d])) lost = true;  // too far out
            }
            else if (a.error_if_leave)
            {
                if (X < a.xlo[d] || X > a.xhi[d]) escaped = true;
            }
            else
            {
                X = fmax(X, a.xlo[d]);
                X = fmin(X, a.xhi[d] - (a.xhi[d] - a.xlo[d]) * a.tol);
            }
            a.X[NDIM * (size_t)p + d] = X;
            if (a.offset) a.offset[NDIM * (size_t)p + d] = lost ? 0 : (int)round((Xr - X) / a.dx[d]);
            // IndexUtilities::getCellIndex on the level domain
            const double v = cell_index_d(X, a.xlo[d], a.xhi[d], a.dx[d], a.dom_lo[d], a.dom_hi[d]);
            const bool inside = v >= (double)a.dom_lo[d] && v <= (double)a.dom_hi[d];
            ok = ok && !lost && !escaped && inside;
            c[d] = ok ? (int)v : a.dom_lo[d];
        }
        if (lost) atomicAdd(&a.status[3], 1);
        if (escaped && !lost) atomicAdd(&a.status[0], 1);
        int q = -1;
        if (ok)
        {
            int b[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < NDIM; ++d) b[d] = (c[d] - a.dom_lo[d]) / a.bin;
            const int key = bin_mem(NDIM, b[0], b[1], b[2], a.nb, a.nsb);
            for (int j = a.bin_patch_start[key]; j < a.bin_patch_start[key + 1] && q < 0; ++j)
            {
                const PatchDev& Q = a.patches[a.bin_patches[j]];
                bool in = true;
#pragma unroll
                for (int d = 0; d < NDIM; ++d) in = in && c[d] >= Q.ilo[d] && c[d] <= Q.ihi[d];
                if (in) q = a.bin_patches[j];
            }
        }
        // a cell outside every patch (or outside the domain, e.g. X == x_upper)
        if (q < 0 && !lost && !escaped) atomicAdd(&a.status[1], 1);
        int cell = 0;
        if (q >= 0)
        {
            const PatchDev& Q = a.patches[q];
            int lin = 0;
#pragma unroll
            for (int d = NDIM - 1; d >= 0; --d) lin = lin * (Q.ihi[d] - Q.ilo[d] + 1) + (c[d] - Q.ilo[d]);
            cell = a.pcell_off[q] + lin;
        }
        lag = a.lag ? a.lag[p] : p;
        if (lag < 0)
        {
            atomicAdd(&a.status[4], 1);  // negative Lagrangian index
            lag = 0;
        }
        a.key_cell[p] = cell;
        a.key_lag[p] = lag;
    }
    // largest Lagrangian index of the block (no device-wide atomic: one value per block)
    __shared__ int s_max[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lag = max(lag, __shfl_xor(lag, off));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = lag;
    __syncthreads();
    if (threadIdx.x == 0) a.blockmax[blockIdx.x] = max(max(s_max[0], s_max[1]), max(s_max[2], s_max[3]));
}

// status[2] = max of the per-block maxima (one block)
__global__ __launch_bounds__(256) void k_max_reduce(const int* __restrict__ v, int n, int* __restrict__ out)
{
    __shared__ int s_w[4];
    int m = 0;
    for (int i = threadIdx.x; i < n; i += 256) m = max(m, v[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = max(max(s_w[0], s_w[1]), max(s_w[2], s_w[3]));
}

__global__ void k_gather_int(const int* __restrict__ src, const int* __restrict__ idx, int n, int* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// out[j] = first position of the sorted keys that is >= vals[j]
__global__ void k_lower_bounds(const int* __restrict__ keys, int n, const int* __restrict__ vals, int nv,
                               int* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nv) return;
    const int v = vals[j];
    int lo = 0, hi = n;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    out[j] = lo;
}

// VecScatter from the old to the new order: row k of dst = row order[k] of src
template <typename E>
__global__ void k_permute_rows(E* __restrict__ dst, const E* __restrict__ src, const int* __restrict__ order, int n,
                               int depth)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * depth) return;
    const int k = (int)(i / depth), j = (int)(i - (long)k * depth);
    dst[i] = src[(size_t)order[k] * depth + j];
}
} // namespace
} // namespace ibamr

extern "C" {

int ibamr_plan_redistribute(ibamr_plan* P, double* X, const int* lag_index, int npoints, int error_if_points_leave,
                            int* new_order, int* patch_offsets, int* periodic_offset, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (npoints < 0) return fail(IBAMR_ERR_INVALID, "npoints < 0");
    if (npoints > P->max_points) return fail(IBAMR_ERR_CAPACITY, "more points than the plan capacity");
    if ((npoints > 0 && (!X || !new_order)) || !patch_offsets) return fail(IBAMR_ERR_INVALID, "null argument");
    if (!P->pcell_off_d)
        return fail(IBAMR_ERR_INVALID, "redistribution needs a level plan whose patches hold fewer than 2^31 cells");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    P->npoints = -1;  // the binning referred to the old order
    P->xs_valid = false;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, P->status_d, 8, 0);
    RedistArgs a{};
    a.X = X;
    a.lag = lag_index;
    a.n = npoints;
    for (int d = 0; d < 3; ++d)
    {
        a.xlo[d] = P->dom_xlo[d];
        a.xhi[d] = P->dom_xhi[d];
        a.dx[d] = P->dx[d];
        a.dom_lo[d] = P->dom_lo[d];
        a.dom_hi[d] = P->dom_hi[d];
        a.pshift[d] = P->pshift[d];
        a.nb[d] = P->nb[d];
        a.nsb[d] = P->nsb[d];
    }
    a.bin = P->bin;
    a.nmem = P->nmem;
    a.npatches = P->npatches;
    a.error_if_leave = error_if_points_leave;
    a.tol = std::sqrt(std::numeric_limits<double>::epsilon());  // LDataManager.cpp:150
    a.patches = P->patches_d;
    a.bin_patch_start = P->bin_patch_start_d;
    a.bin_patches = P->bin_patches_d;
    a.pcell_off = P->pcell_off_d;
    a.key_lag = P->keys_d;
    a.key_cell = reinterpret_cast<int*>(P->cell_d);  // cell_d holds 4 ints per point; the binning is void now
    a.offset = periodic_offset;
    a.status = P->status_d;
    a.blockmax = P->rs_hist_d;  // free until the sorts below; holds >= max_points / 16 ints
    if (npoints > 0)
    {
        static_assert(RS_TILE / RS_DIGITS <= 256, "block maxima must fit the histogram buffer");
        const int G = (npoints + B - 1) / B;
        if (P->ndim == 3)
            hipLaunchKernelGGL(k_redist_prepare<3>, dim3(G), dim3(B), 0, s, a);
        else
            hipLaunchKernelGGL(k_redist_prepare<2>, dim3(G), dim3(B), 0, s, a);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, s, P->rs_hist_d, G, P->status_d + 2);
        HIP_OK(hipGetLastError());
    }
    int st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIP_OK(hipMemcpyAsync(st, P->status_d, sizeof(st), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (st[3] > 0)
        return fail(IBAMR_ERR_OUT_OF_DOMAIN, std::to_string(st[3]) + " point(s) with non-finite positions or more than " +
                                                 std::to_string(REDIST_MAX_WRAPS) + " periodic lengths outside the domain");
    if (st[4] > 0) return fail(IBAMR_ERR_INVALID, "negative Lagrangian index");
    if (st[0] > 0)
        return fail(IBAMR_ERR_OUT_OF_DOMAIN, "IB point has escaped from the computational domain! (" +
                                                 std::to_string(st[0]) + " point(s))");
    if (st[1] > 0)
        return fail(IBAMR_ERR_OUT_OF_DOMAIN, std::to_string(st[1]) + " point(s) lie in no patch of the level");
    if (npoints > 0)
    {
        int* key_cell = reinterpret_cast<int*>(P->cell_d);
        const int G = (npoints + B - 1) / B;
        const int* order1 = nullptr;
        if (lag_index)
        {
            // by Lagrangian index first, then (stable) by cell: patches in level order, cells in box order,
            // nodes of a cell by Lagrangian index (LDataManager.cpp:1498-1503, 2897-2908)
            int rc = radix_sort_pairs(P, P->keys_d, nullptr, npoints, bits_for(st[2]), P->keys_sorted_d, P->perm_d,
                                      P->keys2_d, P->idx_d, s);
            if (rc != IBAMR_OK) return rc;
            hipLaunchKernelGGL(k_gather_int, dim3(G), dim3(B), 0, s, key_cell, P->perm_d, npoints, P->keys_d);
            order1 = P->perm_d;
        }
        else
        {
            // Lagrangian index = position: already in index order
            hipLaunchKernelGGL(k_copy_int, dim3(G), dim3(B), 0, s, key_cell, npoints, P->keys_d);
        }
        int total = 0;
        HIP_OK(hipMemcpyAsync(&total, P->pcell_off_d + P->npatches, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        const int rc = radix_sort_pairs(P, P->keys_d, order1, npoints, bits_for(std::max(0, total - 1)),
                                        P->keys_sorted_d, new_order, P->keys2_d, P->idx_d, s);
        if (rc != IBAMR_OK) return rc;
    }
    hipLaunchKernelGGL(k_lower_bounds, dim3((P->npatches + 1 + B - 1) / B), dim3(B), 0, s, P->keys_sorted_d, npoints,
                       P->pcell_off_d, P->npatches + 1, patch_offsets);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

int ibamr_permute_rows(void* dst, const void* src, const int* order, int n, int depth, int elem_bytes, void* stream)
{
    if (n < 0 || depth < 1) return fail(IBAMR_ERR_INVALID, "bad size");
    if (n == 0) return IBAMR_OK;
    if (!dst || !src || !order) return fail(IBAMR_ERR_INVALID, "null argument");
    if (dst == src) return fail(IBAMR_ERR_INVALID, "dst and src must differ");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long tot = (long)n * depth;
    const int B = 256;
    const dim3 G((unsigned)((tot + B - 1) / B));
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_permute_rows<long long>, G, dim3(B), 0, s, static_cast<long long*>(dst),
                           static_cast<const long long*>(src), order, n, depth);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(k_permute_rows<int>, G, dim3(B), 0, s, static_cast<int*>(dst), static_cast<const int*>(src),
                           order, n, depth);
    else
        return fail(IBAMR_ERR_INVALID, "elem_bytes must be 4 or 8");
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

} // extern "C"
