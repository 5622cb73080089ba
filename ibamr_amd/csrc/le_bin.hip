// le_bin.hip -- point binning: IndexUtilities::getCellIndex per point (FP64, bit-exact,
// IndexUtilities-inl.h:52-72, level form :226-243), then a stable LSD radix sort of
// the points by spread bin (no device-wide atomics) and CSR bin offsets.
#include "le_plan_internal.h"

namespace ibamr
{
namespace
{
struct BinArgs
{
    const double* X;
    int n;
    double xlo[3], xhi[3], dx[3];
    int ref_lo[3], ref_hi[3];
    int sel_lo[3], sel_hi[3];
    int dom_lo[3], dom_hi[3];
    int nb[3];
    int nsb[3];
    int bin;
    int nmem;
    int ksub;   // bits of the plane-in-bin part of a key (0: bins only)
    int* keys;
    int4* cell;  // getCellIndex of every point, in the caller's order
};

// IndexUtilities::getCellIndex (IndexUtilities-inl.h:52-72) on the level domain
// (level overload :226-243), then the spread bin of the cell.
template <int NDIM>
__global__ void k_bin_keys(BinArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    int c[3] = {0, 0, 0};
    int raw[3] = {0, 0, 0};
    bool inside = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
    {
        const double v = cell_index_d(a.X[NDIM * p + d], a.xlo[d], a.xhi[d], a.dx[d], a.ref_lo[d], a.ref_hi[d]);
        // NaN or far outside: not selected
        inside = inside && (v >= (double)a.sel_lo[d]) && (v <= (double)a.sel_hi[d]);
        c[d] = inside ? (int)v : a.dom_lo[d];
        // the index itself where an int holds it (INT_MIN for NaN or beyond 2^30)
        raw[d] = (v >= -1073741824.0 && v <= 1073741824.0) ? (int)v : INT_MIN;
    }
    int key = a.nmem << a.ksub;  // sentinel: sorts after every bin
    if (inside)
    {
        // bin, then the cell's plane inside the bin along the last axis (z in
        // 3-D, y in 2-D), so that a bin's points are sorted by plane whatever the
        // input order and a 64-point chunk of a crowded bin spans few z planes
        int b[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < NDIM; ++d) b[d] = (c[d] - a.dom_lo[d]) / a.bin;
        const int l = c[NDIM - 1] - a.dom_lo[NDIM - 1] - b[NDIM - 1] * a.bin;
        key = (bin_mem(NDIM, b[0], b[1], b[2], a.nb, a.nsb) << a.ksub) | (a.ksub ? l : 0);
    }
    a.keys[p] = key;
    a.cell[p] = make_int4(raw[0], raw[1], raw[2], 0);
}

// Exclusive scan of count[0..m) into start[0..m] (start[m] = total), three passes.

__global__ __launch_bounds__(SCAN_B) void k_scan_blocks(const int* __restrict__ in, int m, int* __restrict__ out,
                                                        int* __restrict__ bsum)
{
    __shared__ int s_w[SCAN_B / 64];
    const int i = blockIdx.x * SCAN_B + threadIdx.x;
    const int v = i < m ? in[i] : 0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1)
    {
        const int t = __shfl_up(inc, off);
        if (lane >= off) inc += t;
    }
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int run = 0;
        for (int w = 0; w < SCAN_B / 64; ++w)
        {
            const int t = s_w[w];
            s_w[w] = run;
            run += t;
        }
        bsum[blockIdx.x] = run;
    }
    __syncthreads();
    if (i < m) out[i] = inc - v + s_w[wv];
}

__global__ __launch_bounds__(SCAN_B) void k_scan_sums(int* __restrict__ bsum, int nb)
{
    // one block: exclusive scan of the block sums, SCAN_B at a time
    __shared__ int s_w[SCAN_B / 64];
    __shared__ int s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int base = 0; base < nb; base += SCAN_B)
    {
        const int i = base + threadIdx.x;
        const int v = i < nb ? bsum[i] : 0;
        int inc = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1)
        {
            const int t = __shfl_up(inc, off);
            if (lane >= off) inc += t;
        }
        if (lane == 63) s_w[wv] = inc;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            int run = s_carry;
            for (int w = 0; w < SCAN_B / 64; ++w)
            {
                const int t = s_w[w];
                s_w[w] = run;
                run += t;
            }
            s_carry = run;
        }
        __syncthreads();
        if (i < nb) bsum[i] = inc - v + s_w[wv];
        __syncthreads();
    }
}

__global__ void k_scan_add(const int* __restrict__ part, int m, const int* __restrict__ bsum, const int* __restrict__ in,
                           int* __restrict__ start)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int v = part[i] + bsum[i / SCAN_B];
    start[i] = v;
    if (i == m - 1) start[m] = v + in[i];
}

// Stable LSD radix sort of the bin keys, carrying the point indices: after it
// the points are grouped by bin and, inside a bin, in index order, so every
// later kernel sees one deterministic order.  Device-wide atomics are avoided
// (on a multi-XCD GPU they are resolved past the per-XCD L2): per pass, a
// block histogram in LDS, one exclusive scan of the (digit, block) counts, and
// a scatter that ranks equal digits inside a block with wave ballots.

// hist[d * nblocks + block] = number of keys of this block's tile with digit d
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const int* __restrict__ keys, int n, int shift,
                                                        int nblocks, int* __restrict__ hist)
{
    __shared__ int s_cnt[RS_DIGITS];
    const int t = threadIdx.x;
    s_cnt[t] = 0;
    __syncthreads();
    const int i0 = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r)
    {
        const int i = i0 + r * RS_THREADS + t;
        if (i < n) atomicAdd(&s_cnt[(keys[i] >> shift) & (RS_DIGITS - 1)], 1);  // LDS: exact, order free
    }
    __syncthreads();
    hist[t * nblocks + blockIdx.x] = s_cnt[t];
}

// off = exclusive scan of hist (digit-major), so block b's keys of digit d go
// to off[d * nblocks + b] onward, in tile order.  vals_in == nullptr: the
// value of element i is i (first pass).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const int* __restrict__ keys_in,
                                                           const int* __restrict__ vals_in, int n, int shift,
                                                           int nblocks, const int* __restrict__ off,
                                                           int* __restrict__ keys_out, int* __restrict__ vals_out)
{
    constexpr int NW = RS_THREADS / 64;
    __shared__ int s_run[RS_DIGITS];       // next output slot per digit
    __shared__ int s_wc[NW][RS_DIGITS];    // per-wave count of a digit in the round, tagged with the round
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    s_run[t] = off[t * nblocks + blockIdx.x];
#pragma unroll
    for (int w = 0; w < NW; ++w) s_wc[w][t] = 0;
    __syncthreads();
    const unsigned long long lt = (1ull << lane) - 1ull;
    const int i0 = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r)
    {
        const int i = i0 + r * RS_THREADS + t;
        const bool valid = i < n;
        const int key = valid ? keys_in[i] : 0;
        const int val = valid ? (vals_in ? vals_in[i] : i) : 0;
        const int dig = (key >> shift) & (RS_DIGITS - 1);
        // lanes of this wave holding the same digit
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < RS_BITS; ++b)
        {
            const unsigned long long m = __ballot((dig >> b) & 1);
            peers &= ((dig >> b) & 1) ? m : ~m;
        }
        const int tag = (r + 1) << 16;
        if (valid && (peers & lt) == 0) s_wc[wv][dig] = tag | __popcll(peers);
        __syncthreads();
        if (valid)
        {
            int pos = s_run[dig] + __popcll(peers & lt);
            for (int w = 0; w < wv; ++w)
            {
                const int c = s_wc[w][dig];
                pos += (c >> 16) == r + 1 ? (c & 0xffff) : 0;
            }
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        int tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w)
        {
            const int c = s_wc[w][t];
            tot += (c >> 16) == r + 1 ? (c & 0xffff) : 0;
        }
        s_run[t] += tot;
        __syncthreads();
    }
}

// start[b] = first sorted position whose key is >= b, for b in [0, m] (keys in
// [0, m)): one binary search per bin, so the work per thread stays bounded
// however the points are spread over the bins.
__global__ void k_bin_bounds(const int* __restrict__ keys_sorted, int n, int m, int ksub, int* __restrict__ start)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b > m) return;
    int lo = 0, hi = n;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if ((keys_sorted[mid] >> ksub) < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[b] = lo;
}

// out[ndim * p + d] = coordinate d of point p's cell (caller's order)
__global__ void k_cells_out(const int4* __restrict__ cell, int n, int ndim, int* __restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int4 c = cell[p];
    out[(size_t)ndim * p] = c.x;
    out[(size_t)ndim * p + 1] = c.y;
    if (ndim == 3) out[(size_t)ndim * p + 2] = c.z;
}

// cells in bin order, and the inverse permutation (slot of every point)
__global__ void k_gather_cells(const int* __restrict__ perm, const int4* __restrict__ cell, int n,
                               int4* __restrict__ out, int* __restrict__ inv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    out[i] = cell[p];
    inv[p] = i;
}
} // namespace

// Stable LSD radix sort of n (key, value) pairs, keys in [0, 2^nbits).  vin ==
// nullptr: value i for element i.  The result lands in (kout, vout); (kt, vt)
// is the ping-pong pair.  Inputs are read in the first pass only.
int radix_sort_pairs(const ibamr_plan* P, const int* kin, const int* vin, int n, int nbits, int* kout, int* vout,
                     int* kt, int* vt, hipStream_t s)
{
    const int B = 256;
    const int passes = std::max(1, (nbits + RS_BITS - 1) / RS_BITS);
    const int nblk = (n + RS_TILE - 1) / RS_TILE;
    const int M = RS_DIGITS * nblk;
    const int nsb = (M + SCAN_B - 1) / SCAN_B;
    for (int p = 0; p < passes; ++p)
    {
        const bool last_parity = ((passes - 1 - p) & 1) == 0;
        int* ko = last_parity ? kout : kt;
        int* vo = last_parity ? vout : vt;
        const int shift = p * RS_BITS;
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(RS_THREADS), 0, s, kin, n, shift, nblk, P->rs_hist_d);
        hipLaunchKernelGGL(k_scan_blocks, dim3(nsb), dim3(SCAN_B), 0, s, P->rs_hist_d, M, P->scan_part_d,
                           P->scan_tmp_d);
        hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_B), 0, s, P->scan_tmp_d, nsb);
        hipLaunchKernelGGL(k_scan_add, dim3((M + B - 1) / B), dim3(B), 0, s, P->scan_part_d, M, P->scan_tmp_d,
                           P->rs_hist_d, P->rs_off_d);
        hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(RS_THREADS), 0, s, kin, vin, n, shift, nblk, P->rs_off_d,
                           ko, vo);
        HIP_OK(hipGetLastError());
        kin = ko;
        vin = vo;
    }
    return IBAMR_OK;
}

int bits_for(int maxkey)
{
    int nbits = 1;
    while (nbits < 31 && (maxkey >> nbits) != 0) ++nbits;
    return nbits;
}
} // namespace ibamr

extern "C" {

int ibamr_plan_bin_async(ibamr_plan* P, const double* X, int npoints, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (npoints < 0) return fail(IBAMR_ERR_INVALID, "npoints < 0");
    if (npoints > P->max_points) return fail(IBAMR_ERR_CAPACITY, "more points than the plan capacity");
    if (npoints > 0 && !X) return fail(IBAMR_ERR_INVALID, "null X");
    hipStream_t s = static_cast<hipStream_t>(stream);
    P->npoints = npoints;
    P->xs_valid = false;
    P->ws_valid = false;
    const int B = 256;
    const int m = P->nmem + 1;  // bins plus the out-of-domain sentinel
    if (npoints == 0)
    {
        hipLaunchKernelGGL(k_fill_int, dim3((m + 1 + B - 1) / B), dim3(B), 0, s, P->bin_start_d, m + 1, 0);
        HIP_OK(hipGetLastError());
        return IBAMR_OK;
    }
    BinArgs a{};
    a.X = X;
    a.n = npoints;
    for (int d = 0; d < 3; ++d)
    {
        a.xlo[d] = P->dom_xlo[d];
        a.xhi[d] = P->dom_xhi[d];
        a.ref_lo[d] = P->ref_lo[d];
        a.ref_hi[d] = P->ref_hi[d];
        a.sel_lo[d] = P->sel_lo[d];
        a.sel_hi[d] = P->sel_hi[d];
        a.dx[d] = P->dx[d];
        a.dom_lo[d] = P->dom_lo[d];
        a.dom_hi[d] = P->dom_hi[d];
        a.nb[d] = P->nb[d];
        a.nsb[d] = P->nsb[d];
    }
    a.bin = P->bin;
    a.nmem = P->nmem;
    a.ksub = P->ksub;
    a.keys = P->keys_d;
    a.cell = P->cell_d;
    const int G = (npoints + B - 1) / B;
    if (P->ndim == 3)
        hipLaunchKernelGGL(k_bin_keys<3>, dim3(G), dim3(B), 0, s, a);
    else
        hipLaunchKernelGGL(k_bin_keys<2>, dim3(G), dim3(B), 0, s, a);
    HIP_OK(hipGetLastError());
    // stable radix sort of keys [0, nmem << ksub] over the bits they use, into (keys_sorted_d, perm_d)
    {
        const int rc = radix_sort_pairs(P, P->keys_d, nullptr, npoints, bits_for(P->nmem << P->ksub), P->keys_sorted_d,
                                        P->perm_d, P->keys2_d, P->idx_d, s);
        if (rc != IBAMR_OK) return rc;
    }
    hipLaunchKernelGGL(k_bin_bounds, dim3((m + 1 + B - 1) / B), dim3(B), 0, s, P->keys_sorted_d, npoints, m,
                       P->ksub, P->bin_start_d);
    HIP_OK(hipGetLastError());
    // idx_d is free once the sort is done: it now holds the inverse permutation
    hipLaunchKernelGGL(k_gather_cells, dim3(G), dim3(B), 0, s, P->perm_d, P->cell_d, npoints, P->cell_sorted_d,
                       P->idx_d);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

int ibamr_plan_check(ibamr_plan* P, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    int first_sentinel = 0;
    HIP_OK(hipMemcpyAsync(&first_sentinel, P->bin_start_d + P->nmem, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (first_sentinel != P->npoints)
        return fail(IBAMR_ERR_OUT_OF_DOMAIN, std::to_string(P->npoints - first_sentinel) +
                                                 " point(s) outside the level domain (or NaN positions)");
    return IBAMR_OK;
}

int ibamr_plan_get_cell_indices(const ibamr_plan* P, int* cells, void* stream)
{
    if (!P || !cells) return fail(IBAMR_ERR_INVALID, "null argument");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    if (P->npoints == 0) return IBAMR_OK;
    const int B = 256;
    hipLaunchKernelGGL(k_cells_out, dim3((P->npoints + B - 1) / B), dim3(B), 0, static_cast<hipStream_t>(stream),
                       P->cell_d, P->npoints, P->ndim, cells);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

int ibamr_plan_bin(ibamr_plan* P, const double* X, int npoints, void* stream)
{
    const int rc = ibamr_plan_bin_async(P, X, npoints, stream);
    if (rc != IBAMR_OK) return rc;
    return ibamr_plan_check(P, stream);
}

} // extern "C"
