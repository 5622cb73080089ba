// le_plan.hip -- level plans: point binning, deterministic gather spread,
// tensor-product interpolation and periodic ghost fill on gfx950.
//
// Reference path (IBAMR): LDataManager::spread/interp (LDataManager.cpp:552-667,
// 699-813) -> LEInteractor::spread/interpolate (side wrappers LEInteractor.cpp:
// 2402-2489, 3627-3714) -> lagrangian_<kernel>_{spread,interp}{2,3}d
// (lagrangian_interaction{2,3}d.f.m4).
//
// Design (MI355X-first, not a translation of the per-point Fortran loops):
//  * binning: IndexUtilities::getCellIndex per point (FP64, bit-exact), then a
//    stable LSD radix sort of the points by spread bin (a cube of cells; no
//    device-wide atomics); CSR offsets by a binary search per bin.
//  * spread: owner-computes gather.  A tile of grid values of one array (side
//    axis or depth component) of one patch belongs to one wavefront (2-D here:
//    a 16x16 tile in LDS; 3-D in le_kernels3d.h: 8x8 columns of z values in
//    registers).  It visits the bins whose points can reach the tile in a fixed
//    order (bin z, image z, bin y, image y, bin x, image x) and the points of
//    each bin in index order, so every grid value sums its contributions in a
//    fixed order: no atomics, bitwise reproducible.  Periodic images
//    (LIndexSetData.cpp:80-117) are generated on the fly from the bin offsets.
//  * interp: 2-D one thread per point in bin order; 3-D one workgroup per
//    superbin with the grid block staged in LDS (le_kernels3d.h).
//  * around the path: redistribution at a regrid (radix sorts by patch, cell and
//    Lagrangian index), same-level and periodic ghost fills, the physical
//    boundary (Robin) operators, the forward-Euler update.
#include "ibamr_amd.h"
#include "le_weights.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace ibamr
{
static thread_local std::string g_err;

static int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// shared with the other translation units of the library
int set_error(int code, const std::string& msg) { return fail(code, msg); }

#define HIP_OK(expr)                                                                          \
    do                                                                                        \
    {                                                                                         \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(IBAMR_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static constexpr int WAVE = 64;
static constexpr int NSLOT = 64;
static constexpr int MAX_SLOTS = 8;

struct PatchDev
{
    int ilo[3];
    int ihi[3];
    double xlo[3];
    int slo[3];   // interpolation ownership box: the patch box (level plans) or
    int shi[3];   // the caller's box (single-patch plans)
};

struct TileDesc
{
    int patch;
    int comp;   // side: axis; cell: depth component
    int t[3];   // tile coordinates inside the array's ghost box
};

// The reference's kernel registry (LEInteractor.cpp:2035-2108): every known
// name and its stencil size.  IB_4, IB_6, BSPLINE_3..6 and PIECEWISE_LINEAR
// have GPU kernels; the other names are known (isKnownKernel, getStencilSize)
// but plans reject them.
struct KernelName
{
    const char* name;
    int stencil;
    int id;   // K_* of the GPU kernel, or -2 (registry only)
};
static const KernelName k_registry[] = {
    {"BSPLINE_3", 4, K_BS3},          {"BSPLINE_4", 4, K_BS4},
    {"BSPLINE_5", 6, K_BS5},          {"BSPLINE_6", 6, K_BS6},
    {"COMPOSITE_BSPLINE_23", 4, -2},  {"COMPOSITE_BSPLINE_32", 4, -2},
    {"COMPOSITE_BSPLINE_34", 4, -2},  {"COMPOSITE_BSPLINE_43", 4, -2},
    {"COMPOSITE_BSPLINE_45", 5, -2},  {"COMPOSITE_BSPLINE_54", 5, -2},
    {"COMPOSITE_BSPLINE_56", 6, -2},  {"COMPOSITE_BSPLINE_65", 6, -2},
    {"DISCONTINUOUS_LINEAR", 2, -2},  {"PIECEWISE_CONSTANT", 1, -2},
    {"PIECEWISE_LINEAR", 2, K_PWL},   {"PIECEWISE_CUBIC", 4, -2},
    {"IB_3", 4, -2},                  {"IB_4", 4, K_IB4},
    {"IB_4_W8", 8, -2},               {"IB_5", 6, -2},
    {"IB_6", 6, K_IB6},               {"USER_DEFINED", 0, -2},
};

static const KernelName* find_kernel(const char* name)
{
    if (!name) return nullptr;
    for (const KernelName& k : k_registry)
        if (!std::strcmp(name, k.name)) return &k;
    return nullptr;
}

// K_* id of a kernel with a GPU implementation; -2 known without one; -1 unknown
static int kernel_id(const char* name)
{
    const KernelName* k = find_kernel(name);
    return k ? k->id : -1;
}

static int stencil_size(int k)
{
    switch (k)
    {
    case K_PWL: return 2;
    case K_IB4: return 4;
    case K_IB6: return 6;
    case K_BS3: return 4;  // LEInteractor.cpp:2057-2058
    case K_BS4: return 4;
    case K_BS5: return 6;  // LEInteractor.cpp:2061-2062
    case K_BS6: return 6;
    default: return -1;
    }
}

static int min_ghost(int k) { return stencil_size(k) < 0 ? -1 : stencil_size(k) / 2 + 1; }

__host__ __device__ inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Memory slot of a bin.  3-D: bins are grouped into superbins of 2x2x2 bins
// whose points are contiguous (the interpolation stages one superbin block);
// 2-D: row-major.  The spread visits bins in logical (z, y, x) order
// whatever their memory slots.
__host__ __device__ inline int bin_mem(int ndim, int bx, int by, int bz, const int* nb, const int* nsb)
{
    if (ndim == 2) return by * nb[0] + bx;
    const int sb = ((bz >> 1) * nsb[1] + (by >> 1)) * nsb[0] + (bx >> 1);
    return sb * 8 + (((bz & 1) << 2) | ((by & 1) << 1) | (bx & 1));
}

} // namespace ibamr

using namespace ibamr;

struct ibamr_plan
{
    int ndim = 3;
    int kernel = K_IB4;
    int centering = IBAMR_SIDE;
    int precision = IBAMR_FP64;
    int depth = 1;
    int narr = 3;           // arrays per patch
    int fdepth = 3;         // F/U components per point
    int data_gcw[3] = {0, 0, 0};
    int idx_gcw = 0;
    int reach = 0;
    int tile = 8;
    int tsz[3] = {8, 8, 8};   // spread tile extent per axis
    int bin = 8;
    int drift = 0;
    double dx[3] = {1, 1, 1};
    int dom_lo[3] = {0, 0, 0}, dom_hi[3] = {0, 0, 0};   // bin grid box
    double dom_xlo[3] = {0, 0, 0}, dom_xhi[3] = {0, 0, 0};
    int ref_lo[3] = {0, 0, 0}, ref_hi[3] = {0, 0, 0};   // box of the getCellIndex formula (x range dom_x*)
    int sel_lo[3] = {0, 0, 0}, sel_hi[3] = {0, 0, 0};   // points kept: cell inside this box
    int pshift[3] = {0, 0, 0};
    int nb[3] = {1, 1, 1};
    int nsb[3] = {1, 1, 1};
    int nbins = 1;
    int nmem = 1;           // bin memory slots (>= nbins)
    int ksub = 0;           // bits of the plane-in-bin part of the binning keys
    int npatches = 0;
    std::vector<PatchDev> patches_h;
    int max_points = 0;
    int npoints = -1;       // points binned (-1: not binned)

    // device buffers
    PatchDev* patches_d = nullptr;
    TileDesc* tiles_d = nullptr;
    int ntiles = 0;
    int* bin_patch_start_d = nullptr;  // [nmem+2]
    int* bin_patches_d = nullptr;
    int4* fill_segs_d = nullptr;       // level ghost fill: (patch, array, first element, count) segments
    int nfill_segs = 0;
    void** arrays_d[MAX_SLOTS] = {};
    std::vector<void*> arrays_h[MAX_SLOTS];   // host copy of each slot's pointer table
    int* keys_d = nullptr;
    int* keys_sorted_d = nullptr;
    int* idx_d = nullptr;              // point indices in bin order before the per-bin sort; afterwards
                                       // the inverse permutation (slot of every point)
    int* keys2_d = nullptr;            // radix sort ping-pong buffer
    int* pcell_off_d = nullptr;        // [npatches+1] first cell number of each patch (box order), or null
                                       // when the level has 2^31 cells or more
    int* status_d = nullptr;           // [8] redistribution counters
    int4* bdry_patch_d = nullptr;      // physical boundary: (patch, first box, number of boxes, 0) per patch
    int* bdry_box_d = nullptr;         // [nbdry][8]: codim, location, lower[3], upper[3]
    double* bdry_abg_d = nullptr;      // [nbdry][narr][3] Robin coefficients
    int nbdry_patches = 0;
    int* rs_hist_d = nullptr;          // [RS_DIGITS * blocks] (digit, block) counts
    int* rs_off_d = nullptr;           // [RS_DIGITS * blocks + 1] their exclusive scan
    int* scan_tmp_d = nullptr;         // block sums of the scan
    int* scan_part_d = nullptr;        // block-local exclusive prefixes of the counts
    int* perm_d = nullptr;
    int4* cell_d = nullptr;
    int4* cell_sorted_d = nullptr;
    double* Xs_d = nullptr;            // positions in bin order
    bool xs_valid = false;             // Xs holds the positions of the last spread/interp (same binning)
    bool ws_valid = false;             // wt_d/st_d hold the stencils of Xs (FP32C 3-D)
    void* Fs_d = nullptr;              // forces in bin order
    // FP32C 3-D: per point in bin order, axis d and variant v, the first stencil
    // index (st_d[(d*nvar+v)*n + k]) and the W weights (wt_d[((d*nvar+v)*n + k)*W + j])
    float* wt_d = nullptr;
    int* st_d = nullptr;
    int nvar = 1;                      // variants per axis: 2 for side data (x_lower, x_lower - dx/2)
    int* bin_start_d = nullptr;        // [nmem+2]

    // optional timing of the main spread / interpolation kernels: event pairs
    // recorded on the launch stream around each launch (0: spread, 1: interp)
    bool timing = false;
    std::vector<hipEvent_t> tev[2];
    int tcount[2] = {0, 0};
};

// ----------------------------------------------------------------------------- binning

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
    int4* cell;
};

// IndexUtilities::getCellIndex (IndexUtilities-inl.h:52-72) on the level domain
// (level overload :226-243), then the spread bin of the cell.
template <int NDIM>
__global__ void k_bin_keys(BinArgs a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    int c[3] = {0, 0, 0};
    bool inside = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
    {
        const double X = a.X[NDIM * p + d];
        const double dXl = X - a.xlo[d], dXu = X - a.xhi[d];
        double v;
        if (fabs(dXl) <= fabs(dXu))
            v = (double)a.ref_lo[d] + floor(dXl / a.dx[d]);
        else
            v = ((double)a.ref_hi[d] + floor(dXu / a.dx[d])) + 1.0;
        // NaN or far outside: not selected
        inside = inside && (v >= (double)a.sel_lo[d]) && (v <= (double)a.sel_hi[d]);
        c[d] = inside ? (int)v : a.dom_lo[d];
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
    a.cell[p] = make_int4(c[0], c[1], c[2], 0);
}

__global__ void k_fill_int(int* __restrict__ a, int n, int v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_copy_int(const int* __restrict__ a, int n, int* __restrict__ b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Exclusive scan of count[0..m) into start[0..m] (start[m] = total), three passes.
constexpr int SCAN_B = 1024;

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
constexpr int RS_BITS = 8;
constexpr int RS_DIGITS = 1 << RS_BITS;
constexpr int RS_THREADS = 256;            // one digit per thread in the bookkeeping
constexpr int RS_ROUNDS = 16;              // rounds of RS_THREADS elements per block
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
static_assert(RS_THREADS == RS_DIGITS, "one thread per digit");

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

// Point data in bin order, so every later kernel reads slot k coalesced.
// Driven by the points: read X and F in the caller's order (coalesced) and
// write each point's record to its slot (a scattered write of a few bytes
// costs less than a scattered read of whole cache lines).
template <typename T>
__global__ void k_scatter_points(const int* __restrict__ inv, int n, int ndim, const double* __restrict__ X,
                                 double* __restrict__ Xs, int fdepth, const T* __restrict__ F, T* __restrict__ Fs)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const size_t k = (size_t)inv[p];
    for (int d = 0; d < ndim; ++d) Xs[(size_t)ndim * k + d] = X[(size_t)ndim * p + d];
    if (F)
        for (int c = 0; c < fdepth; ++c) Fs[(size_t)fdepth * k + c] = F[(size_t)fdepth * p + c];
}

// FP32C 3-D spread set-up: for each point (bin order), axis d and variant v
// (v = 0: the level's x_lower; v = 1: x_lower - dx/2, the normal axis of side
// data, LEInteractor.cpp:3689) the first stencil index in the level's index
// space and the W weights, in the level's frame.  The spread's tiles then stage
// a candidate with loads instead of recomputing its weights for every tile it
// reaches; a patch or periodic image shifts the start by whole cells.  With
// perm, the kernel also performs the bin-order gather of X and F (side data:
// records of 3 positions and 3 force components), so Xs is not read back.
template <int KER>
__global__ void k_point_weights(int n, double* __restrict__ Xs, int nvar, double3 xlo, double3 dx, int3 lo,
                                float* __restrict__ wt, int* __restrict__ st, const int* __restrict__ perm,
                                const double* __restrict__ Xin, const float* __restrict__ Fin,
                                float* __restrict__ Fs)
{
    constexpr int W = KernelTraits<KER>::W;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double xl[3] = {xlo.x, xlo.y, xlo.z}, h[3] = {dx.x, dx.y, dx.z};
    const int il[3] = {lo.x, lo.y, lo.z};
    double x[3];
    if (perm)
    {
        const size_t p = (size_t)perm[k];
        float f[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            x[d] = Xin[3 * p + d];
            f[d] = Fin[3 * p + d];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            Xs[3 * (size_t)k + d] = x[d];
            Fs[3 * (size_t)k + d] = f[d];
        }
    }
    else
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = Xs[3 * (size_t)k + d];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const double X = x[d];
        for (int v = 0; v < nvar; ++v)
        {
            float w[W];
            const int s0 = weights_1d<KER>(X, v ? xl[d] - 0.5 * h[d] : xl[d], h[d], w);
            const size_t row = (size_t)(d * nvar + v) * n + k;
            st[row] = il[d] + s0;
#pragma unroll
            for (int j = 0; j < W; ++j) wt[row * W + j] = w[j];
        }
    }
}

// FP32C 3-D: stencil starts and weights of the bin-ordered positions, used by
// the spread and the interpolation of the same positions
// (with X and F: the gather into bin order of side data's X and F as well)
template <int KER>
int launch_point_weights(ibamr_plan* P, hipStream_t s, const double* X = nullptr, const float* F = nullptr)
{
    if (!P->wt_d || !P->st_d) return fail(IBAMR_ERR_INVALID, "internal: no weight buffers");
    const int B = 256;
    hipLaunchKernelGGL((k_point_weights<KER>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->npoints, P->Xs_d,
                       P->nvar, make_double3(P->dom_xlo[0], P->dom_xlo[1], P->dom_xlo[2]),
                       make_double3(P->dx[0], P->dx[1], P->dx[2]),
                       make_int3(P->dom_lo[0], P->dom_lo[1], P->dom_lo[2]), P->wt_d, P->st_d,
                       X ? P->perm_d : nullptr, X, F, static_cast<float*>(P->Fs_d));
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

int kernel_width(int k)
{
    switch (k)
    {
    case K_IB4: return KernelTraits<K_IB4>::W;
    case K_IB6: return KernelTraits<K_IB6>::W;
    case K_BS3: return KernelTraits<K_BS3>::W;
    case K_PWL: return KernelTraits<K_PWL>::W;
    case K_BS4: return KernelTraits<K_BS4>::W;
    case K_BS5: return KernelTraits<K_BS5>::W;
    case K_BS6: return KernelTraits<K_BS6>::W;
    default: return 6;
    }
}

// The same permutation driven by the sorted order, for records of D positions and
// D force components (side data): each thread gathers one point's record (all
// loads first) and writes it to its slot, so the writes are coalesced.  Measured
// faster than the scattered-write form for these 24 + 12-byte records.
template <int D, typename T>
__global__ void k_gather_points(const int* __restrict__ perm, int n, const double* __restrict__ X,
                                double* __restrict__ Xs, const T* __restrict__ F, T* __restrict__ Fs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t p = (size_t)perm[i];
    double x[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = X[D * p + d];
#pragma unroll
    for (int d = 0; d < D; ++d) Xs[D * (size_t)i + d] = x[d];
    if (F)
    {
        T f[D];
#pragma unroll
        for (int d = 0; d < D; ++d) f[d] = F[D * p + d];
#pragma unroll
        for (int d = 0; d < D; ++d) Fs[D * (size_t)i + d] = f[d];
    }
}

// ----------------------------------------------------------------------------- spread

template <typename T>
struct SpreadArgs
{
    const TileDesc* tiles;
    const PatchDev* patches;
    T* const* arrays;
    const double* X;   // positions in bin order (X[perm[k]] at slot k)
    const T* F;        // forces in bin order
    const int4* cells;
    const int* bin_start;
    int nb[3];
    int nsb[3];
    int bin;
    int dom_lo[3], dom_hi[3];
    int pshift[3];
    double dx[3];
    double vol;
    int fdepth;
    int narr;
    int gcw[3];
    int idx_gcw;
    int reach;
    int side;
    int mode;
    // FP32C 3-D: the points' stencil starts and weights in the level's frame
    const float* wt;
    const int* st;
    int nvar;
    int npts;
};

template <int NDIM, int TILE>
struct TileShape
{
    static constexpr int N = NDIM == 3 ? TILE * TILE * TILE : TILE * TILE;
};

template <typename T>
struct Acc
{
};

// Kahan-compensated add in FP32; plain add in FP64 (bitwise with the reference loop).
__device__ __forceinline__ void acc_add(double* s, double* /*c*/, int i, double v) { s[i] = s[i] + v; }
__device__ __forceinline__ void acc_add(float* s, float* c, int i, float v)
{
    const float y = v - c[i];
    const float t = s[i] + y;
    c[i] = (t - s[i]) - y;
    s[i] = t;
}

template <int NDIM, int KER, typename T, int TILE>
__global__ __launch_bounds__(WAVE) void k_spread_tiles(SpreadArgs<T> a)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr int TN = TileShape<NDIM, TILE>::N;
    constexpr bool FP64 = sizeof(T) == 8;
    __shared__ T s_acc[TN];
    __shared__ T s_cmp[FP64 ? 1 : TN];
    __shared__ T s_w[NDIM][W][NSLOT];
    __shared__ T s_F[NSLOT];
    __shared__ int s_base[NDIM][NSLOT];
    __shared__ int s_lo[NDIM][NSLOT];
    __shared__ int s_n[NDIM][NSLOT];
    __shared__ int s_act[NSLOT];

    const int lane = threadIdx.x;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;

    // array box (ghosted, side box along the axis) and this tile's index range
    int alo[3] = {0, 0, 0}, ahi[3] = {0, 0, 0}, an[3] = {1, 1, 1};
    int flo[3] = {0, 0, 0}, fhi[3] = {0, 0, 0}, fn[3] = {1, 1, 1};
    double xl[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
    {
        alo[d] = P.ilo[d] - a.gcw[d];
        ahi[d] = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
        an[d] = ahi[d] - alo[d] + 1;
        flo[d] = alo[d] + td.t[d] * TILE;
        fhi[d] = min(flo[d] + TILE - 1, ahi[d]);
        if (a.mode == IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        {
            flo[d] = max(flo[d], P.ilo[d]);
            fhi[d] = min(fhi[d], P.ihi[d] + (d == axis ? 1 : 0));
        }
        fn[d] = fhi[d] - flo[d] + 1;
        xl[d] = P.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);  // LEInteractor.cpp:3689
    }
    if (fn[0] < 1 || fn[1] < 1 || fn[2] < 1) return;  // tile wholly in the ghost region
    T* __restrict__ arr = a.arrays[td.patch * a.narr + (a.side ? td.comp : 0)];
    const size_t comp_off = a.side ? 0 : (size_t)td.comp * an[0] * an[1] * (NDIM == 3 ? an[2] : 1);
    const int fcomp = td.comp;
    const int tcount = fn[0] * fn[1] * (NDIM == 3 ? fn[2] : 1);

    // Cells whose points can reach this tile, inside the patch's Lagrangian ghost box.
    int clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
    bool any = true;
#pragma unroll
    for (int d = 0; d < NDIM; ++d)
    {
        clo[d] = max(flo[d] - a.reach, P.ilo[d] - a.idx_gcw);
        chi[d] = min(fhi[d] + a.reach, P.ihi[d] + a.idx_gcw);
        any = any && clo[d] <= chi[d];
    }
    // Per-axis (bin, image offset) lists in increasing (bin, offset) order.
    int lb[3][16], lo_[3][16], ln[3] = {0, 0, 0};
    if (any)
    {
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
        {
            const int N = a.pshift[d];
            const int omax = N > 0 ? 1 : 0;
            for (int o = -omax; o <= omax; ++o)
            {
                const int pl = max(clo[d] - o * N, a.dom_lo[d]);
                const int ph = min(chi[d] - o * N, a.dom_hi[d]);
                if (pl > ph) continue;
                const int bl = floor_div(pl - a.dom_lo[d], a.bin);
                const int bh = floor_div(ph - a.dom_lo[d], a.bin);
                for (int b = bl; b <= bh && ln[d] < 16; ++b)
                {
                    // insertion in (b, o) order
                    int k = ln[d]++;
                    while (k > 0 && (lb[d][k - 1] > b || (lb[d][k - 1] == b && lo_[d][k - 1] > o)))
                    {
                        lb[d][k] = lb[d][k - 1];
                        lo_[d][k] = lo_[d][k - 1];
                        --k;
                    }
                    lb[d][k] = b;
                    lo_[d][k] = o;
                }
            }
            any = any && ln[d] > 0;
        }
    }
    if (NDIM == 2) ln[2] = 1, lb[2][0] = 0, lo_[2][0] = 0;

    // Total points in the visited bins (uniform across the wavefront).
    int total = 0;
    if (any)
    {
        for (int kz = 0; kz < ln[2]; ++kz)
            for (int ky = 0; ky < ln[1]; ++ky)
                for (int kx = 0; kx < ln[0]; ++kx)
                {
                    const int b = bin_mem(NDIM, lb[0][kx], lb[1][ky], lb[2][kz], a.nb, a.nsb);
                    total += a.bin_start[b + 1] - a.bin_start[b];
                }
    }
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;
    if (total == 0 && !overwrite) return;

    // Load (accumulate) or clear (overwrite) the tile; LDS layout is TILE-strided.
    for (int e = lane; e < TN; e += WAVE)
    {
        T v = T(0);
        const int i0 = e % TILE;
        const int i1 = (e / TILE) % TILE;
        const int i2 = NDIM == 3 ? e / (TILE * TILE) : 0;
        if (!overwrite && i0 < fn[0] && i1 < fn[1] && i2 < fn[2])
        {
            const size_t g = (size_t)(flo[0] + i0 - alo[0]) +
                             (size_t)an[0] * ((size_t)(flo[1] + i1 - alo[1]) +
                                              (NDIM == 3 ? (size_t)an[1] * (size_t)(flo[2] + i2 - alo[2]) : 0));
            v = arr[comp_off + g];
        }
        s_acc[e] = v;
        if (!FP64) s_cmp[e] = T(0);
    }
    __syncthreads();

    const T inv_vol = T(1.0 / a.vol);
    for (int kz = 0; kz < ln[2] && total > 0; ++kz)
        for (int ky = 0; ky < ln[1]; ++ky)
            for (int kx = 0; kx < ln[0]; ++kx)
            {
                const int b = bin_mem(NDIM, lb[0][kx], lb[1][ky], lb[2][kz], a.nb, a.nsb);
                const int o[3] = {lo_[0][kx], lo_[1][ky], lo_[2][kz]};
                const int s0 = a.bin_start[b], s1 = a.bin_start[b + 1];
                for (int base = s0; base < s1; base += NSLOT)
                {
                    // ---- stage up to 64 points: weights and clipped, tile-intersected ranges
                    const int k = base + lane;
                    int act = 0;
                    if (k < s1)
                    {
                        const int4 c4 = a.cells[k];
                        const int cc[3] = {c4.x, c4.y, c4.z};
                        act = 1;
#pragma unroll
                        for (int d = 0; d < NDIM; ++d)
                        {
                            const int ci = cc[d] + o[d] * a.pshift[d];
                            act = act && ci >= clo[d] && ci <= chi[d];
                        }
                        if (act)
                        {
#pragma unroll
                            for (int d = 0; d < NDIM; ++d)
                            {
                                // X + Xshift with Xshift = (double)(offset) * dx  (LIndexSetData.cpp:116)
                                const double Xu = a.X[NDIM * (size_t)k + d];
                                const double Xs = Xu + (double)(o[d] * a.pshift[d]) * a.dx[d];
                                T w[W];
                                const int rel = weights_1d_u<KER>(Xs, Xu, xl[d], a.dx[d], w);
                                const int bse = P.ilo[d] + rel;
                                // clip to the array's ghost box (istart/istop, f.m4:1452-1457), then to the tile
                                const int l = max(max(bse, alo[d]), flo[d]);
                                const int h = min(min(bse + W - 1, ahi[d]), fhi[d]);
                                act = act && l <= h;
                                s_base[d][lane] = bse;
                                s_lo[d][lane] = l;
                                s_n[d][lane] = h - l + 1;
#pragma unroll
                                for (int j = 0; j < W; ++j) s_w[d][j][lane] = w[j];
                            }
                            const T Fv = a.F[(size_t)a.fdepth * k + fcomp];
                            s_F[lane] = FP64 ? Fv : Fv * inv_vol;
                        }
                    }
                    s_act[lane] = act;
                    __syncthreads();
                    // ---- apply the staged points one after another (single wavefront: program order)
                    const int nst = min(NSLOT, s1 - base);
                    for (int j = 0; j < nst; ++j)
                    {
                        if (!s_act[j]) continue;
                        const int n0 = s_n[0][j], n1 = s_n[1][j];
                        const int n2 = NDIM == 3 ? s_n[2][j] : 1;
                        const int cnt = n0 * n1 * n2;
                        const T Fv = s_F[j];
                        for (int e = lane; e < cnt; e += WAVE)
                        {
                            const int e0 = e % n0;
                            const int e1 = (e / n0) % n1;
                            const int e2 = e / (n0 * n1);
                            const int i0 = s_lo[0][j] + e0, i1 = s_lo[1][j] + e1;
                            const T w0 = s_w[0][i0 - s_base[0][j]][j];
                            const T w1 = s_w[1][i1 - s_base[1][j]][j];
                            int li = (i0 - flo[0]) + TILE * (i1 - flo[1]);
                            T c;
                            if (NDIM == 3)
                            {
                                const int i2 = s_lo[2][j] + e2;
                                const T w2 = s_w[2][i2 - s_base[2][j]][j];
                                li += TILE * TILE * (i2 - flo[2]);
                                if (FP64)
                                {
                                    const T vol = T(a.vol);
                                    // IB: (w0*(w1*(w2/vol)))*V (f.m4:1432-1463); BS: (((w0*w1)*w2)*V)/vol (:51-64)
                                    c = KT::IB ? (w0 * (w1 * (w2 / vol))) * Fv : (((w0 * w1) * w2) * Fv) / vol;
                                }
                                else
                                    c = (w0 * (w1 * w2)) * Fv;
                            }
                            else
                            {
                                if (FP64)
                                {
                                    const T vol = T(a.vol);
                                    c = KT::IB ? (w0 * (w1 / vol)) * Fv : ((w0 * w1) * Fv) / vol;
                                }
                                else
                                    c = (w0 * w1) * Fv;
                            }
                            // remap the compact tile index to the TILE-strided LDS layout
                            acc_add(s_acc, s_cmp, li, c);
                        }
                    }
                    __syncthreads();
                }
            }

    // Write the tile back (strided LDS layout: TILE per row).
    for (int e = lane; e < tcount; e += WAVE)
    {
        const int i0 = e % fn[0];
        const int i1 = (e / fn[0]) % fn[1];
        const int i2 = NDIM == 3 ? e / (fn[0] * fn[1]) : 0;
        const int li = i0 + TILE * (i1 + TILE * i2);
        const size_t g = (size_t)(flo[0] + i0 - alo[0]) +
                         (size_t)an[0] * ((size_t)(flo[1] + i1 - alo[1]) +
                                          (NDIM == 3 ? (size_t)an[1] * (size_t)(flo[2] + i2 - alo[2]) : 0));
        T v = s_acc[li];
        if (!FP64) v = v - s_cmp[li];
        arr[comp_off + g] = v;
    }
}

// ----------------------------------------------------------------------------- interp

template <typename T>
struct InterpArgs
{
    const PatchDev* patches;
    int npatches;
    const int* bin_patch_start;
    const int* bin_patches;
    T* const* arrays;
    const double* X;   // positions in bin order
    T* U;
    const int* perm;
    const int4* cells;
    const int* keys_sorted;
    const int* bin_start;
    int ksub;          // keys_sorted >> ksub is the bin
    int n;
    double dx[3];
    int gcw[3];
    int narr;
    int fdepth;
    int depth;
    int side;
    // FP32C 3-D: the points' stencil starts and weights in the level's frame
    // (null: evaluate them from X and the level's x_lower / lower index)
    const float* wt;
    const int* st;
    int nvar;
    double lev_xlo[3];
    int lev_lo[3];
};

// Compensated (FP32) or plain (FP64, same order as the reference) sum.
template <typename T>
struct Sum;
template <>
struct Sum<double>
{
    double s = 0.0;
    __device__ void add(double v) { s = s + v; }
    __device__ double get() const { return s; }
};
template <>
struct Sum<float>
{
    float s = 0.0f, c = 0.0f;
    __device__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ float get() const { return s - c; }
};

template <int NDIM, int KER, typename T>
__global__ __launch_bounds__(256) void k_interp(InterpArgs<T> a)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr bool FP64 = sizeof(T) == 8;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const int key = a.keys_sorted[k] >> a.ksub;
    const int4 c4 = a.cells[k];
    const int cc[3] = {c4.x, c4.y, c4.z};
    // owning patch: the one whose interior holds the cell (LDataManager.cpp:748-753)
    int pid = -1;
    if (key >= 0)
    {
        const int b0 = a.bin_patch_start[key], b1 = a.bin_patch_start[key + 1];
        for (int j = b0; j < b1 && pid < 0; ++j)
        {
            const PatchDev& Q = a.patches[a.bin_patches[j]];
            bool in = true;
#pragma unroll
            for (int d = 0; d < NDIM; ++d) in = in && cc[d] >= Q.slo[d] && cc[d] <= Q.shi[d];
            if (in) pid = a.bin_patches[j];
        }
    }
    if (pid < 0) return;
    const PatchDev P = a.patches[pid];
    const int p = a.perm[k];
    double Xp[3];
#pragma unroll
    for (int d = 0; d < NDIM; ++d) Xp[d] = a.X[NDIM * (size_t)k + d];

    const int narrays = a.side ? NDIM : 1;
    for (int ax = 0; ax < narrays; ++ax)
    {
        const int axis = a.side ? ax : -1;
        int alo[3] = {0, 0, 0}, ahi[3] = {0, 0, 0}, an[3] = {1, 1, 1};
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, bse[3] = {0, 0, 0};
        T w[3][W];
#pragma unroll
        for (int d = 0; d < NDIM; ++d)
        {
            alo[d] = P.ilo[d] - a.gcw[d];
            ahi[d] = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
            an[d] = ahi[d] - alo[d] + 1;
            const double xl = P.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);  // LEInteractor.cpp:2465
            const int rel = weights_1d<KER>(Xp[d], xl, a.dx[d], w[d]);
            bse[d] = P.ilo[d] + rel;
            lo[d] = max(bse[d], alo[d]);
            hi[d] = min(bse[d] + W - 1, ahi[d]);
        }
        const T* __restrict__ arr = a.arrays[pid * a.narr + (a.side ? ax : 0)];
        const size_t ncomp = (size_t)an[0] * an[1] * (NDIM == 3 ? an[2] : 1);
        const int ncomps = a.side ? 1 : a.depth;
        for (int cd = 0; cd < ncomps; ++cd)
        {
            const T* __restrict__ u = arr + cd * ncomp;
            Sum<T> acc;
            if (NDIM == 3)
            {
                for (int i2 = lo[2]; i2 <= hi[2]; ++i2)
                {
                    const T w2 = w[2][i2 - bse[2]];
                    for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                    {
                        const T w1 = w[1][i1 - bse[1]];
                        const T* row = u + (size_t)(i1 - alo[1]) * an[0] + (size_t)(i2 - alo[2]) * an[0] * an[1] - alo[0];
                        if (FP64)
                        {
                            for (int i0 = lo[0]; i0 <= hi[0]; ++i0)
                            {
                                const T w0 = w[0][i0 - bse[0]];
                                // IB: V + (w0*(w1*w2))*u (f.m4:1288-1328); BS: V + ((w0*w1)*w2)*u (:24-37)
                                acc.add(KT::IB ? (w0 * (w1 * w2)) * row[i0] : ((w0 * w1) * w2) * row[i0]);
                            }
                        }
                        else
                        {
                            // FP32: one partial sum per grid row (W terms), compensated across rows
                            const T w12 = w1 * w2;
                            T part = T(0);
                            for (int i0 = lo[0]; i0 <= hi[0]; ++i0) part = fmaf(w[0][i0 - bse[0]], row[i0], part);
                            acc.add(w12 * part);
                        }
                    }
                }
            }
            else
            {
                for (int i1 = lo[1]; i1 <= hi[1]; ++i1)
                {
                    const T w1 = w[1][i1 - bse[1]];
                    const T* row = u + (size_t)(i1 - alo[1]) * an[0] - alo[0];
                    if (FP64)
                    {
                        for (int i0 = lo[0]; i0 <= hi[0]; ++i0) acc.add((w[0][i0 - bse[0]] * w1) * row[i0]);
                    }
                    else
                    {
                        T part = T(0);
                        for (int i0 = lo[0]; i0 <= hi[0]; ++i0) part = fmaf(w[0][i0 - bse[0]], row[i0], part);
                        acc.add(w1 * part);
                    }
                }
            }
            a.U[(size_t)a.fdepth * p + (a.side ? ax : cd)] = acc.get();
        }
    }
}

// ----------------------------------------------------------------------------- periodic ghost fill

// Periodic ghost fill of one array of a single-patch, fully periodic level:
// every value outside the core [g, g+N) of each axis takes the value of its
// periodic image, which lies in the core (the ghost width never exceeds the
// period).  Only the ghost shell is enumerated: z slabs (whole planes), then y
// slabs inside the core z range, then x slabs inside the core y and z ranges.
template <typename T>
__global__ void k_fill_periodic(T* __restrict__ arr, int n0, int n1, int n2, int g0, int g1, int g2, int N0, int N1,
                                int N2, long sz, long sy, long sx)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i0, i1, i2;
    if (t < sz)
    {
        const long plane = (long)n0 * n1;
        const int q = (int)(t / plane);                       // which ghost plane
        const long r = t - (long)q * plane;
        i2 = q < g2 ? q : q - g2 + g2 + N2;                   // lower planes, then upper planes
        i1 = (int)(r / n0);
        i0 = (int)(r - (long)i1 * n0);
    }
    else if ((t -= sz) < sy)
    {
        const long slab = (long)(n1 - N1) * n0;              // per core z plane
        const int z = (int)(t / slab);
        const long r = t - (long)z * slab;
        const int q = (int)(r / n0);
        i0 = (int)(r - (long)q * n0);
        i1 = q < g1 ? q : q - g1 + g1 + N1;
        i2 = g2 + z;
    }
    else if ((t -= sy) < sx)
    {
        const int w = n0 - N0;                               // per core (y, z) row
        const long row = t / w;
        const int q = (int)(t - row * w);
        i0 = q < g0 ? q : q - g0 + g0 + N0;
        i1 = g1 + (int)(row % N1);
        i2 = g2 + (int)(row / N1);
    }
    else
        return;
    int j0 = i0 - g0, j1 = i1 - g1, j2 = i2 - g2;
    j0 += j0 < 0 ? N0 : (j0 >= N0 ? -N0 : 0);
    j1 += j1 < 0 ? N1 : (j1 >= N1 ? -N1 : 0);
    j2 += j2 < 0 ? N2 : (j2 >= N2 ? -N2 : 0);
    arr[(long)i0 + (long)n0 * (i1 + (long)n1 * i2)] =
        arr[(long)(j0 + g0) + (long)n0 * ((j1 + g1) + (long)n1 * (j2 + g2))];
}

// Same-level ghost fill (the RefineSchedule::fillData step before interp,
// LDataManager.cpp:742-745, restricted to copies between patches of one level):
// every ghost value of every patch array whose index -- wrapped across periodic
// directions -- lies in another patch's (or, periodically, its own) interior
// side box is copied from there.  Ghost values outside the domain of a
// non-periodic direction, or not covered by the level (coarse-fine boundary),
// are left as they are.
template <typename T>
struct FillArgs
{
    const int4* segs;
    const PatchDev* patches;
    T* const* arrays;
    const int* bin_patch_start;
    const int* bin_patches;
    int ndim, narr, side, depth;
    int gcw[3];
    int dom_lo[3], dom_hi[3];
    int pshift[3];
    int nb[3], nsb[3];
    int bin;
};

template <typename T>
__global__ __launch_bounds__(256) void k_fill_level(FillArgs<T> a)
{
    const int4 sg = a.segs[blockIdx.x];
    if ((int)threadIdx.x >= sg.w) return;
    const int pid = sg.x, arr_i = sg.y;
    const int axis = a.side ? arr_i : -1;
    const PatchDev P = a.patches[pid];
    int n[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
    for (int d = 0; d < a.ndim; ++d)
    {
        lo[d] = P.ilo[d] - a.gcw[d];
        n[d] = P.ihi[d] - P.ilo[d] + 1 + 2 * a.gcw[d] + (d == axis ? 1 : 0);
    }
    const long e = (long)sg.z + threadIdx.x;
    int idx[3] = {(int)(e % n[0]), (int)((e / n[0]) % n[1]), (int)(e / ((long)n[0] * n[1]))};
    bool interior = true;
    int w[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
    {
        if (d >= a.ndim)
            continue;
        const int i = lo[d] + idx[d];
        const int top = P.ihi[d] + (d == axis ? 1 : 0);
        interior = interior && i >= P.ilo[d] && i <= top;
        const int N = a.pshift[d];
        const int dtop = a.dom_hi[d] + (d == axis ? 1 : 0);
        if (N > 0)
            w[d] = a.dom_lo[d] + (((i - a.dom_lo[d]) % N) + N) % N;
        else
        {
            if (i < a.dom_lo[d] || i > dtop) return;  // physical boundary ghost: not filled here
            w[d] = i;
        }
    }
    if (interior) return;
    // owner: a patch whose interior side box holds the wrapped index; its cell,
    // or for a side axis also the cell below (the face is that cell's upper face)
    int q = -1;
    for (int tries = 0; tries < 2 && q < 0; ++tries)
    {
        if (tries == 1 && axis < 0) break;
        int c[3] = {0, 0, 0};
        bool ok = true;
        for (int d = 0; d < a.ndim; ++d)
        {
            c[d] = w[d] - (tries == 1 && d == axis ? 1 : 0);
            ok = ok && c[d] >= a.dom_lo[d] && c[d] <= a.dom_hi[d];
        }
        if (!ok) continue;
        int b[3] = {0, 0, 0};
        for (int d = 0; d < a.ndim; ++d) b[d] = (c[d] - a.dom_lo[d]) / a.bin;
        const int key = bin_mem(a.ndim, b[0], b[1], b[2], a.nb, a.nsb);
        for (int j = a.bin_patch_start[key]; j < a.bin_patch_start[key + 1] && q < 0; ++j)
        {
            const PatchDev& Q = a.patches[a.bin_patches[j]];
            bool in = true;
            for (int d = 0; d < a.ndim; ++d)
                in = in && w[d] >= Q.ilo[d] && w[d] <= Q.ihi[d] + (d == axis ? 1 : 0);
            if (in) q = a.bin_patches[j];
        }
    }
    if (q < 0) return;  // not covered by this level
    const PatchDev Q = a.patches[q];
    int qn[3] = {1, 1, 1};
    long src = 0, stride = 1;
    for (int d = 0; d < a.ndim; ++d)
    {
        qn[d] = Q.ihi[d] - Q.ilo[d] + 1 + 2 * a.gcw[d] + (d == axis ? 1 : 0);
        src += stride * (long)(w[d] - (Q.ilo[d] - a.gcw[d]));
        stride *= qn[d];
    }
    const long ptot = (long)n[0] * n[1] * n[2], qtot = stride;
    T* __restrict__ dst = a.arrays[pid * a.narr + arr_i];
    const T* __restrict__ from = a.arrays[q * a.narr + arr_i];
    const int ncomp = a.side ? 1 : a.depth;
    for (int cd = 0; cd < ncomp; ++cd) dst[(long)cd * ptot + e] = from[(long)cd * qtot + src];
}

} // namespace

#include "le_kernels3d.h"

// ----------------------------------------------------------------------------- host side

namespace
{
template <typename F>
int dispatch(int ndim, int kernel, int precision, F&& f)
{
    auto by_kernel = [&](auto nd, auto prec) -> int {
        switch (kernel)
        {
        case K_IB4: return f(nd, std::integral_constant<int, K_IB4>{}, prec);
        case K_IB6: return f(nd, std::integral_constant<int, K_IB6>{}, prec);
        case K_BS3: return f(nd, std::integral_constant<int, K_BS3>{}, prec);
        case K_PWL: return f(nd, std::integral_constant<int, K_PWL>{}, prec);
        case K_BS4: return f(nd, std::integral_constant<int, K_BS4>{}, prec);
        case K_BS5: return f(nd, std::integral_constant<int, K_BS5>{}, prec);
        case K_BS6: return f(nd, std::integral_constant<int, K_BS6>{}, prec);
        default: return fail(IBAMR_ERR_BAD_KERNEL, "kernel has no GPU spread/interp");
        }
    };
    auto by_prec = [&](auto nd) -> int {
        if (precision == IBAMR_FP64) return by_kernel(nd, (double)0);
        return by_kernel(nd, (float)0);
    };
    if (ndim == 3) return by_prec(std::integral_constant<int, 3>{});
    return by_prec(std::integral_constant<int, 2>{});
}

template <typename T>
void fill_spread_args(const ibamr_plan* P, int slot, const void* F, const double* X, int mode, SpreadArgs<T>& a)
{
    a.tiles = P->tiles_d;
    a.patches = P->patches_d;
    a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
    a.X = X;
    a.F = static_cast<const T*>(F);
    a.cells = P->cell_sorted_d;
    a.bin_start = P->bin_start_d;
    for (int d = 0; d < 3; ++d)
    {
        a.nb[d] = P->nb[d];
        a.nsb[d] = P->nsb[d];
        a.dom_lo[d] = P->dom_lo[d];
        a.dom_hi[d] = P->dom_hi[d];
        a.pshift[d] = P->pshift[d];
        a.dx[d] = P->dx[d];
        a.gcw[d] = P->data_gcw[d];
    }
    a.bin = P->bin;
    a.vol = P->ndim == 3 ? (P->dx[0] * P->dx[1]) * P->dx[2] : P->dx[0] * P->dx[1];
    a.fdepth = P->fdepth;
    a.narr = P->narr;
    a.idx_gcw = P->idx_gcw;
    a.reach = P->reach;
    a.side = P->centering == IBAMR_SIDE;
    a.mode = mode;
    a.wt = P->wt_d;
    a.st = P->st_d;
    a.nvar = P->nvar;
    a.npts = P->npoints;
}
} // namespace

namespace
{
void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

// The spread's work list for the plan's tile shape, uploaded to tiles_d
// (replacing any previous list).
int build_tiles(ibamr_plan* P)
{
    // work list: every tile of every array of every patch.  XCD-aware launch
    // order: workgroup i runs on XCD i % 8.  Tiles are grouped by (patch, z row,
    // y row), each group holding that row of every array, and the groups are
    // dealt to the XCDs in turn; each XCD runs its groups in (patch, z, y) order.
    // A group's tiles read the same point data through one L2 (all arrays, x
    // neighbours), every XCD gets a share of every region (point clouds are
    // rarely uniform, so contiguous eighths of the domain would leave some XCDs
    // idle), and all XCDs sweep the domain together, so the data they re-read is
    // recent in the shared last-level cache.
    std::vector<TileDesc> tiles;
    {
        const int ncomp_tiles = P->centering == IBAMR_SIDE ? P->ndim : P->depth;
        std::vector<std::vector<TileDesc>> q(8);
        size_t g = 0;
        for (int i = 0; i < P->npatches; ++i)
        {
            const PatchDev& pd = P->patches_h[i];
            std::vector<std::array<int, 3>> nt(ncomp_tiles, std::array<int, 3>{1, 1, 1});
            int mt[3] = {1, 1, 1};
            for (int c = 0; c < ncomp_tiles; ++c)
            {
                const int axis = P->centering == IBAMR_SIDE ? c : -1;
                for (int d = 0; d < P->ndim; ++d)
                {
                    const int n = pd.ihi[d] - pd.ilo[d] + 1 + 2 * P->data_gcw[d] + (d == axis ? 1 : 0);
                    nt[c][d] = (n + P->tsz[d] - 1) / P->tsz[d];
                    mt[d] = std::max(mt[d], nt[c][d]);
                }
            }
            for (int tz = 0; tz < mt[2]; ++tz)
                for (int ty = 0; ty < mt[1]; ++ty, ++g)
                    for (int c = 0; c < ncomp_tiles; ++c)
                        if (ty < nt[c][1] && tz < nt[c][2])
                            for (int tx = 0; tx < nt[c][0]; ++tx) q[g % 8].push_back(TileDesc{i, c, {tx, ty, tz}});
        }
        size_t per = 0;
        for (const auto& v : q) per = std::max(per, v.size());
        tiles.assign(8 * per, TileDesc{-1, 0, {0, 0, 0}});
        for (int x = 0; x < 8; ++x)
            for (size_t k = 0; k < q[x].size(); ++k) tiles[8 * k + x] = q[x][k];
    }
    dfree(P->tiles_d);
    P->tiles_d = nullptr;
    P->ntiles = (int)tiles.size();
    HIP_OK(hipMalloc(&P->tiles_d, sizeof(TileDesc) * std::max<size_t>(1, tiles.size())));
    if (!tiles.empty()) HIP_OK(hipMemcpy(P->tiles_d, tiles.data(), sizeof(TileDesc) * tiles.size(), hipMemcpyHostToDevice));
    return IBAMR_OK;
}

// Spread tile heights instantiated for each form (3-D); the first is the default.
const std::vector<int>& tile_heights(int precision)
{
    static const std::vector<int> fp64 = {32, 16}, fp32 = {16, 8, 24};
    return precision == IBAMR_FP64 ? fp64 : fp32;
}

} // namespace

// ----------------------------------------------------------------------------- redistribution

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
            // IndexUtilities::getCellIndex (IndexUtilities-inl.h:52-72) on the level domain
            const double dXl = X - a.xlo[d], dXu = X - a.xhi[d];
            const double v = fabs(dXl) <= fabs(dXu) ? (double)a.dom_lo[d] + floor(dXl / a.dx[d])
                                                    : ((double)a.dom_hi[d] + floor(dXu / a.dx[d])) + 1.0;
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
} // namespace

// ----------------------------------------------------------------------------- physical boundary

namespace
{
// The LINEAR Robin ghost operators of CartCellRobinPhysBdryOp / CartSideRobinPhysBdryOp
// (cartphysbdryop{2,3}d.f.m4), adjoint (accumulateFromPhysicalBoundaryData) and
// forward (setPhysicalBoundaryConditions).  One workgroup per patch walks the
// patch's boundary boxes in order; inside an operator, threads take the index
// sets that the operator's updates keep fixed (the columns normal to a face,
// the index along an edge, ...) and each runs the reference's loop over the
// rest, so every value sees the reference's sequence of operations.  Corners
// (a few gcw^3 values) run on one thread.
template <typename T>
struct BdryArgs
{
    const int4* patches_list;
    const int* boxes;
    const double* abg;
    const PatchDev* patches;
    T* const* arrays;
    int ndim, side, depth, narr, ncoef, gcw, adjoint;  // ncoef: coefficient triples per box
    double dx[3];
};

template <typename T>
struct BPatch
{
    int ndim, side, gcw;
    int ilo[3], ihi[3];
    long vol;  // values per depth component (cell data)
    T* arr[3];
    __device__ T* at(int a, int comp, const int* x) const
    {
        long off = 0, mul = 1;
        for (int d = 0; d < ndim; ++d)
        {
            const int n = ihi[d] - ilo[d] + 1 + 2 * gcw + ((side && d == a) ? 1 : 0);
            off += (long)(x[d] - (ilo[d] - gcw)) * mul;
            mul *= n;
        }
        return side ? arr[a] + off : arr[0] + (long)comp * vol + off;
    }
};

__device__ inline int bp_tangential_dev(int ndim, int dir, int* t)
{
    int m = 0;
    for (int d = ndim - 1; d >= 0; --d)
        if (d != dir) t[m++] = d;
    return m;
}

// column c of a face: coordinates of the tangential dims
__device__ inline void face_column(int ndim, int dir, const int* blo, const int* bhi, long c, int* x)
{
    int t[2];
    const int nt = bp_tangential_dev(ndim, dir, t);
    const int n_in = bhi[t[nt - 1]] - blo[t[nt - 1]] + 1;
    x[t[nt - 1]] = blo[t[nt - 1]] + (int)(c % n_in);
    if (nt > 1) x[t[0]] = blo[t[0]] + (int)(c / n_in);
}

__device__ inline long face_columns(int ndim, int dir, const int* blo, const int* bhi)
{
    long m = 1;
    for (int d = 0; d < ndim; ++d)
        if (d != dir) m *= (long)(bhi[d] - blo[d] + 1);
    return m;
}

template <typename T>
__device__ void bd_cc_op1(const BPatch<T>& P, int a, int comp, int loc, const int* dhi, const int* blo,
                          const int* bhi, double ca, double cb, double cg, int adjoint, const double* dx)
{
    const int ndim = P.ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = dx[loc / ndim];  // as the Fortran indexes it
    const int sgn = lower ? -1 : 1;
    const int ig = lower ? P.ilo[dir] - 1 : dhi[dir] + 1, ii = lower ? P.ilo[dir] : dhi[dir];
    const long ncol = face_columns(ndim, dir, blo, bhi);
    for (long c = threadIdx.x; c < ncol; c += blockDim.x)
    {
        int x[3] = {0, 0, 0};
        face_column(ndim, dir, blo, bhi, c, x);
        for (int i = 0; i < P.gcw; ++i)
        {
            const double n = 1.0 + 2.0 * i;
            const double f_i = -(ca * n * h - 2.0 * cb) / (ca * n * h + 2.0 * cb);
            const double f_g = 2.0 * n * h / (ca * n * h + 2.0 * cb);
            int xg[3] = {x[0], x[1], x[2]}, xi[3] = {x[0], x[1], x[2]};
            xg[dir] = ig + sgn * i;
            xi[dir] = ii - sgn * i;
            if (adjoint)
            {
                const double u_g = (double)*P.at(a, comp, xg);
                T* t = P.at(a, comp, xi);
                *t = (T)((double)*t + f_i * u_g);
            }
            else
            {
                const double u_i = (double)*P.at(a, comp, xi);
                *P.at(a, comp, xg) = (T)(f_i * u_i + f_g * cg);
            }
        }
    }
}

template <typename T>
__device__ void bd_sc_op1(const BPatch<T>& P, int loc, const int* blo, const int* bhi, double ca, double cb, double cg,
                          int adjoint, const double* dx)
{
    const int ndim = P.ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = dx[loc / ndim];
    const int sgn = lower ? -1 : 1;
    const int ib = lower ? P.ilo[dir] : P.ihi[dir] + 1;
    const bool dirichlet = fabs(cb) < 1.0e-12;
    const long ncol = face_columns(ndim, dir, blo, bhi);
    for (long c = threadIdx.x; c < ncol; c += blockDim.x)
    {
        int x[3] = {0, 0, 0};
        face_column(ndim, dir, blo, bhi, c, x);
        int xb[3] = {x[0], x[1], x[2]};
        xb[dir] = ib;
        T* ub = P.at(dir, 0, xb);
        double u_b;
        if (dirichlet)
        {
            u_b = cg / ca;
            *ub = (T)u_b;
        }
        else
            u_b = (double)*ub;
        for (int i = 1; i <= P.gcw; ++i)
        {
            double f_i, f_b, f_g = 0.0;
            if (dirichlet)
            {
                f_i = -1.0;
                f_b = 2.0;
            }
            else
            {
                const double n = 2.0 * i;
                f_i = 1.0;
                f_b = -ca * n * h / cb;
                f_g = n * h / cb;
            }
            int xg[3] = {x[0], x[1], x[2]}, xi[3] = {x[0], x[1], x[2]};
            xg[dir] = ib + sgn * i;
            xi[dir] = ib - sgn * i;
            if (adjoint)
            {
                const double u_g = (double)*P.at(dir, 0, xg);
                T* t = P.at(dir, 0, xi);
                *t = (T)((double)*t + f_i * u_g);
                *ub = (T)((double)*ub + f_b * u_g);
            }
            else
            {
                const double u_i = (double)*P.at(dir, 0, xi);
                *P.at(dir, 0, xg) = (T)(dirichlet ? f_i * u_i + f_b * u_b : f_i * u_i + f_b * u_b + f_g * cg);
            }
        }
    }
}

__device__ inline void bd_edge(int loc, int* e, int* n1, int* n2, int* up1, int* up2)
{
    if (loc < 4)
    {
        *e = 0, *n1 = 1, *n2 = 2, *up1 = loc & 1, *up2 = (loc >> 1) & 1;
    }
    else if (loc < 8)
    {
        *e = 1, *n1 = 0, *n2 = 2, *up1 = (loc - 4) >> 1, *up2 = (loc - 4) & 1;
    }
    else
    {
        *e = 2, *n1 = 0, *n2 = 1, *up1 = (loc - 8) & 1, *up2 = (loc - 8) >> 1;
    }
}

// cell data, 3-D edges: every update keeps the edge index (one thread each)
template <typename T>
__device__ void bd_cc_edge(const BPatch<T>& P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bd_edge(loc, &e, &n1, &n2, &u1, &u2);
    const int b1 = u1 ? P.ihi[n1] : P.ilo[n1], b2 = u2 ? P.ihi[n2] : P.ilo[n2];
    const int s1 = u1 ? 1 : -1, s2 = u2 ? 1 : -1;
    for (int ce = flo[e] + (int)threadIdx.x; ce <= fhi[e]; ce += blockDim.x)
        for (int c2 = flo[n2]; c2 <= fhi[n2]; ++c2)
            for (int c1 = flo[n1]; c1 <= fhi[n1]; ++c1)
            {
                const int m1 = b1 + (b1 - c1 + s1), m2 = b2 + (b2 - c2 + s2);
                int x[3], xm[3], xa[3], xb[3], xc[3], xd[3];
                x[e] = xm[e] = xa[e] = xb[e] = xc[e] = xd[e] = ce;
                x[n1] = c1, x[n2] = c2;
                xm[n1] = m1, xm[n2] = m2;
                xa[n1] = b1, xa[n2] = c2;
                xb[n1] = c1, xb[n2] = b2;
                xc[n1] = b1, xc[n2] = m2;
                xd[n1] = m1, xd[n2] = b2;
                if (adjoint)
                {
                    const double Ug = (double)*P.at(0, comp, x);
                    T* t;
                    t = P.at(0, comp, xm), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xa), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xb), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xc), *t = (T)((double)*t - Ug);
                    t = P.at(0, comp, xd), *t = (T)((double)*t - Ug);
                }
                else
                {
                    const double v = (double)*P.at(0, comp, xm) +
                                     ((double)*P.at(0, comp, xa) - (double)*P.at(0, comp, xc)) +
                                     ((double)*P.at(0, comp, xb) - (double)*P.at(0, comp, xd));
                    *P.at(0, comp, x) = (T)v;
                }
            }
}

// cell data, corners (codim == ndim): one thread
template <typename T>
__device__ void bd_cc_corner(const BPatch<T>& P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    if (threadIdx.x != 0) return;
    const int nd = P.ndim;
    int b[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        b[d] = up ? P.ihi[d] : P.ilo[d];
        s[d] = up ? 1 : -1;
    }
    const int zl = nd == 3 ? flo[2] : 0, zh = nd == 3 ? fhi[2] : 0;
    for (int z = zl; z <= zh; ++z)
        for (int y = flo[1]; y <= fhi[1]; ++y)
            for (int xx = flo[0]; xx <= fhi[0]; ++xx)
            {
                const int x[3] = {xx, y, z};
                int m[3] = {0, 0, 0};
                for (int d = 0; d < nd; ++d) m[d] = b[d] + (b[d] - x[d] + s[d]);
                int pk[3][3], mk[3][3];
                for (int d = 0; d < 3; ++d)
                    for (int o = 0; o < 3; ++o)
                    {
                        pk[d][o] = o == d ? x[o] : b[o];
                        mk[d][o] = o == d ? m[o] : b[o];
                    }
                if (adjoint)
                {
                    const double Ug = (double)*P.at(0, comp, x);
                    T* t = P.at(0, comp, m);
                    *t = (T)((double)*t + Ug);
                    for (int d = 0; d < nd; ++d) t = P.at(0, comp, pk[d]), *t = (T)((double)*t + Ug);
                    for (int d = 0; d < nd; ++d) t = P.at(0, comp, mk[d]), *t = (T)((double)*t - Ug);
                }
                else
                {
                    double v = (double)*P.at(0, comp, m);
                    for (int d = 0; d < nd; ++d)
                        v = v + ((double)*P.at(0, comp, pk[d]) - (double)*P.at(0, comp, mk[d]));
                    *P.at(0, comp, x) = (T)v;
                }
            }
}

// side data, 3-D edges: component n1 extrapolated along n2 (columns keep the edge
// index and n1), then component n2 along n1 (columns keep the edge index and n2)
template <typename T>
__device__ void bd_sc_edge(const BPatch<T>& P, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bd_edge(loc, &e, &n1, &n2, &u1, &u2);
    int bd[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
    bd[n1] = u1 ? P.ihi[n1] : P.ilo[n1];
    bd[n2] = u2 ? P.ihi[n2] : P.ilo[n2];
    sh[n1] = u1 ? -1 : 1;
    sh[n2] = u2 ? -1 : 1;
    for (int part = 0; part < 2; ++part)
    {
        const int comp = part == 0 ? n1 : n2, other = part == 0 ? n2 : n1;
        int lo[3] = {flo[0], flo[1], flo[2]}, hi[3] = {fhi[0], fhi[1], fhi[2]};
        hi[comp] += 1;
        const int ne = hi[e] - lo[e] + 1, nc = hi[comp] - lo[comp] + 1;
        for (int c = threadIdx.x; c < ne * nc; c += blockDim.x)
        {
            int x[3];
            x[e] = lo[e] + c % ne;
            x[comp] = lo[comp] + c / ne;
            if (!(x[comp] < P.ilo[comp] || x[comp] > P.ihi[comp] + 1)) continue;
            for (int o = lo[other]; o <= hi[other]; ++o)
            {
                x[other] = o;
                const double del = (double)abs(o - bd[other]);
                int xb[3] = {x[0], x[1], x[2]}, xs[3] = {x[0], x[1], x[2]};
                xb[other] = bd[other];
                xs[other] = bd[other] + sh[other];
                if (adjoint)
                {
                    const double u_g = (double)*P.at(comp, 0, x);
                    T* t = P.at(comp, 0, xb);
                    *t = (T)((double)*t + (1.0 + del) * u_g);
                    t = P.at(comp, 0, xs);
                    *t = (T)((double)*t - del * u_g);
                }
                else
                    *P.at(comp, 0, x) =
                        (T)((1.0 + del) * (double)*P.at(comp, 0, xb) - del * (double)*P.at(comp, 0, xs));
            }
        }
        __syncthreads();
    }
}

// side data, corners: one thread
template <typename T>
__device__ void bd_sc_corner(const BPatch<T>& P, int loc, const int* flo, const int* fhi, int adjoint)
{
    if (threadIdx.x != 0) return;
    const int nd = P.ndim;
    int bd[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        bd[d] = up ? P.ihi[d] : P.ilo[d];
        sh[d] = up ? -1 : 1;
    }
    for (int comp = 0; comp < nd; ++comp)
    {
        int lo[3] = {flo[0], flo[1], nd == 3 ? flo[2] : 0}, hi[3] = {fhi[0], fhi[1], nd == 3 ? fhi[2] : 0};
        hi[comp] += 1;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int xx = lo[0]; xx <= hi[0]; ++xx)
                {
                    const int x[3] = {xx, y, z};
                    if (!(x[comp] < P.ilo[comp] || x[comp] > P.ihi[comp] + 1)) continue;
                    double dsum = 1.0, del[3] = {0.0, 0.0, 0.0};
                    for (int o = 0; o < nd; ++o)
                        if (o != comp)
                        {
                            del[o] = (double)abs(x[o] - bd[o]);
                            dsum = dsum + del[o];
                        }
                    int xb[3] = {x[0], x[1], x[2]};
                    for (int o = 0; o < nd; ++o)
                        if (o != comp) xb[o] = bd[o];
                    if (adjoint)
                    {
                        const double u_g = (double)*P.at(comp, 0, x);
                        T* t = P.at(comp, 0, xb);
                        *t = (T)((double)*t + dsum * u_g);
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                t = P.at(comp, 0, xs);
                                *t = (T)((double)*t - del[o] * u_g);
                            }
                    }
                    else
                    {
                        double v = dsum * (double)*P.at(comp, 0, xb);
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                v = v - del[o] * (double)*P.at(comp, 0, xs);
                            }
                        *P.at(comp, 0, x) = (T)v;
                    }
                }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_phys_bdry(BdryArgs<T> a)
{
    const int4 pl = a.patches_list[blockIdx.x];
    const int pid = pl.x;
    const PatchDev& Q = a.patches[pid];
    BPatch<T> P;
    P.ndim = a.ndim;
    P.side = a.side;
    P.gcw = a.gcw;
    long vol = 1;
    for (int d = 0; d < 3; ++d)
    {
        P.ilo[d] = d < a.ndim ? Q.ilo[d] : 0;
        P.ihi[d] = d < a.ndim ? Q.ihi[d] : 0;
        if (d < a.ndim) vol *= (long)(Q.ihi[d] - Q.ilo[d] + 1 + 2 * a.gcw);
    }
    P.vol = vol;
    for (int d = 0; d < 3; ++d) P.arr[d] = d < a.narr ? a.arrays[pid * a.narr + d] : nullptr;
    if (!a.side) P.arr[0] = a.arrays[pid];
    const int phases_adj[4] = {3, 2, 10, 11}, phases_fwd[4] = {11, 10, 2, 3};
    for (int ph = 0; ph < 4; ++ph)
    {
        const int phase = a.adjoint ? phases_adj[ph] : phases_fwd[ph];
        for (int k = pl.y; k < pl.y + pl.z; ++k)
        {
            const int* bx = a.boxes + 8 * k;
            const int codim = bx[0], loc = bx[1];
            const int lo[3] = {bx[2], bx[3], bx[4]}, hi[3] = {bx[5], bx[6], bx[7]};
            const int ncomp = a.side ? 1 : a.depth;
            bool did = false;
            if (phase == 3 && codim == 3 && a.ndim == 3)
            {
                for (int c = 0; c < ncomp; ++c)
                {
                    if (a.side)
                        bd_sc_corner(P, loc, lo, hi, a.adjoint);
                    else
                        bd_cc_corner(P, c, loc, lo, hi, a.adjoint);
                    __syncthreads();
                }
                did = true;
            }
            else if (phase == 2 && codim == 2)
            {
                for (int c = 0; c < ncomp; ++c)
                {
                    if (a.ndim == 2)
                    {
                        if (a.side)
                            bd_sc_corner(P, loc, lo, hi, a.adjoint);
                        else
                            bd_cc_corner(P, c, loc, lo, hi, a.adjoint);
                    }
                    else if (a.side)
                        bd_sc_edge(P, loc, lo, hi, a.adjoint);
                    else
                        bd_cc_edge(P, c, loc, lo, hi, a.adjoint);
                    __syncthreads();
                }
                did = true;
            }
            else if ((phase == 10 || phase == 11) && codim == 1)
            {
                const int dir = loc / 2;
                const double* q = a.abg + 3 * (long)k * a.ncoef;
                if (!a.side)
                {
                    if (phase == 10)
                        for (int c = 0; c < a.depth; ++c)
                        {
                            bd_cc_op1(P, 0, c, loc, P.ihi, lo, hi, q[3 * c], q[3 * c + 1], q[3 * c + 2], a.adjoint,
                                      a.dx);
                            __syncthreads();
                        }
                }
                else if (phase == 10)
                {
                    for (int ax = 0; ax < a.ndim; ++ax)
                    {
                        if (ax == dir) continue;
                        int shi[3] = {P.ihi[0], P.ihi[1], P.ihi[2]}, bhi[3] = {hi[0], hi[1], hi[2]};
                        shi[ax] += 1;
                        bhi[ax] += 1;
                        bd_cc_op1(P, ax, 0, loc, shi, lo, bhi, q[3 * ax], q[3 * ax + 1], q[3 * ax + 2], a.adjoint,
                                  a.dx);
                        __syncthreads();
                    }
                }
                else
                {
                    bd_sc_op1(P, loc, lo, hi, q[3 * dir], q[3 * dir + 1], q[3 * dir + 2], a.adjoint, a.dx);
                    __syncthreads();
                }
                did = true;
            }
            (void)did;
        }
    }
}
} // namespace

// IBMethod::forwardEulerStep position update (IBMethod.cpp:713-739, a VecWAXPY):
// X_new = X + dt U, elementwise over n values; U in FP64 or FP32 (FP32C plans).
template <typename TU>
__global__ void k_euler(double* __restrict__ Xn, const double* __restrict__ X, const TU* __restrict__ U, double dt,
                        long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Xn[i] = X[i] + dt * (double)U[i];
}

extern "C" {

const char* ibamr_last_error(void) { return ibamr::g_err.c_str(); }

int ibamr_is_known_kernel(const char* k) { return find_kernel(k) ? 1 : 0; }

// USER_DEFINED reports the stencil size of the user's kernel in the reference
// (LEInteractor.cpp:2100-2101); no user kernel is registered here, so it reports 0.
int ibamr_get_stencil_size(const char* k)
{
    const KernelName* e = find_kernel(k);
    if (!e) return fail(IBAMR_ERR_BAD_KERNEL, std::string("LEInteractor::getStencilSize()\n  Unknown kernel function ") + (k ? k : "(null)"));
    return e->stencil;
}

int ibamr_get_minimum_ghost_width(const char* k)
{
    const KernelName* e = find_kernel(k);
    if (!e) return fail(IBAMR_ERR_BAD_KERNEL, std::string("LEInteractor::getStencilSize()\n  Unknown kernel function ") + (k ? k : "(null)"));
    return e->stencil / 2 + 1;
}

int ibamr_plan_destroy(ibamr_plan* P)
{
    if (!P) return IBAMR_OK;
    for (auto& v : P->tev)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    dfree(P->patches_d);
    dfree(P->tiles_d);
    dfree(P->bin_patch_start_d);
    dfree(P->fill_segs_d);
    dfree(P->bin_patches_d);
    for (int s = 0; s < MAX_SLOTS; ++s) dfree(P->arrays_d[s]);
    dfree(P->keys_d);
    dfree(P->keys_sorted_d);
    dfree(P->idx_d);
    dfree(P->perm_d);
    dfree(P->cell_d);
    dfree(P->cell_sorted_d);
    dfree(P->Xs_d);
    dfree(P->Fs_d);
    dfree(P->wt_d);
    dfree(P->st_d);
    dfree(P->bin_start_d);
    dfree(P->keys2_d);
    dfree(P->pcell_off_d);
    dfree(P->bdry_patch_d);
    dfree(P->bdry_box_d);
    dfree(P->bdry_abg_d);
    dfree(P->status_d);
    dfree(P->rs_hist_d);
    dfree(P->rs_off_d);
    dfree(P->scan_tmp_d);
    dfree(P->scan_part_d);
    delete P;
    return IBAMR_OK;
}

// Single-patch selection (the raw-array LEInteractor form): getCellIndex on the
// patch box and x range, points kept when their cell lies in sel_lo..sel_hi.
struct PatchSelect
{
    int ref_lo[3], ref_hi[3];
    int sel_lo[3], sel_hi[3];
};

static int plan_create_impl(ibamr_plan** out, const ibamr_level_geom* L, const ibamr_plan_config* C,
                            const PatchSelect* ps)
{
    if (!out || !L || !C) return fail(IBAMR_ERR_INVALID, "null argument");
    *out = nullptr;
    const int k = kernel_id(C->kernel);
    if (k == -1) return fail(IBAMR_ERR_BAD_KERNEL, std::string("Unknown kernel function ") + (C->kernel ? C->kernel : "(null)"));
    if (k < 0)
        return fail(IBAMR_ERR_BAD_KERNEL, std::string("kernel function ") + C->kernel +
                                              " has no GPU spread/interpolation (PIECEWISE_LINEAR, IB_4, IB_6 and "
                                              "BSPLINE_3 do)");
    if (L->ndim != 2 && L->ndim != 3) return fail(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (C->centering != IBAMR_CELL && C->centering != IBAMR_SIDE) return fail(IBAMR_ERR_INVALID, "bad centering");
    if (C->precision != IBAMR_FP64 && C->precision != IBAMR_FP32C) return fail(IBAMR_ERR_INVALID, "bad precision");
    if (C->centering == IBAMR_SIDE && C->depth != 1)
        return fail(IBAMR_ERR_DEPTH, "side-centered spreading/interpolation requires vector-valued data (q depth 1, Q depth NDIM)");
    if (C->depth < 1) return fail(IBAMR_ERR_DEPTH, "depth must be >= 1");
    if (L->npatches < 1 || !L->patches) return fail(IBAMR_ERR_INVALID, "level has no patches");
    if (C->max_points < 0) return fail(IBAMR_ERR_INVALID, "max_points < 0");
    const int ndim = L->ndim;
    const int mg = min_ghost(k);
    int gmin = 1 << 30;
    for (int d = 0; d < ndim; ++d) gmin = std::min(gmin, C->data_gcw[d]);
    if (gmin < mg)
        return fail(IBAMR_ERR_GHOST_WIDTH, "insufficient ghost cells: kernel " + std::string(C->kernel) + " needs " +
                                               std::to_string(mg) + ", data ghost width " + std::to_string(gmin));
    int igcw = C->idx_gcw > 0 ? C->idx_gcw : mg;
    if (igcw < mg) return fail(IBAMR_ERR_GHOST_WIDTH, "Lagrangian index ghost width below the kernel minimum");
    if (ps)  // the spread must take every selected point
        for (int d = 0; d < ndim; ++d)
            igcw = std::max(igcw, std::max(ps->ref_lo[d] - ps->sel_lo[d], ps->sel_hi[d] - ps->ref_hi[d]));
    for (int d = 0; d < ndim; ++d)
    {
        if (!(L->dx[d] > 0.0)) return fail(IBAMR_ERR_INVALID, "dx must be positive");
        if (L->domain_upper[d] < L->domain_lower[d]) return fail(IBAMR_ERR_INVALID, "empty domain box");
        if (L->periodic_shift[d] < 0) return fail(IBAMR_ERR_INVALID, "negative periodic shift");
        if (L->periodic_shift[d] > 0 && L->periodic_shift[d] < igcw + C->drift_cells)
            return fail(IBAMR_ERR_INVALID, "periodic extent smaller than the ghost width");
    }

    ibamr_plan* P = new ibamr_plan();
    P->ndim = ndim;
    P->kernel = k;
    P->centering = C->centering;
    P->precision = C->precision;
    P->depth = C->depth;
    P->narr = C->centering == IBAMR_SIDE ? ndim : 1;
    P->fdepth = C->centering == IBAMR_SIDE ? ndim : C->depth;
    P->idx_gcw = igcw;
    P->drift = std::max(0, C->drift_cells);
    P->reach = mg + P->drift;
    P->tile = ndim == 3 ? 8 : 16;
    if (C->tile != 0 && C->tile != P->tile)
    {
        delete P;
        return fail(IBAMR_ERR_INVALID, "unsupported tile edge");
    }
    P->bin = P->tile;
    P->tsz[0] = P->tile;
    P->tsz[1] = P->tile;
    P->tsz[2] = ndim == 3 ? (C->precision == IBAMR_FP64 ? S3_ZC_FP64 : S3_ZC_FP32) : 1;
    // per-axis (bin, image) lists of a tile hold at most 16 entries: the image
    // ranges partition the reachable cells, plus one split per domain side
    for (int d = 0; d < ndim; ++d)
        if ((P->tsz[d] + 2 * P->reach + P->bin - 1) / P->bin + 3 > 16)
        {
            delete P;
            return fail(IBAMR_ERR_INVALID, "drift_cells too large for the tile size");
        }
    for (int d = 0; d < 3; ++d)
    {
        const bool on = d < ndim;
        P->data_gcw[d] = on ? C->data_gcw[d] : 0;
        P->dx[d] = on ? L->dx[d] : 1.0;
        P->dom_lo[d] = on ? L->domain_lower[d] : 0;
        P->dom_hi[d] = on ? L->domain_upper[d] : 0;
        P->dom_xlo[d] = on ? L->domain_x_lower[d] : 0.0;
        P->dom_xhi[d] = on ? L->domain_x_upper[d] : 1.0;
        P->pshift[d] = on ? L->periodic_shift[d] : 0;
        P->nb[d] = on ? (P->dom_hi[d] - P->dom_lo[d] + P->bin) / P->bin : 1;
        P->ref_lo[d] = ps && on ? ps->ref_lo[d] : P->dom_lo[d];
        P->ref_hi[d] = ps && on ? ps->ref_hi[d] : P->dom_hi[d];
        P->sel_lo[d] = ps && on ? ps->sel_lo[d] : P->dom_lo[d];
        P->sel_hi[d] = ps && on ? ps->sel_hi[d] : P->dom_hi[d];
    }
    const long nbins_l = (long)P->nb[0] * P->nb[1] * P->nb[2];
    if (nbins_l > (1L << 26))
    {
        delete P;
        return fail(IBAMR_ERR_INVALID, "level too large for 32-bit bin keys");
    }
    P->nbins = (int)nbins_l;
    for (int d = 0; d < 3; ++d) P->nsb[d] = (P->nb[d] + 1) / 2;
    P->nmem = ndim == 3 ? P->nsb[0] * P->nsb[1] * P->nsb[2] * 8 : P->nbins;
    {
        // keys carry the cell's plane inside the bin when they fit in 31 bits
        int lb = 0;
        while ((1 << lb) < P->bin) ++lb;
        const int sub = lb;
        P->ksub = ((long long)(P->nmem + 1) << sub) < (1LL << 31) && (1 << lb) == P->bin ? sub : 0;
    }
    P->npatches = L->npatches;
    P->max_points = C->max_points;

    // patches, and the patches overlapping each bin (for point -> patch lookup in interp)
    std::vector<std::vector<int>> per_bin(P->nmem);
    for (int i = 0; i < L->npatches; ++i)
    {
        const ibamr_patch_geom& g = L->patches[i];
        PatchDev pd{};
        for (int d = 0; d < 3; ++d)
        {
            const bool on = d < ndim;
            pd.ilo[d] = on ? g.ilower[d] : 0;
            pd.ihi[d] = on ? g.iupper[d] : 0;
            pd.xlo[d] = on ? g.x_lower[d] : 0.0;
            pd.slo[d] = ps && on ? ps->sel_lo[d] : pd.ilo[d];
            pd.shi[d] = ps && on ? ps->sel_hi[d] : pd.ihi[d];
            if (on && (g.iupper[d] < g.ilower[d] || g.ilower[d] < P->dom_lo[d] || g.iupper[d] > P->dom_hi[d]))
            {
                delete P;
                return fail(IBAMR_ERR_INVALID, "patch box empty or outside the domain");
            }
        }
        P->patches_h.push_back(pd);
        int bl[3] = {0, 0, 0}, bh[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d)
        {
            bl[d] = (std::max(pd.slo[d], P->dom_lo[d]) - P->dom_lo[d]) / P->bin;
            bh[d] = (std::min(pd.shi[d], P->dom_hi[d]) - P->dom_lo[d]) / P->bin;
        }
        for (int bz = bl[2]; bz <= bh[2]; ++bz)
            for (int by = bl[1]; by <= bh[1]; ++by)
                for (int bx = bl[0]; bx <= bh[0]; ++bx) per_bin[bin_mem(ndim, bx, by, bz, P->nb, P->nsb)].push_back(i);
    }
    std::vector<int> bps(P->nmem + 2, 0), bpl;
    for (int b = 0; b < P->nmem; ++b)
    {
        bps[b] = (int)bpl.size();
        for (int i : per_bin[b]) bpl.push_back(i);
    }
    bps[P->nmem] = (int)bpl.size();
    bps[P->nmem + 1] = (int)bpl.size();
    if (bpl.empty()) bpl.push_back(0);


    // segments of the same-level ghost fill: every element of every patch array
    std::vector<int4> fsegs;
    for (int i = 0; i < L->npatches; ++i)
    {
        const PatchDev& pd = P->patches_h[i];
        for (int c = 0; c < P->narr; ++c)
        {
            const int axis = C->centering == IBAMR_SIDE ? c : -1;
            long total = 1;
            for (int d = 0; d < ndim; ++d) total *= pd.ihi[d] - pd.ilo[d] + 1 + 2 * P->data_gcw[d] + (d == axis ? 1 : 0);
            for (long st = 0; st < total; st += 256)
                fsegs.push_back(make_int4(i, c, (int)st, (int)std::min<long>(256, total - st)));
        }
    }
    P->nfill_segs = (int)fsegs.size();

    auto cleanup_fail = [&](int code) {
        std::string m = g_err;
        ibamr_plan_destroy(P);
        g_err = m;
        return code;
    };
#define PLAN_HIP(expr)                                                                                  \
    do                                                                                                  \
    {                                                                                                   \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
        {                                                                                               \
            fail(IBAMR_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_));                    \
            return cleanup_fail(IBAMR_ERR_DEVICE);                                                      \
        }                                                                                               \
    } while (0)

    PLAN_HIP(hipMalloc(&P->patches_d, sizeof(PatchDev) * P->patches_h.size()));
    PLAN_HIP(hipMemcpy(P->patches_d, P->patches_h.data(), sizeof(PatchDev) * P->patches_h.size(), hipMemcpyHostToDevice));
    {
        const int rc = build_tiles(P);
        if (rc != IBAMR_OK) return cleanup_fail(rc);
    }
    PLAN_HIP(hipMalloc(&P->bin_patch_start_d, sizeof(int) * bps.size()));
    PLAN_HIP(hipMemcpy(P->bin_patch_start_d, bps.data(), sizeof(int) * bps.size(), hipMemcpyHostToDevice));
    PLAN_HIP(hipMalloc(&P->fill_segs_d, sizeof(int4) * std::max<size_t>(1, fsegs.size())));
    if (!fsegs.empty())
        PLAN_HIP(hipMemcpy(P->fill_segs_d, fsegs.data(), sizeof(int4) * fsegs.size(), hipMemcpyHostToDevice));
    PLAN_HIP(hipMalloc(&P->bin_patches_d, sizeof(int) * bpl.size()));
    PLAN_HIP(hipMemcpy(P->bin_patches_d, bpl.data(), sizeof(int) * bpl.size(), hipMemcpyHostToDevice));
    const size_t np = (size_t)std::max(1, C->max_points);
    PLAN_HIP(hipMalloc(&P->keys_d, sizeof(int) * np));
    PLAN_HIP(hipMalloc(&P->keys_sorted_d, sizeof(int) * np));
    PLAN_HIP(hipMalloc(&P->idx_d, sizeof(int) * np));
    PLAN_HIP(hipMalloc(&P->perm_d, sizeof(int) * np));
    PLAN_HIP(hipMalloc(&P->cell_d, sizeof(int4) * np));
    PLAN_HIP(hipMalloc(&P->cell_sorted_d, sizeof(int4) * np));
    PLAN_HIP(hipMalloc(&P->Xs_d, sizeof(double) * ndim * np));
    PLAN_HIP(hipMalloc(&P->Fs_d, (C->precision == IBAMR_FP64 ? 8 : 4) * (size_t)P->fdepth * np));
    if (ndim == 3 && C->precision != IBAMR_FP64)
    {
        P->nvar = C->centering == IBAMR_SIDE ? 2 : 1;
        const size_t rows = (size_t)3 * P->nvar * np;
        PLAN_HIP(hipMalloc(&P->wt_d, sizeof(float) * kernel_width(k) * rows));
        PLAN_HIP(hipMalloc(&P->st_d, sizeof(int) * rows));
    }
    PLAN_HIP(hipMalloc(&P->bin_start_d, sizeof(int) * (P->nmem + 2)));
    {
        const size_t nrs = (size_t)RS_DIGITS * ((np + RS_TILE - 1) / RS_TILE);
        PLAN_HIP(hipMalloc(&P->keys2_d, sizeof(int) * np));
        PLAN_HIP(hipMalloc(&P->status_d, sizeof(int) * 8));
        std::vector<long long> pco(P->npatches + 1, 0);
        for (int i = 0; i < P->npatches; ++i)
        {
            long long c = 1;
            for (int d = 0; d < ndim; ++d) c *= (long long)(P->patches_h[i].ihi[d] - P->patches_h[i].ilo[d] + 1);
            pco[i + 1] = pco[i] + c;
        }
        if (!ps && pco[P->npatches] < (1LL << 31))
        {
            std::vector<int> pci(pco.begin(), pco.end());
            PLAN_HIP(hipMalloc(&P->pcell_off_d, sizeof(int) * pci.size()));
            PLAN_HIP(hipMemcpy(P->pcell_off_d, pci.data(), sizeof(int) * pci.size(), hipMemcpyHostToDevice));
        }
        PLAN_HIP(hipMalloc(&P->rs_hist_d, sizeof(int) * nrs));
        PLAN_HIP(hipMalloc(&P->rs_off_d, sizeof(int) * (nrs + 1)));
        PLAN_HIP(hipMalloc(&P->scan_tmp_d, sizeof(int) * ((nrs + SCAN_B - 1) / SCAN_B + 1)));
        PLAN_HIP(hipMalloc(&P->scan_part_d, sizeof(int) * nrs));
    }
#undef PLAN_HIP
    if (ndim == 3)
    {
        // raise the dynamic LDS limit of this plan's interpolation kernel here, not
        // per launch: the launches may be captured into a HIP graph
        const int rc = dispatch(3, k, C->precision, [&](auto, auto ker, auto prec) -> int {
            using T = decltype(prec);
            constexpr int KER = decltype(ker)::value;
            // the variant reading the set-up pass's stencils exists for FP32C only
            const void* fs[2] = {reinterpret_cast<const void*>(&k_interp3<KER, T, false>), nullptr};
            if constexpr (sizeof(T) == 4) fs[1] = reinterpret_cast<const void*>(&k_interp3<KER, T, true>);
            for (const void* f : fs)
                if (f)
                    HIP_OK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               interp3_lds_elems<KER, T>() * (int)sizeof(T)));
            return IBAMR_OK;
        });
        if (rc != IBAMR_OK) return cleanup_fail(rc);
    }
    *out = P;
    return IBAMR_OK;
}

int ibamr_plan_create(ibamr_plan** out, const ibamr_level_geom* L, const ibamr_plan_config* C)
{
    return plan_create_impl(out, L, C, nullptr);
}

int ibamr_plan_create_patch(ibamr_plan** out, int ndim, const double* dx, const ibamr_patch_geom* patch,
                            const int* box_lower, const int* box_upper, const ibamr_plan_config* C)
{
    if (!out || !dx || !patch || !box_lower || !box_upper || !C) return fail(IBAMR_ERR_INVALID, "null argument");
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return fail(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    PatchSelect ps{};
    ibamr_level_geom L{};
    L.ndim = ndim;
    L.npatches = 1;
    L.patches = patch;
    for (int d = 0; d < ndim; ++d)
    {
        if (box_upper[d] < box_lower[d]) return fail(IBAMR_ERR_INVALID, "empty box");
        ps.ref_lo[d] = patch->ilower[d];
        ps.ref_hi[d] = patch->iupper[d];
        ps.sel_lo[d] = box_lower[d];
        ps.sel_hi[d] = box_upper[d];
        L.dx[d] = dx[d];
        L.domain_lower[d] = std::min(box_lower[d], patch->ilower[d]);
        L.domain_upper[d] = std::max(box_upper[d], patch->iupper[d]);
        L.domain_x_lower[d] = patch->x_lower[d];
        L.domain_x_upper[d] = patch->x_upper[d];
        L.periodic_shift[d] = 0;
    }
    return plan_create_impl(out, &L, C, &ps);
}

int ibamr_plan_info(const ibamr_plan* P, int* ntiles, int* tile, int* nbins)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (ntiles) *ntiles = P->ntiles;
    if (tile) *tile = P->tile;
    if (nbins) *nbins = P->nbins;
    return IBAMR_OK;
}

int ibamr_plan_bind_data(ibamr_plan* P, int slot, void* const* arrays)
{
    if (!P || !arrays) return fail(IBAMR_ERR_INVALID, "null argument");
    if (slot < 0 || slot >= MAX_SLOTS) return fail(IBAMR_ERR_INVALID, "slot out of range");
    const size_t n = (size_t)P->npatches * P->narr;
    for (size_t i = 0; i < n; ++i)
        if (!arrays[i]) return fail(IBAMR_ERR_INVALID, "null array pointer");
    if (!P->arrays_d[slot]) HIP_OK(hipMalloc(&P->arrays_d[slot], sizeof(void*) * n));
    HIP_OK(hipMemcpy(P->arrays_d[slot], arrays, sizeof(void*) * n, hipMemcpyHostToDevice));
    P->arrays_h[slot].assign(arrays, arrays + n);
    return IBAMR_OK;
}

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
        // a fill kernel, not a memset: safe to capture in a HIP graph
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

int ibamr_plan_bin(ibamr_plan* P, const double* X, int npoints, void* stream)
{
    const int rc = ibamr_plan_bin_async(P, X, npoints, stream);
    if (rc != IBAMR_OK) return rc;
    return ibamr_plan_check(P, stream);
}

// Records the start event of launch `which` when timing is on; *end receives
// the event to record after the launch (nullptr when timing is off).
static int time_begin(ibamr_plan* P, int which, hipStream_t s, hipEvent_t* end)
{
    *end = nullptr;
    if (!P->timing) return IBAMR_OK;
    std::vector<hipEvent_t>& v = P->tev[which];
    const int i = P->tcount[which];
    if ((int)v.size() < 2 * (i + 1))
    {
        hipEvent_t e0, e1;
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        v.push_back(e0);
        v.push_back(e1);
    }
    HIP_OK(hipEventRecord(v[2 * i], s));
    *end = v[2 * i + 1];
    P->tcount[which] = i + 1;
    return IBAMR_OK;
}

int ibamr_plan_set_timing(ibamr_plan* P, int on)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    P->timing = on != 0;
    P->tcount[0] = P->tcount[1] = 0;
    return IBAMR_OK;
}

int ibamr_plan_kernel_times(ibamr_plan* P, double* spread_ms, int* nspread, double* interp_ms, int* ninterp)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    double mean[2] = {0.0, 0.0};
    for (int w = 0; w < 2; ++w)
    {
        double sum = 0.0;
        for (int i = 0; i < P->tcount[w]; ++i)
        {
            HIP_OK(hipEventSynchronize(P->tev[w][2 * i + 1]));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, P->tev[w][2 * i], P->tev[w][2 * i + 1]));
            sum += ms;
        }
        mean[w] = P->tcount[w] ? sum / P->tcount[w] : 0.0;
    }
    if (spread_ms) *spread_ms = mean[0];
    if (nspread) *nspread = P->tcount[0];
    if (interp_ms) *interp_ms = mean[1];
    if (ninterp) *ninterp = P->tcount[1];
    P->tcount[0] = P->tcount[1] = 0;
    return IBAMR_OK;
}

int ibamr_plan_spread(ibamr_plan* P, int slot, const void* F, const double* X, int mode, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE && mode != IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        return fail(IBAMR_ERR_INVALID, "bad mode");
    if (P->npoints > 0 && (!F || !X)) return fail(IBAMR_ERR_INVALID, "null F or X");
    if (P->ntiles == 0) return IBAMR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    return dispatch(P->ndim, P->kernel, P->precision, [&](auto nd, auto ker, auto prec) -> int {
        using T = decltype(prec);
        constexpr int NDIM = decltype(nd)::value;
        constexpr int KER = decltype(ker)::value;
        SpreadArgs<T> a{};
        fill_spread_args<T>(P, slot, F, X, mode, a);
        if (P->npoints > 0)
        {
            const int B = 256;
            // FP32C 3-D: stencils set up here (not for kernels whose side depends on the unshifted X)
            constexpr bool WTS = NDIM == 3 && sizeof(T) == 4 && !side_by_unshifted_x<KER>();
            if (WTS && P->fdepth == NDIM)
            {
                // gather and stencils in one pass
                const int rc = launch_point_weights<KER>(P, s, X, reinterpret_cast<const float*>(F));
                if (rc != IBAMR_OK) return rc;
            }
            else
            {
                if (P->fdepth == NDIM)
                    hipLaunchKernelGGL((k_gather_points<NDIM, T>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s,
                                       P->perm_d, P->npoints, X, P->Xs_d, static_cast<const T*>(F),
                                       static_cast<T*>(P->Fs_d));
                else
                    hipLaunchKernelGGL(k_scatter_points<T>, dim3((P->npoints + B - 1) / B), dim3(B), 0, s,
                                       P->idx_d, P->npoints, NDIM, X, P->Xs_d, P->fdepth, static_cast<const T*>(F),
                                       static_cast<T*>(P->Fs_d));
                HIP_OK(hipGetLastError());
                if (WTS)
                {
                    const int rc = launch_point_weights<KER>(P, s);
                    if (rc != IBAMR_OK) return rc;
                }
            }
            P->xs_valid = true;
            P->ws_valid = WTS;
        }
        a.X = P->Xs_d;
        a.F = static_cast<const T*>(P->Fs_d);
        if (!a.tiles || !a.patches || !a.arrays || !a.X || !a.F || !a.cells || !a.bin_start)
            return fail(IBAMR_ERR_INVALID, "internal: incomplete spread arguments");
        hipEvent_t t_end;
        const int trc = time_begin(P, 0, s, &t_end);
        if (trc != IBAMR_OK) return trc;
        if (NDIM == 3)
        {
            // FP32C packs chunks across bins; FP64 cuts them per bin (the oracle's fixed order)
            constexpr bool MERGE = sizeof(T) == 4;
            switch (P->tsz[2])
            {
            case 8: hipLaunchKernelGGL((k_spread3<KER, T, MERGE, 8>), dim3(P->ntiles), dim3(64), 0, s, a); break;
            case 16: hipLaunchKernelGGL((k_spread3<KER, T, MERGE, 16>), dim3(P->ntiles), dim3(64), 0, s, a); break;
            case 24: hipLaunchKernelGGL((k_spread3<KER, T, MERGE, 24>), dim3(P->ntiles), dim3(64), 0, s, a); break;
            case 32: hipLaunchKernelGGL((k_spread3<KER, T, MERGE, 32>), dim3(P->ntiles), dim3(64), 0, s, a); break;
            default: return fail(IBAMR_ERR_INVALID, "internal: tile height without a kernel");
            }
        }
        else
            hipLaunchKernelGGL((k_spread_tiles<2, KER, T, 16>), dim3(P->ntiles), dim3(WAVE), 0, s, a);
        HIP_OK(hipGetLastError());
        if (t_end) HIP_OK(hipEventRecord(t_end, s));
        return IBAMR_OK;
    });
}

int ibamr_plan_set_tile_height(ibamr_plan* P, int height)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (P->ndim != 3) return height == P->tsz[2] ? IBAMR_OK : fail(IBAMR_ERR_INVALID, "2-D plans have one tile shape");
    const std::vector<int>& hs = tile_heights(P->precision);
    if (std::find(hs.begin(), hs.end(), height) == hs.end())
        return fail(IBAMR_ERR_INVALID, "tile height " + std::to_string(height) + " not built for this precision");
    if ((height + 2 * P->reach + P->bin - 1) / P->bin + 3 > 16)
        return fail(IBAMR_ERR_INVALID, "drift_cells too large for this tile height");
    if (height == P->tsz[2]) return IBAMR_OK;
    const int old = P->tsz[2];
    P->tsz[2] = height;
    const int rc = build_tiles(P);
    if (rc != IBAMR_OK)
    {
        const std::string m = g_err;
        P->tsz[2] = old;
        (void)build_tiles(P);
        g_err = m;
        return rc;
    }
    return IBAMR_OK;
}

int ibamr_plan_get_tile_height(const ibamr_plan* P, int* height)
{
    if (!P || !height) return fail(IBAMR_ERR_INVALID, "null argument");
    *height = P->tsz[2];
    return IBAMR_OK;
}

int ibamr_plan_autotune(ibamr_plan* P, int slot, const void* F, const double* X, int reps, int* heights, float* ms,
                        int* ncand, void* stream)
{
    if (!P || !heights || !ms || !ncand) return fail(IBAMR_ERR_INVALID, "null argument");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    *ncand = 0;
    if (P->ndim != 3)
    {
        heights[0] = P->tsz[2];
        ms[0] = 0.0f;
        *ncand = 1;
        return IBAMR_OK;
    }
    reps = std::max(1, reps);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const std::vector<int> hs = tile_heights(P->precision);
    if ((int)hs.size() > IBAMR_MAX_TILE_HEIGHTS) return fail(IBAMR_ERR_INVALID, "internal: too many tile heights");
    const bool timing = P->timing;
    P->timing = false;  // the trials are not the caller's launches
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = IBAMR_OK, best = P->tsz[2];
    float best_ms = std::numeric_limits<float>::infinity();
    auto run = [&]() -> int {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        for (size_t i = 0; i < hs.size(); ++i)
        {
            heights[i] = hs[i];
            ms[i] = -1.0f;  // not admissible (drift too large for this height)
            if ((hs[i] + 2 * P->reach + P->bin - 1) / P->bin + 3 > 16) continue;
            int r = ibamr_plan_set_tile_height(P, hs[i]);
            if (r != IBAMR_OK) return r;
            r = ibamr_plan_spread(P, slot, F, X, IBAMR_SPREAD_OVERWRITE, stream);
            if (r != IBAMR_OK) return r;
            HIP_OK(hipEventRecord(e0, s));
            for (int k = 0; k < reps; ++k)
            {
                r = ibamr_plan_spread(P, slot, F, X, IBAMR_SPREAD_OVERWRITE, stream);
                if (r != IBAMR_OK) return r;
            }
            HIP_OK(hipEventRecord(e1, s));
            HIP_OK(hipEventSynchronize(e1));
            float t = 0.0f;
            HIP_OK(hipEventElapsedTime(&t, e0, e1));
            ms[i] = t / reps;
            if (ms[i] < best_ms)
            {
                best_ms = ms[i];
                best = hs[i];
            }
        }
        return IBAMR_OK;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    P->timing = timing;
    if (rc != IBAMR_OK) return rc;
    *ncand = (int)hs.size();
    return ibamr_plan_set_tile_height(P, best);
}

int ibamr_plan_interp(ibamr_plan* P, int slot, void* U, const double* X, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    if (P->npoints == 0) return IBAMR_OK;
    if (!U) return fail(IBAMR_ERR_INVALID, "null U");
    if (!X && !P->xs_valid)
        return fail(IBAMR_ERR_INVALID, "X is null and no spread/interp since the last binning permuted positions");
    hipStream_t s = static_cast<hipStream_t>(stream);
    return dispatch(P->ndim, P->kernel, P->precision, [&](auto nd, auto ker, auto prec) -> int {
        using T = decltype(prec);
        constexpr int NDIM = decltype(nd)::value;
        constexpr int KER = decltype(ker)::value;
        InterpArgs<T> a{};
        a.patches = P->patches_d;
        a.npatches = P->npatches;
        a.bin_patch_start = P->bin_patch_start_d;
        a.bin_patches = P->bin_patches_d;
        a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
        a.perm = P->perm_d;
        if (X)  // else: reuse the positions the last spread/interp put in bin order
        {
            const int B = 256;
            hipLaunchKernelGGL((k_gather_points<NDIM, T>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->perm_d,
                               P->npoints, X, P->Xs_d, static_cast<const T*>(nullptr), static_cast<T*>(nullptr));
            HIP_OK(hipGetLastError());
            P->xs_valid = true;
            P->ws_valid = false;
        }
        // FP32C: the spread set-up's stencils of these positions when it computed
        // them, else the kernel evaluates the same ones (level frame)
        a.wt = P->ws_valid ? P->wt_d : nullptr;
        a.st = P->ws_valid ? P->st_d : nullptr;
        a.nvar = P->nvar;
        for (int d = 0; d < 3; ++d)
        {
            a.lev_xlo[d] = P->dom_xlo[d];
            a.lev_lo[d] = P->dom_lo[d];
        }
        a.X = P->Xs_d;
        a.U = static_cast<T*>(U);
        a.cells = P->cell_sorted_d;
        a.keys_sorted = P->keys_sorted_d;
        a.bin_start = P->bin_start_d;
        a.ksub = P->ksub;
        a.n = P->npoints;
        for (int d = 0; d < 3; ++d)
        {
            a.dx[d] = P->dx[d];
            a.gcw[d] = P->data_gcw[d];
        }
        a.narr = P->narr;
        a.fdepth = P->fdepth;
        a.depth = P->depth;
        a.side = P->centering == IBAMR_SIDE;
        if (!a.patches || !a.bin_patch_start || !a.bin_patches || !a.arrays || !a.X || !a.U || !a.perm || !a.cells ||
            !a.keys_sorted || !a.bin_start)
            return fail(IBAMR_ERR_INVALID, "internal: incomplete interpolation arguments");
        hipEvent_t t_end = nullptr;
        if (NDIM == 3)
        {
            // one workgroup per superbin; LDS holds one staged block of the array
            // (the kernel's dynamic LDS limit is raised once, at plan creation)
            const int elems = interp3_lds_elems<KER, T>();
            const int bytes = elems * (int)sizeof(T);
            const int trc = time_begin(P, 1, s, &t_end);
            if (trc != IBAMR_OK) return trc;
            const int nsb = P->nsb[0] * P->nsb[1] * P->nsb[2];
            const int nrow = P->nsb[1] * P->nsb[2];
            const int per = (nrow + 7) / 8 * P->nsb[0];  // superbins per XCD (see k_interp3)
            if constexpr (sizeof(T) == 4)
            {
                if (a.st)  // the spread set-up's stencils of these positions
                {
                    hipLaunchKernelGGL((k_interp3<KER, T, true>), dim3(8 * per), dim3(I3_THREADS), bytes, s, a,
                                       elems, nsb, P->nsb[0]);
                    HIP_OK(hipGetLastError());
                    if (t_end) HIP_OK(hipEventRecord(t_end, s));
                    return IBAMR_OK;
                }
            }
            hipLaunchKernelGGL((k_interp3<KER, T, false>), dim3(8 * per), dim3(I3_THREADS), bytes, s, a, elems, nsb,
                               P->nsb[0]);
        }
        else
        {
            const int trc = time_begin(P, 1, s, &t_end);
            if (trc != IBAMR_OK) return trc;
            const int B = 256;
            hipLaunchKernelGGL((k_interp<2, KER, T>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s, a);
        }
        HIP_OK(hipGetLastError());
        if (t_end) HIP_OK(hipEventRecord(t_end, s));
        return IBAMR_OK;
    });
}

// IBMethod::forwardEulerStep (IBMethod.cpp:713-739)
int ibamr_forward_euler_step(double* X_new, const double* X, const void* U, int u_precision, double dt, long n,
                             void* stream)
{
    if (n < 0) return fail(IBAMR_ERR_INVALID, "n < 0");
    if (n == 0) return IBAMR_OK;
    if (!X_new || !X || !U) return fail(IBAMR_ERR_INVALID, "null array");
    if (u_precision != IBAMR_FP64 && u_precision != IBAMR_FP32C) return fail(IBAMR_ERR_INVALID, "bad precision");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    const dim3 g((unsigned)((n + B - 1) / B));
    if (u_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_euler<double>, g, dim3(B), 0, s, X_new, X, static_cast<const double*>(U), dt, n);
    else
        hipLaunchKernelGGL(k_euler<float>, g, dim3(B), 0, s, X_new, X, static_cast<const float*>(U), dt, n);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

int ibamr_plan_fill_periodic_ghosts(ibamr_plan* P, int slot, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->npatches != 1) return fail(IBAMR_ERR_INVALID, "periodic ghost fill needs a single-patch level");
    const PatchDev& pd = P->patches_h[0];
    for (int d = 0; d < P->ndim; ++d)
    {
        if (P->pshift[d] != pd.ihi[d] - pd.ilo[d] + 1 || pd.ilo[d] != P->dom_lo[d])
            return fail(IBAMR_ERR_INVALID, "periodic ghost fill needs a fully periodic domain covered by the patch");
    }
    // host copy of the pointer table: no device-to-host copy, so this can be captured
    const std::vector<void*>& ptrs = P->arrays_h[slot];
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ncomps = P->centering == IBAMR_SIDE ? 1 : P->depth;
    for (int a = 0; a < P->narr; ++a)
    {
        const int axis = P->centering == IBAMR_SIDE ? a : -1;
        int n[3] = {1, 1, 1}, g[3] = {0, 0, 0}, N[3] = {1, 1, 1};
        for (int d = 0; d < P->ndim; ++d)
        {
            n[d] = pd.ihi[d] - pd.ilo[d] + 1 + 2 * P->data_gcw[d] + (d == axis ? 1 : 0);
            g[d] = P->data_gcw[d];
            N[d] = pd.ihi[d] - pd.ilo[d] + 1;
        }
        const long total = (long)n[0] * n[1] * n[2];
        // the core along a side axis is the N faces [g, g+N): face N is the image of face 0
        const long sz = (long)(n[2] - N[2]) * n[0] * n[1];
        const long sy = (long)N[2] * (n[1] - N[1]) * n[0];
        const long sx = (long)N[2] * N[1] * (n[0] - N[0]);
        const long shell = sz + sy + sx;
        if (shell == 0) continue;
        for (int c = 0; c < ncomps; ++c)
        {
            const int B = 256;
            const long G = (shell + B - 1) / B;
            if (P->precision == IBAMR_FP64)
                hipLaunchKernelGGL(k_fill_periodic<double>, dim3((unsigned)G), dim3(B), 0, s,
                                   static_cast<double*>(ptrs[a]) + c * total, n[0], n[1], n[2], g[0], g[1], g[2],
                                   N[0], N[1], N[2], sz, sy, sx);
            else
                hipLaunchKernelGGL(k_fill_periodic<float>, dim3((unsigned)G), dim3(B), 0, s,
                                   static_cast<float*>(ptrs[a]) + c * total, n[0], n[1], n[2], g[0], g[1], g[2],
                                   N[0], N[1], N[2], sz, sy, sx);
            HIP_OK(hipGetLastError());
        }
    }
    return IBAMR_OK;
}

int ibamr_plan_fill_ghosts(ibamr_plan* P, int slot, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->nfill_segs == 0) return IBAMR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        FillArgs<T> a{};
        a.segs = P->fill_segs_d;
        a.patches = P->patches_d;
        a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
        a.bin_patch_start = P->bin_patch_start_d;
        a.bin_patches = P->bin_patches_d;
        a.ndim = P->ndim;
        a.narr = P->narr;
        a.side = P->centering == IBAMR_SIDE;
        a.depth = P->depth;
        for (int d = 0; d < 3; ++d)
        {
            a.gcw[d] = P->data_gcw[d];
            a.dom_lo[d] = P->dom_lo[d];
            a.dom_hi[d] = P->dom_hi[d];
            a.pshift[d] = P->pshift[d];
            a.nb[d] = P->nb[d];
            a.nsb[d] = P->nsb[d];
        }
        a.bin = P->bin;
        hipLaunchKernelGGL(k_fill_level<T>, dim3((unsigned)P->nfill_segs), dim3(256), 0, s, a);
        HIP_OK(hipGetLastError());
        return IBAMR_OK;
    };
    return P->precision == IBAMR_FP64 ? go((double)0) : go((float)0);
}

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
    P->ws_valid = false;
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

int ibamr_plan_set_physical_boundary(ibamr_plan* P, const ibamr_bdry_box* boxes, int nboxes, const double* abg)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (nboxes < 0 || (nboxes > 0 && (!boxes || !abg))) return fail(IBAMR_ERR_INVALID, "null boundary boxes");
    for (int d = 1; d < P->ndim; ++d)
        if (P->data_gcw[d] != P->data_gcw[0])
            return fail(IBAMR_ERR_GHOST_WIDTH, "physical boundary operators need a uniform ghost width");
    if (P->centering == IBAMR_SIDE && P->depth != 1)
        return fail(IBAMR_ERR_INVALID, "physical boundary operators on side data of depth 1 only");
    const int g = P->data_gcw[0];
    // group by patch, keeping the given order inside each patch
    std::vector<std::vector<int>> per(P->npatches);
    for (int k = 0; k < nboxes; ++k)
    {
        const ibamr_bdry_box& b = boxes[k];
        if (b.patch < 0 || b.patch >= P->npatches)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad patch index");
        if (b.codim < 1 || b.codim > P->ndim)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad codimension");
        const int nloc = b.codim == 1 ? 2 * P->ndim : (b.codim == P->ndim ? (1 << P->ndim) : 12);
        if (b.location < 0 || b.location >= nloc)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad location index");
        // every cell an operator touches must lie in the patch's ghost box
        const PatchDev& Q = P->patches_h[b.patch];
        for (int d = 0; d < P->ndim; ++d)
        {
            if (b.lower[d] > b.upper[d] || b.lower[d] < Q.ilo[d] - g || b.upper[d] > Q.ihi[d] + g)
                return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + " is not inside the ghost box");
        }
        if (b.codim == 1)
        {
            const int dir = b.location / 2;
            const int want = (b.location % 2) ? Q.ihi[dir] + 1 : Q.ilo[dir] - 1;
            if (b.lower[dir] != want || b.upper[dir] != want)
                return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": a codim-1 box is the cell "
                                               "layer just outside its face");
            for (int d = 0; d < P->ndim; ++d)
                if (d != dir && (b.lower[d] < Q.ilo[d] || b.upper[d] > Q.ihi[d]))
                    return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": a codim-1 box spans at "
                                                   "most its face");
        }
        else
        {
            // codim 2 / ndim: outside the patch, within the ghost width, on the sides the location names
            int up[3] = {-1, -1, -1};  // -1: along the box (edge direction)
            if (b.codim == P->ndim)
                for (int d = 0; d < P->ndim; ++d) up[d] = (b.location >> d) & 1;
            else
            {
                const int l = b.location;
                if (l < 4)
                    up[1] = l & 1, up[2] = (l >> 1) & 1;
                else if (l < 8)
                    up[0] = (l - 4) >> 1, up[2] = (l - 4) & 1;
                else
                    up[0] = (l - 8) & 1, up[1] = (l - 8) >> 1;
            }
            for (int d = 0; d < P->ndim; ++d)
            {
                if (up[d] < 0) continue;
                const bool ok = up[d] ? b.lower[d] > Q.ihi[d] : b.upper[d] < Q.ilo[d];
                if (!ok)
                    return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) +
                                                       ": an edge or corner box lies outside the patch on its sides");
            }
        }
        per[b.patch].push_back(k);
    }
    std::vector<int4> pl;
    std::vector<int> bx;
    std::vector<double> ab;
    const int narr = P->centering == IBAMR_SIDE ? P->ndim : P->depth;
    for (int p = 0; p < P->npatches; ++p)
    {
        if (per[p].empty()) continue;
        pl.push_back(make_int4(p, (int)(bx.size() / 8), (int)per[p].size(), 0));
        for (int k : per[p])
        {
            const ibamr_bdry_box& b = boxes[k];
            const int v[8] = {b.codim, b.location, b.lower[0], b.lower[1], P->ndim == 3 ? b.lower[2] : 0,
                              b.upper[0], b.upper[1], P->ndim == 3 ? b.upper[2] : 0};
            bx.insert(bx.end(), v, v + 8);
            ab.insert(ab.end(), abg + 3 * (long)k * narr, abg + 3 * (long)(k + 1) * narr);
        }
    }
    dfree(P->bdry_patch_d);
    dfree(P->bdry_box_d);
    dfree(P->bdry_abg_d);
    P->bdry_patch_d = nullptr;
    P->bdry_box_d = nullptr;
    P->bdry_abg_d = nullptr;
    P->nbdry_patches = (int)pl.size();
    if (pl.empty()) return IBAMR_OK;
    HIP_OK(hipMalloc(&P->bdry_patch_d, sizeof(int4) * pl.size()));
    HIP_OK(hipMalloc(&P->bdry_box_d, sizeof(int) * bx.size()));
    HIP_OK(hipMalloc(&P->bdry_abg_d, sizeof(double) * ab.size()));
    HIP_OK(hipMemcpy(P->bdry_patch_d, pl.data(), sizeof(int4) * pl.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(P->bdry_box_d, bx.data(), sizeof(int) * bx.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(P->bdry_abg_d, ab.data(), sizeof(double) * ab.size(), hipMemcpyHostToDevice));
    return IBAMR_OK;
}

static int phys_bdry_launch(ibamr_plan* P, int slot, int adjoint, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->nbdry_patches == 0) return IBAMR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        BdryArgs<T> a{};
        a.patches_list = P->bdry_patch_d;
        a.boxes = P->bdry_box_d;
        a.abg = P->bdry_abg_d;
        a.patches = P->patches_d;
        a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
        a.ndim = P->ndim;
        a.side = P->centering == IBAMR_SIDE;
        a.depth = P->depth;
        a.narr = P->narr;
        a.ncoef = P->centering == IBAMR_SIDE ? P->ndim : P->depth;
        a.gcw = P->data_gcw[0];
        a.adjoint = adjoint;
        for (int d = 0; d < 3; ++d) a.dx[d] = P->dx[d];
        hipLaunchKernelGGL(k_phys_bdry<T>, dim3((unsigned)P->nbdry_patches), dim3(256), 0, s, a);
        HIP_OK(hipGetLastError());
        return IBAMR_OK;
    };
    return P->precision == IBAMR_FP64 ? go((double)0) : go((float)0);
}

int ibamr_plan_accumulate_physical_boundary(ibamr_plan* P, int slot, void* stream)
{
    return phys_bdry_launch(P, slot, 1, stream);
}

int ibamr_plan_fill_physical_boundary(ibamr_plan* P, int slot, void* stream)
{
    return phys_bdry_launch(P, slot, 0, stream);
}

} // extern "C"
