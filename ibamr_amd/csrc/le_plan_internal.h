// le_plan_internal.h -- declarations shared by the level-plan translation units
// (le_plan.hip, le_bin.hip, le_spread.hip, le_interp.hip, le_ghost.hip,
// le_redist.hip): the plan object, patch/tile descriptors, error helpers, the
// kernel dispatcher and a few small kernels each unit instantiates for itself.
#pragma once
#include "ibamr_amd.h"
#include "le_weights.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <climits>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace ibamr
{
extern thread_local std::string g_err;  // le_plan.hip

inline int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int set_error(int code, const std::string& msg);

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

// The spread's work list for one tile shape: every tile of every array of every
// patch, in XCD-aware launch order (built at plan creation, kept until the plan
// is destroyed, so a HIP graph captured with any shape stays valid).
struct TileList
{
    TileDesc* d = nullptr;
    int n = 0;
    int shape[3] = {1, 1, 1};
};


// radix sort and scan shapes (le_bin.hip; plan buffers are sized from them)
constexpr int RS_BITS = 8;
constexpr int RS_DIGITS = 1 << RS_BITS;
constexpr int RS_THREADS = 256;            // one digit per thread in the bookkeeping
constexpr int RS_ROUNDS = 16;              // rounds of RS_THREADS elements per block
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
static_assert(RS_THREADS == RS_DIGITS, "one thread per digit");
constexpr int SCAN_B = 1024;


// IndexUtilities::getCellIndex (IndexUtilities-inl.h:52-72; the level form
// :226-243 passes the level's domain box and x range) along one axis, as a
// double: the lower side's formula when X is at least as close to x_lower as to
// x_upper, else the upper side's.  The result is an integer value (or NaN for a
// NaN position); callers check its range before converting it.
__device__ __forceinline__ double cell_index_d(double X, double x_lower, double x_upper, double dx, int ilower,
                                               int iupper)
{
    const double dXl = X - x_lower, dXu = X - x_upper;
    if (fabs(dXl) <= fabs(dXu)) return (double)ilower + floor(dXl / dx);
    return ((double)iupper + floor(dXu / dx)) + 1.0;
}

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
    int tile = 8;           // bin edge (cells)
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
    std::vector<TileList> tlists;      // spread work lists, one per admissible tile shape (first: default)
    int tcur = 0;                      // the shape in use
    // FP32C 3-D: the stencils of the bin-ordered positions, set up once per
    // spread (k_point_weights): for axis d and variant v (v = 1: x_lower - dx/2,
    // the normal axis of side data) the first stencil index
    // (st_d[(d*nvar+v)*n + k]) and the W weights (wt_d[((d*nvar+v)*n + k)*W + j])
    float* wt_d = nullptr;
    int* st_d = nullptr;
    int nvar = 1;                      // variants per axis: 2 for side data
    bool ws_valid = false;             // wt_d/st_d hold the stencils of Xs
    // 3-D interpolation: per superbin, the patch owning all its cells when every
    // staged block of it (the superbin's cells widened by sb_off, the kernel's
    // stencil start offsets per variant plus the drift) lies inside that
    // patch's ghosted arrays and fits the LDS block; else -1 (extents from the
    // points)
    int* sb_owner_d = nullptr;
    int sb_off[2][2] = {};
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
    double* bdry_coef_d = nullptr;     // Robin a, b, g over each codim-1 box's columns, per array
    long* bdry_coef_off_d = nullptr;   // [nbdry][narr] where each (box, array)'s coefficients start
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
    void* Fs_d = nullptr;              // forces in bin order
    int* bin_start_d = nullptr;        // [nmem+2]

    // optional timing of the main spread / interpolation kernels: event pairs
    // recorded on the launch stream around each launch (0: spread, 1: interp)
    bool timing = false;
    std::vector<hipEvent_t> tev[2];
    int tcount[2] = {0, 0};
};

namespace ibamr
{
namespace
{
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

// The same permutation driven by the sorted order, for records of D positions and
// D force components (side data): each thread gathers one point's record (all
// loads first) and writes it to its slot, so the writes are coalesced.  Measured
// faster than the scattered-write form for these 24 + 12-byte records.
template <int D, typename T>
__global__ void k_gather_points(const int* __restrict__ perm, int n, const double* __restrict__ X,
                                double* __restrict__ Xs, const T* __restrict__ F, T* __restrict__ Fs,
                                const T* __restrict__ ds = nullptr)
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
        if (ds)  // F_ds[k][d] = F[k][d] * ds[k]  (LDataManager.cpp:439-445)
        {
            const T w = ds[p];
#pragma unroll
            for (int d = 0; d < D; ++d) f[d] = f[d] * w;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) Fs[D * (size_t)i + d] = f[d];
    }
}
} // namespace

// Kernel of a plan as a compile-time constant: f(ndim, kernel, precision tag).
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

inline void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

inline int kernel_width(int k)
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

// le_bin.hip
int radix_sort_pairs(const ibamr_plan* P, const int* kin, const int* vin, int n, int nbits, int* kout, int* vout,
                     int* kt, int* vt, hipStream_t s);
int bits_for(int maxkey);
// le_spread.hip: the spread's work lists for every admissible tile shape of the
// plan's form
int build_tile_lists(ibamr_plan* P);
// le_interp.hip: per-plan interpolation set-up (dynamic LDS limits; superbin owners)
int interp_prepare(int ndim, int kernel, int precision);
int build_superbin_owners(ibamr_plan* P);
// le_plan.hip: timing of the main spread / interpolation launches
int time_begin(ibamr_plan* P, int which, hipStream_t s, hipEvent_t* end);
} // namespace ibamr
