// le_spread.hip -- the spread (LDataManager::spread inner loop,
// LDataManager.cpp:552-667 -> LEInteractor::spread -> lagrangian_<kernel>_spread{2,3}d):
// bin-order gather of the points, the spread kernels, the tile work list and its
// autotuner.
#include "le_plan_internal.h"

namespace ibamr
{
namespace
{
// Point data in bin order, so every later kernel reads slot k coalesced.
// Driven by the points: read X and F in the caller's order (coalesced) and
// write each point's record to its slot (a scattered write of a few bytes
// costs less than a scattered read of whole cache lines).
// ds (optional): F_ds[k][c] = F[k][c] * ds[k] (LDataManager.cpp:439-445).
template <typename T>
__global__ void k_scatter_points(const int* __restrict__ inv, int n, int ndim, const double* __restrict__ X,
                                 double* __restrict__ Xs, int fdepth, const T* __restrict__ F, T* __restrict__ Fs,
                                 const T* __restrict__ ds)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const size_t k = (size_t)inv[p];
    for (int d = 0; d < ndim; ++d) Xs[(size_t)ndim * k + d] = X[(size_t)ndim * p + d];
    if (F)
    {
        const T w = ds ? ds[p] : T(1);
        for (int c = 0; c < fdepth; ++c)
        {
            const T f = F[(size_t)fdepth * p + c];
            Fs[(size_t)fdepth * k + c] = ds ? f * w : f;
        }
    }
}

// FP32C 3-D spread set-up: for each point (bin order), axis d and variant v
// (v = 0: the level's x_lower; v = 1: x_lower - dx/2, the normal axis of side
// data, LEInteractor.cpp:3689) the first stencil index in the level's index
// space and the W weights, in the level's frame.  The spread's tiles then stage
// a candidate with loads instead of recomputing its weights for every tile it
// reaches; a patch or periodic image shifts the start by whole cells.  With
// perm, the kernel also performs the bin-order gather of X and F (side data:
// records of 3 positions and 3 force components, F times ds[p] when ds is
// given), so Xs is not read back.
template <int KER>
__global__ void k_point_weights(int n, double* __restrict__ Xs, int nvar, double3 xlo, double3 dx, int3 lo,
                                float* __restrict__ wt, int* __restrict__ st, const int* __restrict__ perm,
                                const double* __restrict__ Xin, const float* __restrict__ Fin,
                                float* __restrict__ Fs, const float* __restrict__ ds)
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
        if (ds)  // F_ds[k][d] = F[k][d] * ds[k]  (LDataManager.cpp:439-445)
        {
            const float w = ds[p];
#pragma unroll
            for (int d = 0; d < 3; ++d) f[d] = f[d] * w;
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
int launch_point_weights(ibamr_plan* P, hipStream_t s, const double* X = nullptr, const float* F = nullptr,
                         const float* ds = nullptr)
{
    if (!P->wt_d || !P->st_d) return fail(IBAMR_ERR_INVALID, "internal: no weight buffers");
    const int B = 256;
    hipLaunchKernelGGL((k_point_weights<KER>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->npoints, P->Xs_d,
                       P->nvar, make_double3(P->dom_xlo[0], P->dom_xlo[1], P->dom_xlo[2]),
                       make_double3(P->dx[0], P->dx[1], P->dx[2]),
                       make_int3(P->dom_lo[0], P->dom_lo[1], P->dom_lo[2]), P->wt_d, P->st_d,
                       X ? P->perm_d : nullptr, X, F, static_cast<float*>(P->Fs_d), ds);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
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

    // Load (accumulate: the raw LEInteractor::spread adds each contribution to
    // the existing value, f.m4:1463) or clear (overwrite; interior mode, which
    // spreads into zeroed data and adds the old values afterwards as
    // LDataManager::spread does, LDataManager.cpp:594, 661-663) the tile; LDS
    // layout is TILE-strided.
    const bool from_old = a.mode == IBAMR_SPREAD_ACCUMULATE;
    for (int e = lane; e < TN; e += WAVE)
    {
        T v = T(0);
        const int i0 = e % TILE;
        const int i1 = (e / TILE) % TILE;
        const int i2 = NDIM == 3 ? e / (TILE * TILE) : 0;
        if (from_old && i0 < fn[0] && i1 < fn[1] && i2 < fn[2])
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
        if (!overwrite && !from_old) v = arr[comp_off + g] + v;
        arr[comp_off + g] = v;
    }
}
} // namespace

#include "le_spread3d.h"

namespace
{
template <typename T>
void fill_spread_args(const ibamr_plan* P, int slot, const void* F, const double* X, int mode, SpreadArgs<T>& a)
{
    a.tiles = P->tlists[P->tcur].d;
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

// Spread tile shapes built for each form; the first is the default.
//   2-D: a 16 x 16 tile in LDS (k_spread_tiles).
//   3-D: 8 x 8 or 16 x 4 grid lines (one per lane) times a column of z values
//     in registers (k_spread3); FP32C keeps two register sets per z value (the
//     FP64 total and the chunk's FP32 partial), so its columns are shorter.
//     FP32C z extents of SWEEP_MIN_Z or more are z segments swept through a
//     register window (k_spread3s).
using Shape = std::array<int, 3>;
constexpr int SWEEP_MIN_Z = 48;
const std::vector<Shape>& tile_shapes(int ndim, int precision)
{
    static const std::vector<Shape> s2 = {{16, 16, 1}};
    static const std::vector<Shape> s64 = {{8, 8, 32}, {8, 8, 16}, {16, 4, 32}};
    static const std::vector<Shape> s32 = {{8, 8, 16}, {8, 8, 8}, {8, 8, 24}, {16, 4, 16},
                                           {8, 8, 64}, {8, 8, 128}, {8, 8, 256}};
    if (ndim == 2) return s2;
    return precision == IBAMR_FP64 ? s64 : s32;
}

// A shape is admissible when each axis of a tile reaches at most 16 (bin,
// image) entries (the image ranges partition the reachable cells, plus one
// split per domain side).  A z sweep takes up to S3S_MAXZ z entries and at
// most 64 (x, y) combinations per z layer, needs every stencil of a bin layer
// inside its register window (bin + 2 reach <= S3S_ZW) and the set-up pass's
// stencils (not computed for kernels whose stencil side depends on the
// unshifted X).
bool shape_admissible(const ibamr_plan* P, const Shape& sh)
{
    const bool sweep = P->ndim == 3 && sh[2] >= SWEEP_MIN_Z;
    if (sweep && (P->precision == IBAMR_FP64 || P->kernel == K_BS4 || P->kernel == K_BS6 ||
                  P->bin + 2 * P->reach > S3S_ZW))
        return false;
    int n[3] = {1, 1, 1};
    for (int d = 0; d < P->ndim; ++d)
    {
        n[d] = (sh[d] + 2 * P->reach + P->bin - 1) / P->bin + 3;
        if (n[d] > (sweep && d == 2 ? S3S_MAXZ : 16)) return false;
    }
    return !sweep || n[0] * n[1] <= 64;
}

// Work list of one tile shape: every tile of every array of every patch.
// XCD-aware launch order: workgroup i runs on XCD i % 8.  Tiles are grouped by
// (patch, z row, y row), each group holding that row of every array, and the
// groups are dealt to the XCDs in turn; each XCD runs its groups in (patch, z,
// y) order.  A group's tiles read the same point data through one L2 (all
// arrays, x neighbours), every XCD gets a share of every region (point clouds
// are rarely uniform, so contiguous eighths of the domain would leave some XCDs
// idle), and all XCDs sweep the domain together, so the data they re-read is
// recent in the shared last-level cache.
int build_tiles(ibamr_plan* P, const Shape& sh, TileList& out)
{
    std::vector<TileDesc> tiles;
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
                nt[c][d] = (n + sh[d] - 1) / sh[d];
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
    out.n = (int)tiles.size();
    for (int d = 0; d < 3; ++d) out.shape[d] = sh[d];
    HIP_OK(hipMalloc(&out.d, sizeof(TileDesc) * std::max<size_t>(1, tiles.size())));
    if (!tiles.empty()) HIP_OK(hipMemcpy(out.d, tiles.data(), sizeof(TileDesc) * tiles.size(), hipMemcpyHostToDevice));
    return IBAMR_OK;
}

int find_shape(const ibamr_plan* P, const int* shape)
{
    for (size_t i = 0; i < P->tlists.size(); ++i)
        if (P->tlists[i].shape[0] == shape[0] && P->tlists[i].shape[1] == shape[1] &&
            P->tlists[i].shape[2] == shape[2])
            return (int)i;
    return -1;
}
} // namespace

int build_tile_lists(ibamr_plan* P)
{
    for (const Shape& sh : tile_shapes(P->ndim, P->precision))
    {
        if (!shape_admissible(P, sh)) continue;
        P->tlists.emplace_back();
        const int rc = build_tiles(P, sh, P->tlists.back());
        if (rc != IBAMR_OK) return rc;
    }
    if (P->tlists.empty()) return fail(IBAMR_ERR_INVALID, "drift_cells too large for the tile size");
    P->tcur = 0;
    return IBAMR_OK;
}
} // namespace ibamr

extern "C" {

int ibamr_plan_spread(ibamr_plan* P, int slot, const void* F, const double* X, int mode, void* stream)
{
    return ibamr_plan_spread_ds(P, slot, F, X, nullptr, mode, stream);
}

int ibamr_plan_spread_ds(ibamr_plan* P, int slot, const void* F, const double* X, const void* ds, int mode,
                         void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE && mode != IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        return fail(IBAMR_ERR_INVALID, "bad mode");
    if (P->npoints > 0 && (!F || !X)) return fail(IBAMR_ERR_INVALID, "null F or X");
    if (P->tlists.empty() || P->tlists[P->tcur].n == 0) return IBAMR_OK;
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
            if constexpr (WTS)
            {
                if (P->fdepth == NDIM)
                {
                    // gather (with F*ds) and stencils in one pass
                    const int rc = launch_point_weights<KER>(P, s, X, reinterpret_cast<const float*>(F),
                                                             reinterpret_cast<const float*>(ds));
                    if (rc != IBAMR_OK) return rc;
                }
            }
            if (!WTS || P->fdepth != NDIM)
            {
                // the points in bin order
                if (P->fdepth == NDIM)
                    hipLaunchKernelGGL((k_gather_points<NDIM, T>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s,
                                       P->perm_d, P->npoints, X, P->Xs_d, static_cast<const T*>(F),
                                       static_cast<T*>(P->Fs_d), static_cast<const T*>(ds));
                else
                    hipLaunchKernelGGL(k_scatter_points<T>, dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->idx_d,
                                       P->npoints, NDIM, X, P->Xs_d, P->fdepth, static_cast<const T*>(F),
                                       static_cast<T*>(P->Fs_d), static_cast<const T*>(ds));
                HIP_OK(hipGetLastError());
                if constexpr (WTS)
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
        const TileList& tl = P->tlists[P->tcur];
        if constexpr (NDIM == 3)
        {
            // FP32C packs chunks across bins; FP64 cuts them per bin (the oracle's fixed order)
            constexpr bool MERGE = sizeof(T) == 4;
            const dim3 G(tl.n), B(64);
#define IBAMR_S3(TX, ZC)                                                              \
    if (tl.shape[0] == TX && tl.shape[2] == ZC)                                       \
    {                                                                                 \
        hipLaunchKernelGGL((k_spread3<KER, T, MERGE, ZC, TX>), G, B, 0, s, a);        \
        HIP_OK(hipGetLastError());                                                    \
        if (t_end) HIP_OK(hipEventRecord(t_end, s));                                  \
        return IBAMR_OK;                                                              \
    }
            if constexpr (sizeof(T) == 8)
            {
                IBAMR_S3(8, 32)
                IBAMR_S3(8, 16)
                IBAMR_S3(16, 32)
            }
            else
            {
                IBAMR_S3(8, 16)
                IBAMR_S3(8, 8)
                IBAMR_S3(8, 24)
                IBAMR_S3(16, 16)
                if constexpr (!side_by_unshifted_x<KER>())
                    if (tl.shape[0] == 8 && tl.shape[2] >= SWEEP_MIN_Z)
                    {
                        if (P->npoints > 0 && !P->ws_valid)
                            return fail(IBAMR_ERR_INVALID, "internal: z sweep without stencils");
                        hipLaunchKernelGGL((k_spread3s<KER, 8>), G, B, 0, s, a, tl.shape[2]);
                        HIP_OK(hipGetLastError());
                        if (t_end) HIP_OK(hipEventRecord(t_end, s));
                        return IBAMR_OK;
                    }
            }
#undef IBAMR_S3
            return fail(IBAMR_ERR_INVALID, "internal: tile shape without a kernel");
        }
        else
            hipLaunchKernelGGL((k_spread_tiles<2, KER, T, 16>), dim3(tl.n), dim3(WAVE), 0, s, a);
        HIP_OK(hipGetLastError());
        if (t_end) HIP_OK(hipEventRecord(t_end, s));
        return IBAMR_OK;
    });
}

int ibamr_plan_set_tile_shape(ibamr_plan* P, const int* shape)
{
    if (!P || !shape) return fail(IBAMR_ERR_INVALID, "null argument");
    const int i = find_shape(P, shape);
    if (i < 0)
        return fail(IBAMR_ERR_INVALID, "tile shape " + std::to_string(shape[0]) + "x" + std::to_string(shape[1]) + "x" +
                                           std::to_string(shape[2]) + " not built for this plan");
    P->tcur = i;
    return IBAMR_OK;
}

int ibamr_plan_get_tile_shape(const ibamr_plan* P, int* shape)
{
    if (!P || !shape) return fail(IBAMR_ERR_INVALID, "null argument");
    for (int d = 0; d < 3; ++d) shape[d] = P->tlists.empty() ? 0 : P->tlists[P->tcur].shape[d];
    return IBAMR_OK;
}

int ibamr_plan_set_tile_height(ibamr_plan* P, int height)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    for (size_t i = 0; i < P->tlists.size(); ++i)
        if (P->tlists[i].shape[2] == height)
        {
            P->tcur = (int)i;
            return IBAMR_OK;
        }
    return fail(IBAMR_ERR_INVALID, "tile height " + std::to_string(height) + " not built for this plan");
}

int ibamr_plan_get_tile_height(const ibamr_plan* P, int* height)
{
    if (!P || !height) return fail(IBAMR_ERR_INVALID, "null argument");
    *height = P->tlists.empty() ? 0 : P->tlists[P->tcur].shape[2];
    return IBAMR_OK;
}

int ibamr_plan_autotune(ibamr_plan* P, int slot, const void* F, const double* X, int reps, int* shapes, float* ms,
                        int* ncand, void* stream)
{
    if (!P || !shapes || !ms || !ncand) return fail(IBAMR_ERR_INVALID, "null argument");
    if (P->npoints < 0) return fail(IBAMR_ERR_NOT_BINNED, "points not binned");
    *ncand = 0;
    if ((int)P->tlists.size() > IBAMR_MAX_TILE_SHAPES) return fail(IBAMR_ERR_INVALID, "internal: too many tile shapes");
    reps = std::max(1, reps);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool timing = P->timing;
    P->timing = false;  // the trials are not the caller's launches
    const int start = P->tcur;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best = start;
    float best_ms = std::numeric_limits<float>::infinity();
    // rounds over all shapes, each timing `reps` spreads per shape; a shape's
    // time is its fastest round (the shapes are interleaved, so a slow spell of
    // the device does not fall on one shape only)
    constexpr int ROUNDS = 3;
    auto run = [&]() -> int {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        const size_t ns = P->tlists.size();
        for (size_t i = 0; i < ns; ++i)
        {
            for (int d = 0; d < 3; ++d) shapes[3 * i + d] = P->tlists[i].shape[d];
            ms[i] = std::numeric_limits<float>::infinity();
            P->tcur = (int)i;
            const int r = ibamr_plan_spread(P, slot, F, X, IBAMR_SPREAD_OVERWRITE, stream);  // untimed
            if (r != IBAMR_OK) return r;
        }
        for (int round = 0; round < ROUNDS; ++round)
            for (size_t i = 0; i < ns; ++i)
            {
                P->tcur = (int)i;
                HIP_OK(hipEventRecord(e0, s));
                for (int k = 0; k < reps; ++k)
                {
                    const int r = ibamr_plan_spread(P, slot, F, X, IBAMR_SPREAD_OVERWRITE, stream);
                    if (r != IBAMR_OK) return r;
                }
                HIP_OK(hipEventRecord(e1, s));
                HIP_OK(hipEventSynchronize(e1));
                float t = 0.0f;
                HIP_OK(hipEventElapsedTime(&t, e0, e1));
                ms[i] = std::min(ms[i], t / reps);
            }
        for (size_t i = 0; i < ns; ++i)
            if (ms[i] < best_ms)
            {
                best_ms = ms[i];
                best = (int)i;
            }
        return IBAMR_OK;
    };
    const int rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    P->timing = timing;
    if (rc != IBAMR_OK)
    {
        P->tcur = start;
        return rc;
    }
    *ncand = (int)P->tlists.size();
    P->tcur = best;
    return IBAMR_OK;
}

} // extern "C"
