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
    // FP32C 3-D (fixed point): the bound buffer (largest bin count, max |F| per
    // component) and the bins whose points can reach one grid value
    const unsigned* fx;
    double fx_factor;
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
// fmax[c] = max over the points of |F_c| as float bits (non-negative floats
// order like their bit patterns; a NaN sorts above every finite value), for
// the fixed-point scale of the FP32C spread.  Order-independent: atomicMax of
// integers.
__global__ __launch_bounds__(256) void k_fmax(const float* __restrict__ F, int n, int fdepth,
                                              unsigned* __restrict__ fmax)
{
    __shared__ unsigned s_m[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = 0; c < fdepth; ++c)
    {
        unsigned m = 0u;
        for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
            m = max(m, __float_as_uint(fabsf(F[(size_t)fdepth * k + c])));
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
        if (lane == 0) s_m[wv] = m;
        __syncthreads();
        if (threadIdx.x == 0)
        {
            const unsigned b = max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3]));
            if (b) atomicMax(fmax + c, b);
        }
        __syncthreads();
    }
}

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
    a.fx = P->fx_d;
    a.fx_factor = P->fx_factor;
}

// Spread tile shapes built for each form; the first is the default.
//   2-D: a 16 x 16 tile in LDS (k_spread_tiles).
//   3-D FP64: 8 x 8 grid lines times a column of z values in registers (k_spread3).
//   3-D FP32C: a box of 64-bit fixed-point values in LDS (k_spread3_fx); the
//     kernels of the path (IB_4, IB_6, BSPLINE_3) have several shapes, the
//     others the default only.
using Shape = std::array<int, 3>;
const std::vector<Shape>& tile_shapes(int ndim, int precision, int kernel)
{
    static const std::vector<Shape> s2 = {{16, 16, 1}};
    static const std::vector<Shape> s64 = {{8, 8, 32}, {8, 8, 16}};
    static const std::vector<Shape> sfx = {{16, 16, 64}, {16, 16, 16}, {16, 16, 32}, {32, 8, 64}, {32, 16, 8}};
    static const std::vector<Shape> sfx1 = {{16, 16, 64}, {16, 16, 16}};
    if (ndim == 2) return s2;
    if (precision == IBAMR_FP64) return s64;
    return (kernel == K_IB4 || kernel == K_IB6 || kernel == K_BS3) ? sfx : sfx1;
}

// FP32C 3-D: z extents up to 16 are LDS boxes (k_spread3_fx); 32 and more are
// column segments swept through an LDS window of FXZ_PLANES planes (k_spread3_fxz)
constexpr int FXZ_PLANES = 16;
bool is_column(const int* sh) { return sh[2] >= 32; }

// The FP32C kernel of one tile shape (the shapes of tile_shapes)
template <int KER>
int launch_fx(const ibamr_plan* P, const SpreadArgs<float>& a, hipStream_t s)
{
    const int* sh = P->tlists[P->tcur].shape;
    const dim3 G(P->tlists[P->tcur].n), B(FX_THREADS);
#define IBAMR_FX_BOX(X, Y, Z)                                                            \
    if (sh[0] == X && sh[1] == Y && sh[2] == Z)                                          \
    {                                                                                    \
        hipLaunchKernelGGL((k_spread3_fx<KER, X, Y, Z>), G, B, 0, s, a);                 \
        return IBAMR_OK;                                                                 \
    }
#define IBAMR_FX_COLUMN(X, Y)                                                            \
    if (sh[0] == X && sh[1] == Y && is_column(sh))                                       \
    {                                                                                    \
        hipLaunchKernelGGL((k_spread3_fxz<KER, X, Y, FXZ_PLANES>), G, B, 0, s, a, sh[2]); \
        return IBAMR_OK;                                                                 \
    }
    IBAMR_FX_COLUMN(16, 16)
    IBAMR_FX_BOX(16, 16, 16)
    if constexpr (KER == K_IB4 || KER == K_IB6 || KER == K_BS3)
    {
        IBAMR_FX_COLUMN(32, 8)
        IBAMR_FX_BOX(32, 16, 8)
    }
#undef IBAMR_FX_BOX
#undef IBAMR_FX_COLUMN
    return fail(IBAMR_ERR_INVALID, "internal: tile shape without a kernel");
}

// A shape is admissible when each axis of a tile reaches at most 16 (bin,
// image) entries (the image ranges partition the reachable cells, plus one
// split per domain side), and for a column segment when one bin layer's points
// reach no more planes than the window holds.
bool shape_admissible(const ibamr_plan* P, const Shape& sh)
{
    for (int d = 0; d < P->ndim; ++d)
        if ((sh[d] + 2 * P->reach + P->bin - 1) / P->bin + 3 > 16) return false;
    if (P->ndim == 3 && P->precision != IBAMR_FP64 && is_column(sh.data()) && P->bin + 2 * P->reach > FXZ_PLANES)
        return false;
    return true;
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
    for (const Shape& sh : tile_shapes(P->ndim, P->precision, P->kernel))
    {
        if (!shape_admissible(P, sh)) continue;
        P->tlists.emplace_back();
        const int rc = build_tiles(P, sh, P->tlists.back());
        if (rc != IBAMR_OK) return rc;
    }
    if (P->tlists.empty()) return fail(IBAMR_ERR_INVALID, "drift_cells too large for the tile size");
    P->tcur = 0;
    // FP32C fixed-point bound: the cells whose points can reach one grid value
    // span 2 reach + 1 cells per axis, so at most nblk bins, and with periodic
    // images as many more copies as that window can wrap around the period
    double f = 1.0;
    for (int d = 0; d < P->ndim; ++d)
    {
        const int win = 2 * P->reach + 1;
        const int nblk = (win - 1 + P->bin - 1) / P->bin + 1;
        const int wraps = P->pshift[d] > 0 ? 1 + (win - 1 + P->pshift[d] - 1) / P->pshift[d] : 1;
        f *= (double)nblk * wraps;
    }
    P->fx_factor = f;
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
        constexpr bool FX = NDIM == 3 && sizeof(T) == 4;
        SpreadArgs<T> a{};
        fill_spread_args<T>(P, slot, F, X, mode, a);
        const int B = 256;
        if (P->npoints > 0)
        {
            // the points in bin order
            if (P->fdepth == NDIM)
                hipLaunchKernelGGL((k_gather_points<NDIM, T>), dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->perm_d,
                                   P->npoints, X, P->Xs_d, static_cast<const T*>(F), static_cast<T*>(P->Fs_d),
                                   static_cast<const T*>(ds));
            else
                hipLaunchKernelGGL(k_scatter_points<T>, dim3((P->npoints + B - 1) / B), dim3(B), 0, s, P->idx_d,
                                   P->npoints, NDIM, X, P->Xs_d, P->fdepth, static_cast<const T*>(F),
                                   static_cast<T*>(P->Fs_d), static_cast<const T*>(ds));
            HIP_OK(hipGetLastError());
            P->xs_valid = true;
        }
        if constexpr (FX)
        {
            // fixed-point scale: max |F| per component (a fill kernel, not a
            // memset: the launches may be captured in a HIP graph)
            if (!P->fx_d) return fail(IBAMR_ERR_INVALID, "internal: no fixed-point bound buffer");
            hipLaunchKernelGGL(k_fill_int, dim3((P->fdepth + B - 1) / B), dim3(B), 0, s,
                               reinterpret_cast<int*>(P->fx_d + 1), P->fdepth, 0);
            if (P->npoints > 0)
            {
                const int G = std::min((P->npoints + B - 1) / B, 1024);
                hipLaunchKernelGGL(k_fmax, dim3(G), dim3(B), 0, s, static_cast<const float*>(P->Fs_d), P->npoints,
                                   P->fdepth, P->fx_d + 1);
            }
            HIP_OK(hipGetLastError());
        }
        a.X = P->Xs_d;
        a.F = static_cast<const T*>(P->Fs_d);
        if (!a.tiles || !a.patches || !a.arrays || !a.X || !a.F || !a.cells || !a.bin_start)
            return fail(IBAMR_ERR_INVALID, "internal: incomplete spread arguments");
        hipEvent_t t_end;
        const int trc = time_begin(P, 0, s, &t_end);
        if (trc != IBAMR_OK) return trc;
        const TileList& tl = P->tlists[P->tcur];
        if constexpr (FX)
        {
            const int rc = launch_fx<KER>(P, a, s);
            if (rc != IBAMR_OK) return rc;
        }
        else if constexpr (NDIM == 3)
        {
            switch (tl.shape[2])
            {
            case 16: hipLaunchKernelGGL((k_spread3<KER, 16>), dim3(tl.n), dim3(64), 0, s, a); break;
            case 32: hipLaunchKernelGGL((k_spread3<KER, 32>), dim3(tl.n), dim3(64), 0, s, a); break;
            default: return fail(IBAMR_ERR_INVALID, "internal: tile height without a kernel");
            }
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
    auto run = [&]() -> int {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
        for (size_t i = 0; i < P->tlists.size(); ++i)
        {
            for (int d = 0; d < 3; ++d) shapes[3 * i + d] = P->tlists[i].shape[d];
            P->tcur = (int)i;
            int r = ibamr_plan_spread(P, slot, F, X, IBAMR_SPREAD_OVERWRITE, stream);
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
                best = (int)i;
            }
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
