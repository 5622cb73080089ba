// le_interp.hip -- tensor-product interpolation (LDataManager::interp
// inner loop, LDataManager.cpp:699-813 -> LEInteractor::interpolate ->
// lagrangian_<kernel>_interp{2,3}d).
#include "le_plan_internal.h"

namespace ibamr
{
namespace
{
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
    // 3-D: superbins with a static owner and block (see build_superbin_owners)
    const int* sb_owner;
    int sb_off[2][2];
    int sbc;         // cells per superbin edge
    int nsy;         // superbins along y
    int dom_hi[3];
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
} // namespace

#include "le_interp3d.h"

// Raises the dynamic LDS limit of the 3-D interpolation kernels of a plan's
// kernel and precision (done once at plan creation, not per launch: the
// launches may be captured into a HIP graph).
int interp_prepare(int ndim, int kernel, int precision)
{
    if (ndim != 3) return IBAMR_OK;
    return dispatch(3, kernel, precision, [&](auto, auto ker, auto prec) -> int {
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
}

// Lowest and highest stencil start minus the point's cell over positions in the
// cell, for the cell-centred [0] and the side-centred [1] placement along an
// axis (weights_1d: IB_4/IB_6 nint((X - x_lower)/dx) - 2/3, BSPLINE floor based,
// PIECEWISE_LINEAR by the nearest cell centre).  Rounding of X / dx at a cell
// boundary can move a start by one more; the kernel checks each point against
// its staged block and reads the others from the array.
void stencil_start_offsets(int k, int off[2][2])
{
    static const int tab[7][2][2] = {
        {{-1, 0}, {0, 0}},     // PIECEWISE_LINEAR
        {{-2, -1}, {-1, -1}},  // IB_4
        {{-3, -2}, {-2, -2}},  // IB_6
        {{-1, -1}, {-1, 0}},   // BSPLINE_3
        {{-2, -1}, {-1, -1}},  // BSPLINE_4
        {{-2, -2}, {-2, -1}},  // BSPLINE_5
        {{-3, -2}, {-2, -2}},  // BSPLINE_6
    };
    for (int v = 0; v < 2; ++v)
        for (int i = 0; i < 2; ++i) off[v][i] = tab[k][v][i];
}

// Static superbins: the 3-D interpolation stages a superbin's block of each
// array from the superbin's cells when one patch owns all of them and every
// block lies in that patch's ghosted arrays and fits the LDS block (edge 16 + W);
// the kernel then skips the pass over the points that finds the owner and the
// exact extents.
int build_superbin_owners(ibamr_plan* P)
{
    if (P->ndim != 3) return IBAMR_OK;
    int off[2][2];
    stencil_start_offsets(P->kernel, off);
    for (int v = 0; v < 2; ++v)
    {
        P->sb_off[v][0] = off[v][0] - P->drift;
        P->sb_off[v][1] = off[v][1] + P->drift;
    }
    const int W = kernel_width(P->kernel), edge = 2 * P->bin + W, sbc = 2 * P->bin;
    const int nsb = P->nsb[0] * P->nsb[1] * P->nsb[2];
    std::vector<int> owner(nsb, -1);
    const bool side = P->centering == IBAMR_SIDE;
    for (int i = 0; i < P->npatches; ++i)
    {
        const PatchDev& g = P->patches_h[i];
        int s_lo[3], s_hi[3];
        for (int d = 0; d < 3; ++d)
        {
            s_lo[d] = std::max(0, (g.slo[d] - P->dom_lo[d]) / sbc);
            s_hi[d] = std::min(P->nsb[d] - 1, (g.shi[d] - P->dom_lo[d]) / sbc);
        }
        for (int sz = s_lo[2]; sz <= s_hi[2]; ++sz)
            for (int sy = s_lo[1]; sy <= s_hi[1]; ++sy)
                for (int sx = s_lo[0]; sx <= s_hi[0]; ++sx)
                {
                    const int sc[3] = {sx, sy, sz};
                    bool ok = true;
                    for (int d = 0; d < 3 && ok; ++d)
                    {
                        const int clo = P->dom_lo[d] + sbc * sc[d], chi = std::min(clo + sbc - 1, P->dom_hi[d]);
                        ok = clo >= g.slo[d] && chi <= g.shi[d];
                        for (int ax = 0; ax < (side ? 3 : 1) && ok; ++ax)
                        {
                            const int v = side && d == ax ? 1 : 0;
                            const int blo = clo + P->sb_off[v][0], bhi = chi + P->sb_off[v][1] + W - 1;
                            ok = blo >= g.ilo[d] - P->data_gcw[d] && bhi <= g.ihi[d] + P->data_gcw[d] + v &&
                                 bhi - blo + 1 <= edge;
                        }
                    }
                    if (ok) owner[(sz * P->nsb[1] + sy) * P->nsb[0] + sx] = i;
                }
    }
    HIP_OK(hipMalloc(&P->sb_owner_d, sizeof(int) * std::max(1, nsb)));
    HIP_OK(hipMemcpy(P->sb_owner_d, owner.data(), sizeof(int) * nsb, hipMemcpyHostToDevice));
    return IBAMR_OK;
}

} // namespace ibamr

extern "C" {

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
        a.sb_owner = P->sb_owner_d;
        for (int v = 0; v < 2; ++v)
            for (int i = 0; i < 2; ++i) a.sb_off[v][i] = P->sb_off[v][i];
        a.sbc = 2 * P->bin;
        a.nsy = P->nsb[1];
        for (int d = 0; d < 3; ++d) a.dom_hi[d] = P->dom_hi[d];
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

} // extern "C"
