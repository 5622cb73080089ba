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
#include "le_plan_internal.h"

namespace ibamr
{
thread_local std::string g_err;

// shared with the other translation units of the library
int set_error(int code, const std::string& msg) { return fail(code, msg); }

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
} // namespace ibamr

namespace ibamr
{
// IBMethod::forwardEulerStep position update (IBMethod.cpp:713-739, a VecWAXPY):
// X_new = X + dt U, elementwise over n values; U in FP64 or FP32 (FP32C plans).
template <typename TU>
__global__ void k_euler(double* __restrict__ Xn, const double* __restrict__ X, const TU* __restrict__ U, double dt,
                        long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Xn[i] = X[i] + dt * (double)U[i];
}

// X + dt U, then a position outside a periodic axis's [lo, hi) moved back by
// one period L = hi - lo (x < lo: x + L; then x >= hi: x - L), as
// LDataManager::beginDataRedistribution wraps positions (LDataManager.cpp:1403-1407;
// one period suffices for a step that moves a point by less than one).
template <typename TU>
__global__ void k_euler_wrap(double* __restrict__ Xn, const double* __restrict__ X, const TU* __restrict__ U,
                             double dt, long n, int ndim, double3 lo, double3 hi, double3 L, int3 per)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = (int)(i % ndim);
    double x = X[i] + dt * (double)U[i];
    const int p = d == 0 ? per.x : (d == 1 ? per.y : per.z);
    if (p)
    {
        const double l = d == 0 ? lo.x : (d == 1 ? lo.y : lo.z);
        const double h = d == 0 ? hi.x : (d == 1 ? hi.y : hi.z);
        const double len = d == 0 ? L.x : (d == 1 ? L.y : L.z);
        if (x < l) x = x + len;
        if (x >= h) x = x - len;
    }
    Xn[i] = x;
}

// Records the start event of launch `which` when timing is on; *end receives
// the event to record after the launch (nullptr when timing is off).
int time_begin(ibamr_plan* P, int which, hipStream_t s, hipEvent_t* end)
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
} // namespace ibamr

using namespace ibamr;

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
    for (TileList& t : P->tlists) dfree(t.d);
    dfree(P->wt_d);
    dfree(P->sb_owner_d);
    dfree(P->st_d);
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
    dfree(P->bin_start_d);
    dfree(P->keys2_d);
    dfree(P->pcell_off_d);
    dfree(P->bdry_patch_d);
    dfree(P->bdry_box_d);
    dfree(P->bdry_coef_d);
    dfree(P->bdry_coef_off_d);
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
        // the spread's work lists of every admissible tile shape (a shape change
        // later only selects another list)
        const int rc = build_tile_lists(P);
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
        // FP32C: stencil starts and weights of every point, per axis and variant
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
        int rc = interp_prepare(3, k, C->precision);
        if (rc == IBAMR_OK) rc = build_superbin_owners(P);
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
    if (ntiles) *ntiles = P->tlists.empty() ? 0 : P->tlists[P->tcur].n;
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

int ibamr_forward_euler_step_periodic(double* X_new, const double* X, const void* U, int u_precision, double dt,
                                      long npoints, int ndim, const double* x_lower, const double* x_upper,
                                      const int* periodic, void* stream)
{
    if (npoints < 0) return fail(IBAMR_ERR_INVALID, "npoints < 0");
    if (ndim != 2 && ndim != 3) return fail(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (!x_lower || !x_upper || !periodic) return fail(IBAMR_ERR_INVALID, "null domain description");
    if (npoints == 0) return IBAMR_OK;
    if (!X_new || !X || !U) return fail(IBAMR_ERR_INVALID, "null array");
    if (u_precision != IBAMR_FP64 && u_precision != IBAMR_FP32C) return fail(IBAMR_ERR_INVALID, "bad precision");
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, len[3] = {0, 0, 0};
    int per[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d)
    {
        lo[d] = x_lower[d];
        hi[d] = x_upper[d];
        len[d] = x_upper[d] - x_lower[d];
        per[d] = periodic[d] != 0;
        if (per[d] && !(len[d] > 0.0)) return fail(IBAMR_ERR_INVALID, "periodic axis with x_upper <= x_lower");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long n = npoints * ndim;
    const int B = 256;
    const dim3 g((unsigned)((n + B - 1) / B));
    const double3 l3 = make_double3(lo[0], lo[1], lo[2]), h3 = make_double3(hi[0], hi[1], hi[2]),
                  L3 = make_double3(len[0], len[1], len[2]);
    const int3 p3 = make_int3(per[0], per[1], per[2]);
    if (u_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_euler_wrap<double>, g, dim3(B), 0, s, X_new, X, static_cast<const double*>(U), dt, n,
                           ndim, l3, h3, L3, p3);
    else
        hipLaunchKernelGGL(k_euler_wrap<float>, g, dim3(B), 0, s, X_new, X, static_cast<const float*>(U), dt, n, ndim,
                           l3, h3, L3, p3);
    HIP_OK(hipGetLastError());
    return IBAMR_OK;
}

} // extern "C"
