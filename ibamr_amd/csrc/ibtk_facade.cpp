// ibtk_facade.cpp -- the IBTK-shaped C++ layer (ibtk_amd.hpp) over the C ABI.
//
// LEInteractor's raw-array overloads take host std::vectors (as the reference
// does) and device patch data; each call copies X and Q to the device, bins
// the points on a cached single-patch plan (ibamr_plan_create_patch) and runs
// the GPU spread/interp.  LDataManager keeps one level plan and works on
// device-resident LData.  Errors that the reference raises with TBOX_ERROR
// throw ibtk_amd::Error with the reference's text.
#include "ibtk_amd.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <sstream>
#include <type_traits>

namespace ibtk_amd
{
namespace
{
void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

void abi_check(int rc, const char* where)
{
    if (rc < 0) throw Error(std::string(where) + ": " + ibamr_last_error());
}

template <class T>
T* dev_alloc(size_t n)
{
    void* p = nullptr;
    if (n) hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    return static_cast<T*>(p);
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

template <class T>
void fill_device(T* d, size_t n, T v)
{
    if (!n) return;
    if (v == T(0))
    {
        hip_check(hipMemset(d, 0, n * sizeof(T)), "hipMemset");
        return;
    }
    std::vector<T> h(n, v);
    hip_check(hipMemcpy(d, h.data(), n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

// LEInteractor::getStencilSize error text (LEInteractor.cpp:2103-2104)
void require_known(const std::string& k, const char* where)
{
    if (!ibamr_is_known_kernel(k.c_str()))
    {
        std::ostringstream m;
        m << where << "\n  Unknown kernel function " << k << "\n";
        throw Error(m.str());
    }
}
} // namespace

// ----------------------------------------------------------------------------- boxes and data

bool Box::contains(const Index& i) const
{
    for (int d = 0; d < d_ndim; ++d)
        if (i(d) < d_lower(d) || i(d) > d_upper(d)) return false;
    return true;
}

bool Box::operator==(const Box& o) const
{
    if (d_ndim != o.d_ndim) return false;
    for (int d = 0; d < d_ndim; ++d)
        if (d_lower(d) != o.d_lower(d) || d_upper(d) != o.d_upper(d)) return false;
    return true;
}

Box Box::grown(const IntVector& g) const
{
    Box b = *this;
    for (int d = 0; d < d_ndim; ++d)
    {
        b.d_lower(d) -= g(d);
        b.d_upper(d) += g(d);
    }
    return b;
}

long Box::size() const
{
    long n = 1;
    for (int d = 0; d < d_ndim; ++d) n *= std::max(0, d_upper(d) - d_lower(d) + 1);
    return n;
}

CartesianPatchGeometry::CartesianPatchGeometry(int ndim, const double* dx, const double* x_lower,
                                               const double* x_upper)
{
    for (int d = 0; d < ndim; ++d)
    {
        d_dx[d] = dx[d];
        d_xlo[d] = x_lower[d];
        d_xhi[d] = x_upper[d];
    }
}

template <class T>
CellData<T>::CellData(const Box& box, int depth, const IntVector& ghosts) : d_box(box), d_depth(depth), d_gcw(ghosts)
{
    if (depth < 1) throw Error("CellData: depth must be >= 1");
    d_data = dev_alloc<T>(components() * depth);
}

template <class T>
CellData<T>::~CellData()
{
    dev_free(d_data);
}

template <class T>
size_t CellData<T>::components() const
{
    return (size_t)getGhostBox().size();
}

template <class T>
void CellData<T>::fillAll(T v)
{
    fill_device(d_data, components() * d_depth, v);
}

template <class T>
void CellData<T>::copyFromHost(const std::vector<T>& h)
{
    if (h.size() != components() * d_depth) throw Error("CellData::copyFromHost: size mismatch");
    hip_check(hipMemcpy(d_data, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

template <class T>
void CellData<T>::copyToHost(std::vector<T>& h) const
{
    h.resize(components() * d_depth);
    hip_check(hipMemcpy(h.data(), d_data, h.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
}

template <class T>
SideData<T>::SideData(const Box& box, int depth, const IntVector& ghosts) : d_box(box), d_depth(depth), d_gcw(ghosts)
{
    if (depth < 1) throw Error("SideData: depth must be >= 1");
    for (int a = 0; a < box.getDim(); ++a) d_data[a] = dev_alloc<T>(components(a) * depth);
}

template <class T>
SideData<T>::~SideData()
{
    for (T* p : d_data) dev_free(p);
}

template <class T>
size_t SideData<T>::components(int axis) const
{
    Box g = d_box.grown(d_gcw);
    size_t n = 1;
    for (int d = 0; d < d_box.getDim(); ++d) n *= (size_t)(g.upper()(d) - g.lower()(d) + 1 + (d == axis ? 1 : 0));
    return n;
}

template <class T>
void SideData<T>::fillAll(T v)
{
    for (int a = 0; a < d_box.getDim(); ++a) fill_device(d_data[a], components(a) * d_depth, v);
}

template <class T>
void SideData<T>::copyFromHost(int axis, const std::vector<T>& h)
{
    if (h.size() != components(axis) * d_depth) throw Error("SideData::copyFromHost: size mismatch");
    hip_check(hipMemcpy(d_data[axis], h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

template <class T>
void SideData<T>::copyToHost(int axis, std::vector<T>& h) const
{
    h.resize(components(axis) * d_depth);
    hip_check(hipMemcpy(h.data(), d_data[axis], h.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
}

template <class T>
LData<T>::LData(int n, int depth) : d_n(n), d_depth(depth)
{
    if (n < 0 || depth < 1) throw Error("LData: bad size");
    d_data = dev_alloc<T>((size_t)n * depth);
}

template <class T>
LData<T>::~LData()
{
    dev_free(d_data);
}

template <class T>
void LData<T>::copyFromHost(const std::vector<T>& h)
{
    if (h.size() != (size_t)d_n * d_depth) throw Error("LData::copyFromHost: size mismatch");
    if (!h.empty()) hip_check(hipMemcpy(d_data, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

template <class T>
void LData<T>::copyToHost(std::vector<T>& h) const
{
    h.resize((size_t)d_n * d_depth);
    if (!h.empty()) hip_check(hipMemcpy(h.data(), d_data, h.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
}

template class CellData<double>;
template class CellData<float>;
template class SideData<double>;
template class SideData<float>;
template class LData<double>;
template class LData<float>;
template class LData<int>;

// ----------------------------------------------------------------------------- LEInteractor

bool LEInteractor::isKnownKernel(const std::string& kernel_fcn) { return ibamr_is_known_kernel(kernel_fcn.c_str()) != 0; }

int LEInteractor::getStencilSize(const std::string& kernel_fcn)
{
    require_known(kernel_fcn, "LEInteractor::getStencilSize()");
    return ibamr_get_stencil_size(kernel_fcn.c_str());
}

int LEInteractor::getMinimumGhostWidth(const std::string& kernel_fcn)
{
    require_known(kernel_fcn, "LEInteractor::getStencilSize()");
    return ibamr_get_minimum_ghost_width(kernel_fcn.c_str());
}

namespace
{
// One cached single-patch plan per (geometry, box, kernel, centering, depth,
// ghost width); capacity grows to the largest call.  Device staging for X and Q.
struct RawPlan
{
    ibamr_plan* plan = nullptr;
    int capacity = 0;
};

struct RawCache
{
    std::mutex mu;
    std::vector<std::pair<std::string, RawPlan>> plans;
    double* X_d = nullptr;
    double* Q_d = nullptr;
    size_t X_cap = 0, Q_cap = 0;
};

RawCache& raw_cache()
{
    // Plans and staging live until the process ends (no teardown after the HIP
    // runtime has gone).
    static RawCache* c = new RawCache();
    return *c;
}

std::string plan_key(const Patch& patch, const Box& box, const std::string& k, int centering, int depth,
                     const IntVector& gcw)
{
    std::ostringstream s;
    s.precision(17);
    const int nd = patch.getBox().getDim();
    s << nd << '|' << k << '|' << centering << '|' << depth;
    for (int d = 0; d < nd; ++d)
        s << '|' << patch.getBox().lower()(d) << ',' << patch.getBox().upper()(d) << ',' << box.lower()(d) << ','
          << box.upper()(d) << ',' << gcw(d) << ',' << patch.getPatchGeometry().getDx()[d] << ','
          << patch.getPatchGeometry().getXLower()[d] << ',' << patch.getPatchGeometry().getXUpper()[d];
    return s.str();
}

ibamr_plan* get_raw_plan(RawCache& c, const Patch& patch, const Box& box, const std::string& k, int centering,
                         int depth, const IntVector& gcw, int npoints)
{
    const std::string key = plan_key(patch, box, k, centering, depth, gcw);
    RawPlan* rp = nullptr;
    for (auto& e : c.plans)
        if (e.first == key) rp = &e.second;
    if (rp && rp->capacity >= npoints) return rp->plan;
    if (rp)
    {
        ibamr_plan_destroy(rp->plan);
        rp->plan = nullptr;
    }
    const int nd = patch.getBox().getDim();
    ibamr_patch_geom g{};
    int blo[3] = {0, 0, 0}, bhi[3] = {0, 0, 0};
    double dx[3] = {1, 1, 1};
    for (int d = 0; d < nd; ++d)
    {
        g.ilower[d] = patch.getBox().lower()(d);
        g.iupper[d] = patch.getBox().upper()(d);
        g.x_lower[d] = patch.getPatchGeometry().getXLower()[d];
        g.x_upper[d] = patch.getPatchGeometry().getXUpper()[d];
        blo[d] = box.lower()(d);
        bhi[d] = box.upper()(d);
        dx[d] = patch.getPatchGeometry().getDx()[d];
    }
    ibamr_plan_config C{};
    C.kernel = k.c_str();
    C.centering = centering;
    C.precision = IBAMR_FP64;
    C.depth = depth;
    for (int d = 0; d < nd; ++d) C.data_gcw[d] = gcw(d);
    C.idx_gcw = 0;
    const int cap = std::max(npoints, rp ? 2 * rp->capacity : 1024);
    C.max_points = cap;
    ibamr_plan* p = nullptr;
    abi_check(ibamr_plan_create_patch(&p, nd, dx, &g, blo, bhi, &C), "LEInteractor");
    if (rp)
    {
        rp->plan = p;
        rp->capacity = cap;
    }
    else
        c.plans.push_back({key, RawPlan{p, cap}});
    return p;
}

template <class T>
void ensure(T*& p, size_t& cap, size_t n)
{
    if (n <= cap) return;
    dev_free(p);
    p = nullptr;
    cap = 0;
    p = dev_alloc<T>(n);
    cap = n;
}

void ghost_check(const char* where, const std::string& k, const IntVector& gcw, int nd)
{
    const int mg = ibamr_get_minimum_ghost_width(k.c_str());
    int gmin = gcw(0);
    for (int d = 1; d < nd; ++d) gmin = std::min(gmin, gcw(d));
    if (gmin < mg)
    {
        std::ostringstream m;
        m << where << ": insufficient ghost cells:"
          << "  kernel function          = " << k << "\n"
          << "  kernel stencil size      = " << ibamr_get_stencil_size(k.c_str()) << "\n"
          << "  minimum ghost cell width = " << mg << "\n"
          << "  ghost cell width         = " << gmin << "\n";
        throw Error(m.str());
    }
}

// The common body of the raw-array overloads: arrays = one pointer (cell) or
// one per axis (side); interp (spread == false) or spread.
void raw_call(bool spread, std::vector<double>* Qw, const std::vector<double>* Qr, int Q_depth,
              const std::vector<double>& X_data, int X_depth, void* const* arrays, int centering, int q_depth,
              const IntVector& gcw, const Patch& patch, const Box& box, const std::string& k)
{
    const char* where = spread ? "LEInteractor::spread()" : "LEInteractor::interpolate()";
    const int nd = patch.getBox().getDim();
    require_known(k, where);
    if (X_depth != nd) throw Error(std::string(where) + ": X_depth must equal NDIM");
    if (centering == IBAMR_SIDE && (Q_depth != nd || q_depth != 1))
        throw Error(std::string(where) + ":\n  side-centered " + (spread ? "spreading" : "interpolation") +
                    " requires vector-valued data.\n");
    if (centering == IBAMR_CELL && Q_depth != q_depth) throw Error(std::string(where) + ": Q_depth != q_data depth");
    const std::vector<double>& Q = spread ? *Qr : *Qw;
    const int n = (int)(X_data.size() / nd);
    if (Q.size() / Q_depth != (size_t)n) throw Error(std::string(where) + ": Q_data and X_data sizes disagree");
    ghost_check(where, k, gcw, nd);
    if (n == 0) return;

    RawCache& c = raw_cache();
    std::lock_guard<std::mutex> lock(c.mu);
    ibamr_plan* p = get_raw_plan(c, patch, box, k, centering, centering == IBAMR_SIDE ? 1 : q_depth, gcw, n);
    ensure(c.X_d, c.X_cap, (size_t)n * nd);
    ensure(c.Q_d, c.Q_cap, (size_t)n * Q_depth);
    hip_check(hipMemcpy(c.X_d, X_data.data(), (size_t)n * nd * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(c.Q_d, Q.data(), (size_t)n * Q_depth * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    abi_check(ibamr_plan_bind_data(p, 0, arrays), where);
    abi_check(ibamr_plan_bin_async(p, c.X_d, n, nullptr), where);
    if (spread)
        abi_check(ibamr_plan_spread(p, 0, c.Q_d, c.X_d, IBAMR_SPREAD_ACCUMULATE, nullptr), where);
    else
    {
        abi_check(ibamr_plan_interp(p, 0, c.Q_d, c.X_d, nullptr), where);
        hip_check(hipMemcpy(Qw->data(), c.Q_d, (size_t)n * Q_depth * sizeof(double), hipMemcpyDeviceToHost),
                  "hipMemcpy");
    }
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
} // namespace

void LEInteractor::interpolate(std::vector<double>& Q_data, int Q_depth, const std::vector<double>& X_data,
                               int X_depth, const std::shared_ptr<CellData<double>>& q_data,
                               const std::shared_ptr<Patch>& patch, const Box& interp_box,
                               const std::string& interp_fcn)
{
    if (Q_data.empty()) return;
    void* arr[1] = {q_data->getPointer()};
    raw_call(false, &Q_data, nullptr, Q_depth, X_data, X_depth, arr, IBAMR_CELL, q_data->getDepth(),
             q_data->getGhostCellWidth(), *patch, interp_box, interp_fcn);
}

void LEInteractor::interpolate(std::vector<double>& Q_data, int Q_depth, const std::vector<double>& X_data,
                               int X_depth, const std::shared_ptr<SideData<double>>& q_data,
                               const std::shared_ptr<Patch>& patch, const Box& interp_box,
                               const std::string& interp_fcn)
{
    if (Q_data.empty()) return;
    void* arr[3] = {q_data->getPointer(0), q_data->getPointer(1),
                    patch->getBox().getDim() == 3 ? q_data->getPointer(2) : nullptr};
    raw_call(false, &Q_data, nullptr, Q_depth, X_data, X_depth, arr, IBAMR_SIDE, q_data->getDepth(),
             q_data->getGhostCellWidth(), *patch, interp_box, interp_fcn);
}

void LEInteractor::spread(const std::shared_ptr<CellData<double>>& q_data, const std::vector<double>& Q_data,
                          int Q_depth, const std::vector<double>& X_data, int X_depth,
                          const std::shared_ptr<Patch>& patch, const Box& spread_box, const std::string& spread_fcn)
{
    if (Q_data.empty()) return;
    void* arr[1] = {q_data->getPointer()};
    raw_call(true, nullptr, &Q_data, Q_depth, X_data, X_depth, arr, IBAMR_CELL, q_data->getDepth(),
             q_data->getGhostCellWidth(), *patch, spread_box, spread_fcn);
}

void LEInteractor::spread(const std::shared_ptr<SideData<double>>& q_data, const std::vector<double>& Q_data,
                          int Q_depth, const std::vector<double>& X_data, int X_depth,
                          const std::shared_ptr<Patch>& patch, const Box& spread_box, const std::string& spread_fcn)
{
    if (Q_data.empty()) return;
    void* arr[3] = {q_data->getPointer(0), q_data->getPointer(1),
                    patch->getBox().getDim() == 3 ? q_data->getPointer(2) : nullptr};
    raw_call(true, nullptr, &Q_data, Q_depth, X_data, X_depth, arr, IBAMR_SIDE, q_data->getDepth(),
             q_data->getGhostCellWidth(), *patch, spread_box, spread_fcn);
}

// ----------------------------------------------------------------------------- LDataManager

struct LDataManager::Impl
{
    Level level;
    int ndim = 3;
    ibamr_plan* spread_plan = nullptr;
    ibamr_plan* interp_plan = nullptr;  // same kernel: shares the spread plan
    int npoints = -1;
    int max_points = 0;
    std::vector<void*> bound[2];
};

namespace
{
ibamr_plan* make_level_plan(const Level& L, const std::string& k, const IntVector& gw, int max_points)
{
    require_known(k, "LDataManager");
    std::vector<ibamr_patch_geom> pg;
    for (const auto& p : L.patches)
    {
        ibamr_patch_geom g{};
        for (int d = 0; d < L.ndim; ++d)
        {
            g.ilower[d] = p->getBox().lower()(d);
            g.iupper[d] = p->getBox().upper()(d);
            g.x_lower[d] = p->getPatchGeometry().getXLower()[d];
            g.x_upper[d] = p->getPatchGeometry().getXUpper()[d];
        }
        pg.push_back(g);
    }
    ibamr_level_geom G{};
    G.ndim = L.ndim;
    G.npatches = (int)pg.size();
    G.patches = pg.data();
    int gmax = 0;
    for (int d = 0; d < L.ndim; ++d)
    {
        G.dx[d] = L.dx[d];
        G.domain_lower[d] = L.domain.lower()(d);
        G.domain_upper[d] = L.domain.upper()(d);
        G.domain_x_lower[d] = L.x_lower[d];
        G.domain_x_upper[d] = L.x_upper[d];
        G.periodic_shift[d] = L.periodic_shift(d);
        gmax = std::max(gmax, gw(d));
    }
    ibamr_plan_config C{};
    C.kernel = k.c_str();
    C.centering = IBAMR_SIDE;
    C.precision = IBAMR_FP64;
    C.depth = 1;
    for (int d = 0; d < L.ndim; ++d) C.data_gcw[d] = gw(d);
    C.idx_gcw = gmax;
    C.max_points = max_points;
    ibamr_plan* p = nullptr;
    abi_check(ibamr_plan_create(&p, &G, &C), "LDataManager");
    return p;
}

std::vector<void*> side_arrays(const std::vector<std::shared_ptr<SideData<double>>>& f, int nd, size_t npatches)
{
    if (f.size() != npatches) throw Error("LDataManager: one SideData per patch expected");
    std::vector<void*> a;
    for (const auto& s : f)
        for (int d = 0; d < nd; ++d) a.push_back(const_cast<double*>(s->getPointer(d)));
    return a;
}
} // namespace

LDataManager::LDataManager(const Level& level, const std::string& default_interp_kernel_fcn,
                           const std::string& default_spread_kernel_fcn, const IntVector& ghost_width, int max_points)
    : d_impl(new Impl())
{
    d_impl->level = level;
    d_impl->ndim = level.ndim;
    d_impl->max_points = max_points;
    d_impl->spread_plan = make_level_plan(level, default_spread_kernel_fcn, ghost_width, max_points);
    d_impl->interp_plan = default_interp_kernel_fcn == default_spread_kernel_fcn
                              ? d_impl->spread_plan
                              : make_level_plan(level, default_interp_kernel_fcn, ghost_width, max_points);
}

LDataManager::~LDataManager()
{
    if (d_impl->interp_plan != d_impl->spread_plan) ibamr_plan_destroy(d_impl->interp_plan);
    ibamr_plan_destroy(d_impl->spread_plan);
}

void LDataManager::endDataRedistribution(const LData<double>& X_data)
{
    const int n = X_data.getLocalNodeCount();
    abi_check(ibamr_plan_bin(d_impl->spread_plan, X_data.getArray(), n, nullptr), "LDataManager");
    if (d_impl->interp_plan != d_impl->spread_plan)
        abi_check(ibamr_plan_bin(d_impl->interp_plan, X_data.getArray(), n, nullptr), "LDataManager");
    d_impl->npoints = n;
}

std::vector<int> LDataManager::redistribute(LData<double>& X_data, const std::vector<LData<double>*>& data,
                                            LData<int>* lag_index, bool error_if_points_leave_domain)
{
    const int n = X_data.getLocalNodeCount();
    if (X_data.getDepth() != d_impl->ndim) throw Error("LDataManager::redistribute: X must have depth NDIM");
    for (const LData<double>* d : data)
        if (!d || d->getLocalNodeCount() != n) throw Error("LDataManager::redistribute: data sizes differ from X");
    if (lag_index && lag_index->getLocalNodeCount() != n)
        throw Error("LDataManager::redistribute: Lagrangian index size differs from X");
    const int np = (int)d_impl->level.patches.size();
    int* order = dev_alloc<int>((size_t)std::max(n, 1));
    int* offsets = dev_alloc<int>((size_t)np + 1);
    try
    {
        abi_check(ibamr_plan_redistribute(d_impl->spread_plan, X_data.getArray(),
                                          lag_index ? lag_index->getArray() : nullptr, n,
                                          error_if_points_leave_domain ? 1 : 0, order, offsets, nullptr, nullptr),
                  "LDataManager::redistribute");
        auto move = [&](auto& L) {
            using T = typename std::remove_reference<decltype(*L.getArray())>::type;
            LData<T> tmp(n, L.getDepth());
            abi_check(ibamr_permute_rows(tmp.getArray(), L.getArray(), order, n, L.getDepth(), (int)sizeof(T), nullptr),
                      "LDataManager::redistribute");
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            L.swap(tmp);
        };
        move(X_data);
        for (LData<double>* d : data) move(*d);
        if (lag_index) move(*lag_index);
    }
    catch (...)
    {
        dev_free(order);
        dev_free(offsets);
        throw;
    }
    std::vector<int> h(np + 1);
    hip_check(hipMemcpy(h.data(), offsets, sizeof(int) * (np + 1), hipMemcpyDeviceToHost), "hipMemcpy");
    dev_free(order);
    dev_free(offsets);
    d_impl->npoints = -1;
    endDataRedistribution(X_data);
    return h;
}

IBStandardForceGen::IBStandardForceGen(int ndim, int num_nodes, const std::vector<Spring>& springs,
                                       const std::vector<Beam>& beams, const std::vector<TargetPoint>& targets)
    : d_ndim(ndim), d_n(num_nodes)
{
    std::vector<int> sm, ss, bm, bn, bp, tn;
    std::vector<double> sk, sr, bK, bc, tK, tE, tX;
    d_par_off.push_back(0);
    for (const Spring& s : springs)
    {
        sm.push_back(s.master);
        ss.push_back(s.slave);
        sk.push_back(s.stiffness);
        sr.push_back(s.rest_length);
        d_fcn_idx.push_back(s.force_fcn_idx);
        d_par.push_back(s.stiffness);
        d_par.push_back(s.rest_length);
        d_par.insert(d_par.end(), s.extra_parameters.begin(), s.extra_parameters.end());
        d_par_off.push_back((int)d_par.size());
        if (s.force_fcn_idx != 0) d_resolve = true;
    }
    for (const Beam& b : beams)
    {
        bm.push_back(b.master);
        bn.push_back(b.next);
        bp.push_back(b.prev);
        bK.push_back(b.rigidity);
        for (int d = 0; d < ndim; ++d) bc.push_back(b.curvature[d]);
    }
    for (const TargetPoint& t : targets)
    {
        tn.push_back(t.node);
        tK.push_back(t.stiffness);
        tE.push_back(t.damping);
        for (int d = 0; d < ndim; ++d) tX.push_back(t.X0[d]);
    }
    abi_check(ibamr_forces_create(&d_forces, ndim, num_nodes, (int)springs.size(), sm.data(), ss.data(), sk.data(),
                                  sr.data(), (int)beams.size(), bm.data(), bn.data(), bp.data(), bK.data(), bc.data(),
                                  (int)targets.size(), tn.data(), tK.data(), tE.data(), tX.data()),
              "IBStandardForceGen");
}

IBStandardForceGen::~IBStandardForceGen()
{
    ibamr_forces_destroy(d_forces);
}

void IBStandardForceGen::registerSpringForceFunction(int force_fcn_index, int law)
{
    abi_check(ibamr_forces_register_spring_function(d_forces, force_fcn_index, law),
              "IBStandardForceGen::registerSpringForceFunction");
    d_resolve = true;
}

void IBStandardForceGen::initializeLevelData()
{
    abi_check(ibamr_forces_set_spring_parameters(d_forces, d_fcn_idx.data(), d_par_off.data(), d_par.data()),
              "IBStandardForceGen::initializeLevelData");
    d_resolve = false;
}

void IBStandardForceGen::computeLagrangianForce(LData<double>& F, const LData<double>& X, const LData<double>& U)
{
    if (d_resolve) initializeLevelData();
    if (F.getLocalNodeCount() != d_n || X.getLocalNodeCount() != d_n || U.getLocalNodeCount() != d_n ||
        F.getDepth() != d_ndim || X.getDepth() != d_ndim || U.getDepth() != d_ndim)
        throw Error("IBStandardForceGen::computeLagrangianForce: F, X and U must hold NDIM values per node");
    abi_check(ibamr_forces_compute(d_forces, F.getArray(), IBAMR_FP64, X.getArray(), U.getArray(), IBAMR_FP64,
                                   IBAMR_SPREAD_ACCUMULATE, nullptr),
              "IBStandardForceGen::computeLagrangianForce");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void LDataManager::spread(std::vector<std::shared_ptr<SideData<double>>>& f_data, const LData<double>& F_data,
                          const LData<double>& X_data)
{
    if (d_impl->npoints != X_data.getLocalNodeCount()) endDataRedistribution(X_data);
    std::vector<void*> a = side_arrays(f_data, d_impl->ndim, d_impl->level.patches.size());
    if (a != d_impl->bound[0])
    {
        abi_check(ibamr_plan_bind_data(d_impl->spread_plan, 0, a.data()), "LDataManager::spread");
        d_impl->bound[0] = a;
    }
    abi_check(ibamr_plan_spread(d_impl->spread_plan, 0, F_data.getArray(), X_data.getArray(),
                                IBAMR_SPREAD_ACCUMULATE_INTERIOR, nullptr),
              "LDataManager::spread");
}

void LDataManager::spread(std::vector<std::shared_ptr<SideData<double>>>& f_data, const LData<double>& F_data,
                          const LData<double>& X_data, const LData<double>& ds_data)
{
    if (ds_data.getLocalNodeCount() != F_data.getLocalNodeCount() || ds_data.getDepth() != 1)
        throw Error("LDataManager::spread: ds_data must hold one value per node of F_data");
    if (d_impl->npoints != X_data.getLocalNodeCount()) endDataRedistribution(X_data);
    std::vector<void*> a = side_arrays(f_data, d_impl->ndim, d_impl->level.patches.size());
    if (a != d_impl->bound[0])
    {
        abi_check(ibamr_plan_bind_data(d_impl->spread_plan, 0, a.data()), "LDataManager::spread");
        d_impl->bound[0] = a;
    }
    abi_check(ibamr_plan_spread_ds(d_impl->spread_plan, 0, F_data.getArray(), X_data.getArray(), ds_data.getArray(),
                                   IBAMR_SPREAD_ACCUMULATE_INTERIOR, nullptr),
              "LDataManager::spread");
}

void LDataManager::interp(const std::vector<std::shared_ptr<SideData<double>>>& u_data, LData<double>& U_data,
                          const LData<double>& X_data)
{
    if (d_impl->npoints != X_data.getLocalNodeCount()) endDataRedistribution(X_data);
    std::vector<void*> a = side_arrays(u_data, d_impl->ndim, d_impl->level.patches.size());
    if (a != d_impl->bound[1])
    {
        abi_check(ibamr_plan_bind_data(d_impl->interp_plan, 1, a.data()), "LDataManager::interp");
        d_impl->bound[1] = a;
    }
    abi_check(ibamr_plan_interp(d_impl->interp_plan, 1, U_data.getArray(), X_data.getArray(), nullptr),
              "LDataManager::interp");
}

void LDataManager::fillPeriodicGhosts(std::vector<std::shared_ptr<SideData<double>>>& u_data)
{
    std::vector<void*> a = side_arrays(u_data, d_impl->ndim, d_impl->level.patches.size());
    if (a != d_impl->bound[1])
    {
        abi_check(ibamr_plan_bind_data(d_impl->interp_plan, 1, a.data()), "LDataManager::fillPeriodicGhosts");
        d_impl->bound[1] = a;
    }
    abi_check(ibamr_plan_fill_periodic_ghosts(d_impl->interp_plan, 1, nullptr), "LDataManager::fillPeriodicGhosts");
}

void LDataManager::fillGhosts(std::vector<std::shared_ptr<SideData<double>>>& u_data)
{
    std::vector<void*> a = side_arrays(u_data, d_impl->ndim, d_impl->level.patches.size());
    if (a != d_impl->bound[1])
    {
        abi_check(ibamr_plan_bind_data(d_impl->interp_plan, 1, a.data()), "LDataManager::fillGhosts");
        d_impl->bound[1] = a;
    }
    abi_check(ibamr_plan_fill_ghosts(d_impl->interp_plan, 1, nullptr), "LDataManager::fillGhosts");
}

void IBMethod::forwardEulerStep(LData<double>& X_new, const LData<double>& X, const LData<double>& U, double dt)
{
    const long n = (long)X.getLocalNodeCount() * X.getDepth();
    if ((long)X_new.getLocalNodeCount() * X_new.getDepth() != n || (long)U.getLocalNodeCount() * U.getDepth() != n)
        throw Error("IBMethod::forwardEulerStep: X_new, X and U sizes differ");
    abi_check(ibamr_forward_euler_step(X_new.getArray(), X.getArray(), U.getArray(), IBAMR_FP64, dt, n, nullptr),
              "IBMethod::forwardEulerStep");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

} // namespace ibtk_amd
