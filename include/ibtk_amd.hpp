// ibtk_amd.hpp -- C++ host layer mirroring the IBTK/IBAMR interface of the
// Eulerian-Lagrangian coupling path, on top of the C ABI (ibamr_amd.h).
//
// Names, argument meaning and error behaviour follow the reference:
//   LEInteractor      ibtk/include/ibtk/LEInteractor.h (isKnownKernel :105,
//                     getStencilSize :99, getMinimumGhostWidth :117, the
//                     std::vector raw-array overloads of interpolate :342-486
//                     and spread :880-1200)
//   LDataManager      ibtk/include/ibtk/LDataManager.h spread :292-474, interp :491-545
//   IBMethod hooks    include/ibamr/IBStrategy.h interpolateVelocity :276,
//                     spreadForce :339; IBMethod::forwardEulerStep (src/IB/IBMethod.cpp:713-739)
// SAMRAI is not a dependency: Box, Patch, CartesianPatchGeometry, CellData and
// SideData below are small stand-ins holding what the path reads (boxes, ghost
// widths, geometry, device arrays in SAMRAI's ArrayData layout); SAMRAI's
// Pointer<T> becomes std::shared_ptr<T>.  Errors the reference reports with
// TBOX_ERROR throw ibtk_amd::Error with the reference's message.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ibamr_amd.h"

namespace ibtk_amd
{
struct Error : std::runtime_error
{
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// SAMRAI hier::Index / IntVector (up to 3 components; unused components 0).
struct IntVector
{
    std::array<int, 3> v{{0, 0, 0}};
    IntVector() = default;
    explicit IntVector(int a) : v{{a, a, a}} {}
    IntVector(int a, int b, int c = 0) : v{{a, b, c}} {}
    int& operator()(int d) { return v[d]; }
    int operator()(int d) const { return v[d]; }
};
using Index = IntVector;

class Box
{
public:
    Box() = default;
    Box(int ndim, const Index& lower, const Index& upper) : d_ndim(ndim), d_lower(lower), d_upper(upper) {}
    int getDim() const { return d_ndim; }
    const Index& lower() const { return d_lower; }
    const Index& upper() const { return d_upper; }
    bool contains(const Index& i) const;
    bool operator==(const Box& o) const;
    Box grown(const IntVector& g) const;
    long size() const;

private:
    int d_ndim = 3;
    Index d_lower, d_upper;
};

class CartesianPatchGeometry
{
public:
    CartesianPatchGeometry() = default;
    CartesianPatchGeometry(int ndim, const double* dx, const double* x_lower, const double* x_upper);
    const double* getDx() const { return d_dx.data(); }
    const double* getXLower() const { return d_xlo.data(); }
    const double* getXUpper() const { return d_xhi.data(); }

private:
    std::array<double, 3> d_dx{{1, 1, 1}}, d_xlo{{0, 0, 0}}, d_xhi{{1, 1, 1}};
};

class Patch
{
public:
    Patch(const Box& box, const CartesianPatchGeometry& geom) : d_box(box), d_geom(geom) {}
    const Box& getBox() const { return d_box; }
    const CartesianPatchGeometry& getPatchGeometry() const { return d_geom; }

private:
    Box d_box;
    CartesianPatchGeometry d_geom;
};

// Device-resident cell-centred data: one ghosted array, depth slowest.
template <class T>
class CellData
{
public:
    CellData(const Box& box, int depth, const IntVector& ghosts);
    ~CellData();
    CellData(const CellData&) = delete;
    CellData& operator=(const CellData&) = delete;
    T* getPointer(int d = 0) { return d_data + (size_t)d * components(); }
    const T* getPointer(int d = 0) const { return d_data + (size_t)d * components(); }
    const Box& getBox() const { return d_box; }
    Box getGhostBox() const { return d_box.grown(d_gcw); }
    const IntVector& getGhostCellWidth() const { return d_gcw; }
    int getDepth() const { return d_depth; }
    size_t components() const;     // values per depth component (ghost box)
    void fillAll(T v);
    void copyFromHost(const std::vector<T>& h);
    void copyToHost(std::vector<T>& h) const;

private:
    Box d_box;
    int d_depth;
    IntVector d_gcw;
    T* d_data = nullptr;
};

// Device-resident side-centred data: one ghosted array per axis over
// toSideBox(box, axis) (upper index + 1 along the axis), depth slowest.
template <class T>
class SideData
{
public:
    SideData(const Box& box, int depth, const IntVector& ghosts);
    ~SideData();
    SideData(const SideData&) = delete;
    SideData& operator=(const SideData&) = delete;
    T* getPointer(int axis, int d = 0) { return d_data[axis] + (size_t)d * components(axis); }
    const T* getPointer(int axis, int d = 0) const { return d_data[axis] + (size_t)d * components(axis); }
    const Box& getBox() const { return d_box; }
    const IntVector& getGhostCellWidth() const { return d_gcw; }
    int getDepth() const { return d_depth; }
    size_t components(int axis) const;
    void fillAll(T v);
    void copyFromHost(int axis, const std::vector<T>& h);
    void copyToHost(int axis, std::vector<T>& h) const;

private:
    Box d_box;
    int d_depth;
    IntVector d_gcw;
    std::array<T*, 3> d_data{{nullptr, nullptr, nullptr}};
};

// IBTK::LEInteractor: the kernel registry and the raw-array overloads
// (std::vector host data, device patch data).  The points interpolated or
// spread are those whose cell (IndexUtilities::getCellIndex on the patch) lies
// in the given box, with no periodic shifts (LEInteractor.cpp:2725-2802).
class LEInteractor
{
public:
    static bool isKnownKernel(const std::string& kernel_fcn);
    static int getStencilSize(const std::string& kernel_fcn);
    static int getMinimumGhostWidth(const std::string& kernel_fcn);

    static void interpolate(std::vector<double>& Q_data, int Q_depth, const std::vector<double>& X_data,
                            int X_depth, const std::shared_ptr<CellData<double>>& q_data,
                            const std::shared_ptr<Patch>& patch, const Box& interp_box,
                            const std::string& interp_fcn = "IB_4");
    static void interpolate(std::vector<double>& Q_data, int Q_depth, const std::vector<double>& X_data,
                            int X_depth, const std::shared_ptr<SideData<double>>& q_data,
                            const std::shared_ptr<Patch>& patch, const Box& interp_box,
                            const std::string& interp_fcn = "IB_4");
    static void spread(const std::shared_ptr<CellData<double>>& q_data, const std::vector<double>& Q_data,
                       int Q_depth, const std::vector<double>& X_data, int X_depth,
                       const std::shared_ptr<Patch>& patch, const Box& spread_box,
                       const std::string& spread_fcn = "IB_4");
    static void spread(const std::shared_ptr<SideData<double>>& q_data, const std::vector<double>& Q_data,
                       int Q_depth, const std::vector<double>& X_data, int X_depth,
                       const std::shared_ptr<Patch>& patch, const Box& spread_box,
                       const std::string& spread_fcn = "IB_4");
};

// A patch level as LDataManager sees it.
struct Level
{
    int ndim = 3;
    std::array<double, 3> dx{{1, 1, 1}};
    Box domain;                                   // level index space
    std::array<double, 3> x_lower{{0, 0, 0}}, x_upper{{1, 1, 1}};
    IntVector periodic_shift;                     // cells, 0 = not periodic
    std::vector<std::shared_ptr<Patch>> patches;
};

// Lagrangian data (IBTK::LData local form): device array [n][depth].
template <class T>
class LData
{
public:
    LData(int n, int depth);
    ~LData();
    LData(const LData&) = delete;
    LData& operator=(const LData&) = delete;
    T* getArray() { return d_data; }
    const T* getArray() const { return d_data; }
    int getLocalNodeCount() const { return d_n; }
    int getDepth() const { return d_depth; }
    void copyFromHost(const std::vector<T>& h);
    void copyToHost(std::vector<T>& h) const;
    // exchange contents (and sizes) with another LData of the same type
    void swap(LData& o)
    {
        std::swap(d_n, o.d_n);
        std::swap(d_depth, o.d_depth);
        std::swap(d_data, o.d_data);
    }

private:
    int d_n, d_depth;
    T* d_data = nullptr;
};

// IBTK::LDataManager::spread/interp for one level on the GPU.  The plan is
// made once (regrid); beginDataRedistribution/endDataRedistribution re-bin the
// points (LDataManager.cpp:1520-1960, the cell association part).
class LDataManager
{
public:
    LDataManager(const Level& level, const std::string& default_interp_kernel_fcn,
                 const std::string& default_spread_kernel_fcn, const IntVector& ghost_width, int max_points);
    ~LDataManager();

    // re-associates the points with cells (IndexUtilities::getCellIndex on the level)
    void endDataRedistribution(const LData<double>& X_data);

    // beginDataRedistribution + endDataRedistribution (LDataManager.cpp:1349-1960) for
    // the points this GPU holds: X is moved into the domain (periodic wrap, or clamp /
    // error when error_if_points_leave_domain), the nodes are renumbered so that every
    // patch's nodes are contiguous (patch order, cells in box order, Lagrangian index
    // in a cell), and X, the arrays in data and lag_index (if given) are permuted to
    // that order, then the points are binned again.  Returns the per-patch offsets
    // (npatches + 1 entries).  Throws with the reference's message for escaped points.
    std::vector<int> redistribute(LData<double>& X_data, const std::vector<LData<double>*>& data,
                                  LData<int>* lag_index = nullptr, bool error_if_points_leave_domain = false);

    // LDataManager::spread (LDataManager.cpp:552-667): f <- f + S F on the patch
    // interiors, patch ghost values unchanged.  f_data[p] is patch p's data.
    void spread(std::vector<std::shared_ptr<SideData<double>>>& f_data, const LData<double>& F_data,
                const LData<double>& X_data);
    // LDataManager::spread with ds_data (LDataManager.cpp:394-464): spreads
    // F*ds, F_ds[k][d] = F[k][d] * ds[k], formed while the points are put in
    // bin order (no separate pass over the points).
    void spread(std::vector<std::shared_ptr<SideData<double>>>& f_data, const LData<double>& F_data,
                const LData<double>& X_data, const LData<double>& ds_data);
    // LDataManager::interp (LDataManager.cpp:699-813): U <- J u for every point
    // in a patch interior; u's ghosts must be filled (fillPeriodicGhosts does it
    // for a fully periodic single-patch level).
    void interp(const std::vector<std::shared_ptr<SideData<double>>>& u_data, LData<double>& U_data,
                const LData<double>& X_data);
    void fillPeriodicGhosts(std::vector<std::shared_ptr<SideData<double>>>& u_data);
    // same-level ghost fill between the level's patches, periodic wrap included
    // (the RefineSchedule::fillData before interp, LDataManager.cpp:742-745,
    // without coarse-fine interpolation)
    void fillGhosts(std::vector<std::shared_ptr<SideData<double>>>& u_data);

private:
    struct Impl;
    std::unique_ptr<Impl> d_impl;
};

// IBStandardForceGen (src/IB/IBStandardForceGen.cpp:254-305): springs, beams
// and target points of one structure, given as IBStandardInitializer reads them
// (node indices 0-based; beam curvature and target positions [element][NDIM]).
// A spring's parameters are {stiffness, rest_length, extra_parameters...} and
// its force function index selects the law (IBStandardInitializer.cpp:397-434).
struct Spring
{
    int master, slave;
    double stiffness, rest_length;
    int force_fcn_idx = 0;
    std::vector<double> extra_parameters;
};
struct Beam
{
    int prev, master, next;
    double rigidity;
    std::array<double, 3> curvature;
};
struct TargetPoint
{
    int node;
    double stiffness, damping;
    std::array<double, 3> X0;
};

class IBStandardForceGen
{
public:
    IBStandardForceGen(int ndim, int num_nodes, const std::vector<Spring>& springs, const std::vector<Beam>& beams,
                       const std::vector<TargetPoint>& targets);
    ~IBStandardForceGen();
    IBStandardForceGen(const IBStandardForceGen&) = delete;
    IBStandardForceGen& operator=(const IBStandardForceGen&) = delete;
    // computeLagrangianForce: F += springs + beams + target points at X (U damps
    // the target points); IBMethod::computeLagrangianForce zeroes F first.
    void computeLagrangianForce(LData<double>& F, const LData<double>& X, const LData<double>& U);
    // registerSpringForceFunction (IBStandardForceGen.cpp:133-140) with a
    // built-in law (IBAMR_SPRING_HOOKEAN, _LINEAR_STRAIN, _POLYNOMIAL); index 0
    // starts as the Hookean default.  Springs take the laws of their indices at
    // initializeLevelData (:728-790), which computeLagrangianForce calls first
    // when a registration or a non-default index is pending.
    void registerSpringForceFunction(int force_fcn_index, int law);
    void initializeLevelData();

private:
    int d_ndim, d_n;
    ibamr_forces* d_forces = nullptr;
    std::vector<int> d_fcn_idx, d_par_off;
    std::vector<double> d_par;
    bool d_resolve = false;
};

// The IBStrategy hooks of IBMethod on one level.
class IBMethod
{
public:
    IBMethod(std::shared_ptr<LDataManager> l_data_manager) : d_ldm(std::move(l_data_manager)) {}
    // IBStrategy::spreadForce (IBStrategy.h:339)
    void spreadForce(std::vector<std::shared_ptr<SideData<double>>>& f_data, const LData<double>& F,
                     const LData<double>& X)
    {
        d_ldm->spread(f_data, F, X);
    }
    // IBStrategy::interpolateVelocity (IBStrategy.h:276)
    void interpolateVelocity(std::vector<std::shared_ptr<SideData<double>>>& u_data, LData<double>& U,
                             const LData<double>& X, bool fill_ghosts = true)
    {
        if (fill_ghosts) d_ldm->fillGhosts(u_data);
        d_ldm->interp(u_data, U, X);
    }
    // IBMethod::forwardEulerStep (IBMethod.cpp:713-739): X_new = X + dt U
    static void forwardEulerStep(LData<double>& X_new, const LData<double>& X, const LData<double>& U, double dt);

private:
    std::shared_ptr<LDataManager> d_ldm;
};

} // namespace ibtk_amd
