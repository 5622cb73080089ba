/*
 * ibamr_amd.h -- C ABI of the MI355X (gfx950) Eulerian-Lagrangian coupling
 * library: regularized-delta spreading and interpolation (IB_4, IB_6,
 * BSPLINE_3) on cell- and side-centred patch data.
 *
 * All pointers passed to the plan entry points are device pointers; the
 * caller owns them.  Plain C types only: streams are passed as void*
 * (a hipStream_t, NULL = default stream).  Every entry point returns an
 * ibamr_status; ibamr_last_error() describes the last failure of the calling
 * thread.  No exceptions cross this boundary.
 *
 * Data layouts follow the reference exactly, so SAMRAI/PETSc arrays can be
 * handed over without repacking:
 *   - Eulerian data: SAMRAI ArrayData over the ghost box, column-major
 *     (x fastest), depth slowest.  Side data: one array per axis over
 *     toSideBox(ghost box, axis) (upper index + 1 along the axis).
 *   - Lagrangian data: LData local form, AoS: X[n][ndim] (always double),
 *     F/U[n][depth] (side data: depth = ndim).
 *   - FP64 plans use double grids and F/U; FP32C plans use float grids and
 *     F/U (spread sums: FP32 partial sums of at most 64 points folded into
 *     FP64 totals; compensated interpolation sums); positions and stencil
 *     selection stay FP64 in both.
 */
#ifndef IBAMR_AMD_H
#define IBAMR_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ibamr_status
{
    IBAMR_OK = 0,
    IBAMR_ERR_BAD_KERNEL = -1,      /* unknown kernel name (LEInteractor.cpp:1923-2016 TBOX_ERROR) */
    IBAMR_ERR_GHOST_WIDTH = -2,     /* ghost width below getMinimumGhostWidth (LEInteractor.cpp:4491-4498) */
    IBAMR_ERR_DEPTH = -3,           /* side data needs Q_depth == NDIM, q depth 1 (LEInteractor.cpp:2197-2201) */
    IBAMR_ERR_INVALID = -4,         /* invalid descriptor or argument */
    IBAMR_ERR_OUT_OF_DOMAIN = -5,   /* a point's cell lies outside the level domain */
    IBAMR_ERR_CAPACITY = -6,        /* more points than the plan was created for */
    IBAMR_ERR_DEVICE = -7,          /* HIP runtime error */
    IBAMR_ERR_NOT_BINNED = -8       /* spread/interp before ibamr_plan_bin */
} ibamr_status;

enum
{
    IBAMR_CELL = 0,
    IBAMR_SIDE = 1
};
enum
{
    IBAMR_FP64 = 0,  /* plain FP64, results identical to the reference arithmetic */
    IBAMR_FP32C = 1  /* FP32 grid/force data, FP32 weights; spread partial sums in FP32 folded into
                        FP64 totals, interpolation sums compensated; positions and stencil
                        selection FP64 */
};
enum
{
    IBAMR_SPREAD_ACCUMULATE = 0, /* u += S F over the whole ghost box (LEInteractor::spread) */
    IBAMR_SPREAD_OVERWRITE = 1,  /* u = S F over the whole ghost box (zero + spread) */
    IBAMR_SPREAD_ACCUMULATE_INTERIOR = 2 /* u += S F on the patch interiors, ghost values unchanged
                                            (LDataManager::spread, LDataManager.cpp:589-594, 661-663) */
};

/* Kernel registry, LEInteractor::isKnownKernel / getStencilSize /
 * getMinimumGhostWidth (LEInteractor.cpp:2035-2114): the reference's 22 names
 * with their stencil sizes (USER_DEFINED reports 0: no user kernel is
 * registered).  "IB_4", "IB_6", "BSPLINE_3".."BSPLINE_6" and
 * "PIECEWISE_LINEAR" have GPU spread/interp; plans for the other names fail
 * with IBAMR_ERR_BAD_KERNEL.  Unknown names: 0 from
 * ibamr_is_known_kernel, IBAMR_ERR_BAD_KERNEL from the other two. */
int ibamr_is_known_kernel(const char* kernel_fcn);
int ibamr_get_stencil_size(const char* kernel_fcn);
int ibamr_get_minimum_ghost_width(const char* kernel_fcn);

/* One patch of a level: SAMRAI Box (cell indices, level index space) and
 * CartesianPatchGeometry x_lower / x_upper. */
typedef struct ibamr_patch_geom
{
    int ilower[3];
    int iupper[3];
    double x_lower[3];
    double x_upper[3];
} ibamr_patch_geom;

/* A patch level (LDataManager::spread/interp loop over one level,
 * LDataManager.cpp:552-667 and 699-813).  The domain box and x range feed the
 * level form of IndexUtilities::getCellIndex (IndexUtilities-inl.h:226-243);
 * periodic_shift is CartesianGridGeometry::getPeriodicShift(ratio) (0 = not
 * periodic along that axis). */
typedef struct ibamr_level_geom
{
    int ndim;
    double dx[3];
    int domain_lower[3];
    int domain_upper[3];
    double domain_x_lower[3];
    double domain_x_upper[3];
    int periodic_shift[3];
    int npatches;
    const ibamr_patch_geom* patches;
} ibamr_level_geom;

typedef struct ibamr_plan_config
{
    const char* kernel;   /* "IB_4" | "IB_6" | "BSPLINE_3" | "BSPLINE_4" | "BSPLINE_5" | "BSPLINE_6" | "PIECEWISE_LINEAR" */
    int centering;        /* IBAMR_CELL | IBAMR_SIDE */
    int precision;        /* IBAMR_FP64 | IBAMR_FP32C */
    int depth;            /* cell data depth; side data: must be 1 */
    int data_gcw[3];      /* ghost cell width of the Eulerian arrays */
    int idx_gcw;          /* Lagrangian index ghost width (LDataManager ghost width); 0 = kernel minimum */
    int max_points;       /* capacity of ibamr_plan_bin */
    int drift_cells;      /* how far (cells) points may move between bin and spread */
    int tile;             /* spread tile edge in cells (0 = default: 8 in 3-D, 16 in 2-D) */
} ibamr_plan_config;

typedef struct ibamr_plan ibamr_plan;

/* Creates a plan for one level: validates the kernel and ghost widths, builds
 * the (patch, tile) work list and allocates device scratch for max_points. */
int ibamr_plan_create(ibamr_plan** plan, const ibamr_level_geom* level, const ibamr_plan_config* config);

/* A plan for LEInteractor's raw-array form on one patch
 * (LEInteractor::interpolate/spread with double* Q_data, X_data and a box,
 * LEInteractor.cpp:2805-2860, 3950-4009; point selection by
 * LEInteractor::buildLocalIndices, LEInteractor.cpp:6089-6124): no periodic
 * images; a point is taken when its cell, by IndexUtilities::getCellIndex on
 * the patch box and x range, lies in box_lower..box_upper (any box: the patch
 * box, a grown box, a sub-box).  Points not taken are left out silently by
 * ibamr_plan_bin_async (ibamr_plan_check and ibamr_plan_bin report them as
 * IBAMR_ERR_OUT_OF_DOMAIN); interp leaves their U untouched and spread skips
 * them.  Stencils are clipped to the data ghost box as in the reference. */
int ibamr_plan_create_patch(ibamr_plan** plan, int ndim, const double* dx, const ibamr_patch_geom* patch,
                            const int* box_lower, const int* box_upper, const ibamr_plan_config* config);
int ibamr_plan_destroy(ibamr_plan* plan);

/* Binds Eulerian arrays to a slot (0..7): arrays[p*narr + a], narr = ndim for
 * side data (a = axis) and 1 for cell data.  Copies the pointer table to the
 * device (synchronous; do it once per regrid, not per step). */
int ibamr_plan_bind_data(ibamr_plan* plan, int slot, void* const* arrays);

/* Point binning: cell of every point by IndexUtilities::getCellIndex (level
 * form, bit-exact FP64), then a stable sort by spread bin.  This is the GPU
 * counterpart of the LNodeSetData association made at regrid
 * (LDataManager.cpp:1564-1651).  ibamr_plan_bin synchronizes the stream and
 * reports points outside the domain; ibamr_plan_bin_async does not (use
 * ibamr_plan_check afterwards), so it can be captured in a HIP graph. */
int ibamr_plan_bin(ibamr_plan* plan, const double* X, int npoints, void* stream);
int ibamr_plan_bin_async(ibamr_plan* plan, const double* X, int npoints, void* stream);
int ibamr_plan_check(ibamr_plan* plan, void* stream);

/* The cell index of every binned point, as the binning computed it
 * (IndexUtilities::getCellIndex, IndexUtilities-inl.h:52-72; level form
 * :226-243 with the level's domain box and x range, or the patch box and x
 * range for ibamr_plan_create_patch plans): cells (device, [n][ndim] int, the
 * caller's point order) receives it for each point of the last binning, inside
 * the domain or not; INT_MIN where the index does not fit in an int (NaN or
 * far away positions).  Kernel only (graph-safe). */
int ibamr_plan_get_cell_indices(const ibamr_plan* plan, int* cells, void* stream);

/* LDataManager::spread for one level (LDataManager.cpp:623-658): for every
 * patch, LEInteractor::spread over the Lagrangian index ghost box, with
 * periodic images, into the arrays bound to `slot`.  Deterministic, with no
 * floating-point atomics, so results are bitwise reproducible: FP64 sums every
 * grid value's contributions in one fixed order (bin, image, point); FP32C
 * sums the contributions of each chunk of at most 64 points in FP32 and folds
 * the chunk sums into FP64 totals, in an order fixed by the bins, the points'
 * order and the tile shape (the same inputs and shape give the same bits).  F: [n][ndim] (side) or [n][depth] (cell); X: current positions.
 * Modes: IBAMR_SPREAD_ACCUMULATE adds each contribution to the existing value
 * (the raw LEInteractor::spread); IBAMR_SPREAD_ACCUMULATE_INTERIOR sums the
 * contributions from zero and then adds the existing value (LDataManager::
 * spread's zero, spread, add, LDataManager.cpp:594, 661-663). */
int ibamr_plan_spread(ibamr_plan* plan, int slot, const void* F, const double* X, int mode, void* stream);

/* LDataManager::spread with a per-point weight ds (LDataManager.cpp:394-464):
 * spreads F_ds[k][c] = F[k][c] * ds[k], the product formed in the plan's
 * force precision (double for FP64, as the reference forms it; float for
 * FP32C) while the points are put in bin order, so it costs no pass of its
 * own.  ds: [n] of the force's type, the caller's point order; NULL spreads F
 * itself (ibamr_plan_spread). */
int ibamr_plan_spread_ds(ibamr_plan* plan, int slot, const void* F, const double* X, const void* ds, int mode,
                         void* stream);

/* LDataManager::interp for one level (LDataManager.cpp:748-802): each point
 * is interpolated from the patch whose interior holds its cell; ghost values
 * must be filled beforehand (see ibamr_plan_fill_periodic_ghosts).  U is
 * overwritten for those points.  X may be NULL when the positions are those of
 * the previous ibamr_plan_spread / ibamr_plan_interp call on this plan (same
 * binning): the positions that call put in bin order are reused (a step that
 * spreads and interpolates at one configuration, e.g. the midpoint of
 * IBExplicitHierarchyIntegrator). */
int ibamr_plan_interp(ibamr_plan* plan, int slot, void* U, const double* X, void* stream);

/* Fills the ghost cells of the arrays bound to `slot` from the periodic
 * interior for a single-patch level that spans a fully periodic domain (the
 * RefineSchedule::fillData step before interp, LDataManager.cpp:742-745). */
int ibamr_plan_fill_periodic_ghosts(ibamr_plan* plan, int slot, void* stream);

/* Same-level ghost fill for a level of patches (the RefineSchedule::fillData
 * before interp, LDataManager.cpp:742-745, between patches of one level): each
 * ghost value whose index -- wrapped across periodic directions -- lies in the
 * interior (side box for side data) of a patch of the level is copied from it.
 * Ghost values beyond a non-periodic domain face or at coarse-fine boundaries
 * are left unchanged.  Unlike ibamr_plan_fill_periodic_ghosts, interior values
 * (including the face N of a periodic side axis) are never written.  Kernel
 * only (graph-safe). */
int ibamr_plan_fill_ghosts(ibamr_plan* plan, int slot, void* stream);

/* Redistribution of the Lagrangian points over the level's patches, as
 * LDataManager::beginDataRedistribution / endDataRedistribution do at a regrid
 * (src: ibtk/src/lagrangian/LDataManager.cpp:1349-1521 and :1523-1960,
 * computeNodeDistribution :2858-2990), for the points held on this GPU:
 *  - X (device, [n][ndim], in/out) is moved into the domain: in a periodic
 *    direction X += L while X < x_lower and X -= L while X >= x_upper
 *    (:1403-1407); in other directions the point is an error when
 *    error_if_points_leave is set (:1409-1415) and otherwise clamped to
 *    [x_lower, x_upper - L sqrt(eps)] (:1417-1420);
 *  - periodic_offset (device, [n][ndim], may be NULL) receives
 *    round((X_old - X_new) / dx) per coordinate (:1424-1428);
 *  - every point's cell (IndexUtilities::getCellIndex) must lie in a patch of
 *    the level;
 *  - new_order (device, [n]) receives the new local order: new_order[k] is the
 *    old position of the point that becomes k.  Points are grouped by patch in
 *    level order, so each patch's interior points are contiguous
 *    (:1709-1712), patch p holding [patch_offsets[p], patch_offsets[p+1])
 *    (patch_offsets: device, [npatches+1]); inside a patch by cell in box order
 *    (x fastest); inside a cell by Lagrangian index (lag_index, device [n], may
 *    be NULL = the point's position), as the LNodeSet sort does (:1498-1503).
 * Then move every LData array to the new order with ibamr_permute_rows and bin
 * the permuted positions again (the plan's binning is void after this call).
 * Synchronizes the stream (the error checks read device counters).  Errors:
 * IBAMR_ERR_OUT_OF_DOMAIN (escaped, non-finite or farther than 64 periodic
 * lengths, or in no patch), IBAMR_ERR_INVALID (negative lag_index, plan made by
 * ibamr_plan_create_patch, or 2^31 cells or more). */
int ibamr_plan_redistribute(ibamr_plan* plan, double* X, const int* lag_index, int npoints, int error_if_points_leave,
                            int* new_order, int* patch_offsets, int* periodic_offset, void* stream);

/* Finite-element point source (FEDataManager::spread, ibtk/src/lagrangian/
 * FEDataManager.cpp:1169-1204 with sum_weighted_elem_solution :825-857): for
 * every element e and quadrature point q,
 *   X_qp[e*nqp+q][d]     = sum_k X_nodes[elem_nodes[e][k]][d] * phi[k][q]
 *   F_JxW_qp[e*nqp+q][i] = (sum_k F_nodes[elem_nodes[e][k]][i] * phi[k][q]) * JxW[e][q]
 * (sums in basis order from 0.0, as the reference loop).  phi ([nbasis][nqp]:
 * the quadrature rule's shape function table) and JxW ([nelem][nqp]) come from
 * the caller's FE library; JxW NULL means weights 1 (the positions-only pass
 * with weights_are_unity); F_nodes NULL computes X_qp only.  All arrays are
 * device arrays; F_qp is double (IBAMR_FP64) or float (IBAMR_FP32C, rounded
 * once).  Spreading the nelem*nqp points with a level plan then gives what the
 * per-patch spread after zeroExteriorValues (:2738-2783) plus the ghost
 * accumulation gives.  nvars <= 8.  X_nodes / F_nodes hold nnodes rows; an
 * element node index outside [0, nnodes) is never read and turns that
 * element's outputs into NaN.  Kernel only (graph-safe). */
int ibamr_fe_quadrature_points(int ndim, int nvars, int nelem, int nbasis, int nqp, int nnodes, const int* elem_nodes,
                               const double* phi, const double* JxW, const double* X_nodes, const double* F_nodes,
                               double* X_qp, void* F_qp, int f_precision, void* stream);

/* FE interpolation, the weighted part (FEDataManager::interpWeighted,
 * FEDataManager.cpp:2009-2061 with integrate_elem_rhs :1569-1606): after the
 * velocity has been interpolated to the quadrature points (a level plan's
 * interp of the points from ibamr_fe_quadrature_points), every element e forms
 *   rhs[k][i] = sum_q U_qp[e*nqp+q][i] * (phi[k][q] * JxW[e][q])
 * and adds it to node elem_nodes[e][k]; F_nodes [nnodes][nvars] (device,
 * double) = (IBAMR_SPREAD_OVERWRITE) or += (IBAMR_SPREAD_ACCUMULATE) the sum,
 * elements added in element order (deterministic, no atomics).  The L2
 * projection solve that follows in the reference (a mass-matrix solve) is the
 * caller's.  The mesh object holds the node -> (element, basis) lists, built on
 * the host from the host connectivity; node indices are checked there. */
typedef struct ibamr_fe_mesh ibamr_fe_mesh;
int ibamr_fe_mesh_create(ibamr_fe_mesh** mesh, int nelem, int nbasis, int nnodes, const int* elem_nodes_host);
int ibamr_fe_integrate_rhs(const ibamr_fe_mesh* mesh, int nvars, int nqp, const double* phi, const double* JxW,
                           const void* U_qp, int u_precision, double* F_nodes, int mode, void* stream);
int ibamr_fe_mesh_destroy(ibamr_fe_mesh* mesh);

/* A quadrature rule per element.  The reference keys each element's rule
 * separately (FEDataManager.cpp:1153-1165 spread, :1951-1963 interp); with
 * adaptive quadrature the order follows the deformed element, getQuadratureKey
 * (ibtk/src/utilities/libmesh_utilities.cpp:187-225):
 *   hmax  = get_max_edge_length (ibtk/include/ibtk/libmesh_utilities.h:1373-1511):
 *           TET4, HEX8: longest vertex-vertex edge; TET10, HEX27: longest edge as
 *           the sum of its two segments through the mid-edge node; other types:
 *           longest vertex pair on one edge (HEX20: the hexahedron's edges)
 *   npts  = int(ceil(point_density * hmax / dx_min)), at least 2 (first-order
 *           elements) or 3 (second-order)
 *   order = min(2 npts - 1, 43) for IBAMR_FE_QGAUSS, npts for IBAMR_FE_QGRID.
 * ibamr_fe_adaptive_quadrature_orders writes each element's order (device
 * int[nelem]; -1 for an element with a node index outside [0, nnodes)); kernel
 * only.  The caller's FE library gives the shape function table of each rule
 * (libMesh builds the rule from the key, as in the reference); the rule-per-
 * element forms of the point source and the weighted right-hand side then take
 * rule r = elem_rule[e] with rule_nqp[r] points, its table at phi +
 * rule_phi_off[r] ([nbasis][rule_nqp[r]]), element e's points and JxW values
 * from qp_off[e] (qp_off [nelem + 1], qp_off[nelem] = nqp_total), all device
 * arrays.  Arithmetic per point and per node as ibamr_fe_quadrature_points /
 * ibamr_fe_integrate_rhs; an element whose rule index or point count does not
 * fit gives NaN.  Kernels only. */
enum
{
    IBAMR_FE_EDGE2 = 0, IBAMR_FE_EDGE3, IBAMR_FE_TRI3, IBAMR_FE_TRI6, IBAMR_FE_QUAD4, IBAMR_FE_QUAD8,
    IBAMR_FE_QUAD9, IBAMR_FE_TET4, IBAMR_FE_TET10, IBAMR_FE_HEX8, IBAMR_FE_HEX20, IBAMR_FE_HEX27
};
enum
{
    IBAMR_FE_QGAUSS = 0,
    IBAMR_FE_QGRID = 1
};
int ibamr_fe_adaptive_quadrature_orders(int elem_type, int ndim, int nelem, int nbasis, int nnodes,
                                        const int* elem_nodes, const double* X_nodes, int quad_type,
                                        double point_density, double dx_min, int* order, void* stream);
int ibamr_fe_quadrature_points_rules(int ndim, int nvars, int nelem, int nbasis, int nnodes, const int* elem_nodes,
                                     int nrules, const int* elem_rule, const int* rule_nqp, const long* rule_phi_off,
                                     const double* phi, const long* qp_off, long nqp_total, const double* JxW,
                                     const double* X_nodes, const double* F_nodes, double* X_qp, void* F_qp,
                                     int f_precision, void* stream);
int ibamr_fe_integrate_rhs_rules(const ibamr_fe_mesh* mesh, int nvars, int nrules, const int* elem_rule,
                                 const int* rule_nqp, const long* rule_phi_off, const double* phi, const long* qp_off,
                                 const double* JxW, const void* U_qp, int u_precision, double* F_nodes, int mode,
                                 void* stream);

/* Nodal quadrature (spread_spec / interp_spec use_nodal_quadrature,
 * FEDataManager.cpp:996-1075 spread, :1761-1905 interp): the mesh nodes are the
 * points.  ibamr_fe_nodal_mass forms the diagonal mass of a rule,
 * FEProjector::buildDiagonalL2MassMatrix (FEProjector.cpp:587-672): M[v] = sum
 * over v's elements in element order of sum_q (phi[k][q] * phi[k][q]) * JxW[e][q]
 * (phi [nbasis][nqp] and JxW [nelem][nqp] of the nodal rule, build_nodal_qrule).
 * ibamr_fe_nodal_weight forms out[n][i] = (IBAMR_SPREAD_OVERWRITE) or +=
 * (IBAMR_SPREAD_ACCUMULATE) V[n][i] * M[n] over nnodes rows of nvars: the
 * spread's F (x) dX (:1014), whose rows are spread from X_nodes by a level plan,
 * and the interpolation's U * dX added into the right-hand side (:1883-1899)
 * after a level plan's interp to X_nodes.  V and out are double (IBAMR_FP64) or
 * float (IBAMR_FP32C).  Kernels only. */
int ibamr_fe_nodal_mass(const ibamr_fe_mesh* mesh, int nqp, const double* phi, const double* JxW, double* M_diag,
                        void* stream);
int ibamr_fe_nodal_weight(int nnodes, int nvars, const void* V, int v_precision, const double* M_diag, void* out,
                          int out_precision, int mode, void* stream);

/* Physical boundary operators after the spread / before the interpolation on
 * non-periodic domains: CartCellRobinPhysBdryOp / CartSideRobinPhysBdryOp
 * with LINEAR interpolation (ibtk/src/boundary/physical_boundary/
 * CartCellRobinPhysBdryOp.cpp:415-474, CartSideRobinPhysBdryOp.cpp:553-620
 * and :484-545, kernels cartphysbdryop{2,3}d.f.m4).  The boxes are what
 * SAMRAI's PatchGeometry gives each patch: codim 1: the boundary box (the cell
 * layer just outside the face; its tangential extent is the coefficient box);
 * codim 2 and 3: getBoundaryFillBox(box, patch box, ghost width).  A patch's
 * boxes are applied in the order given (SAMRAI's: codim 1, 2, 3 lists, each by
 * location).  abg [nboxes][arrays][3] holds the Robin coefficients a, b, g of
 * codim-1 boxes per array (cell data: depth component d, the reference's
 * bc_coefs[d]; side data: axis, bc_coefs[axis]), uniform over the box.
 * LDataManager::spread calls accumulateFromPhysicalBoundaryData after the
 * ghost accumulation (LDataManager.cpp:647-660): ibamr_plan_accumulate_
 * physical_boundary after ibamr_plan_spread.  The patch's out-of-domain ghost
 * values the fold reads are, after a level spread, the contributions of all
 * points that reach them; SAMRAI's accumulation leaves there only the patch's
 * own points' contributions, so on levels with several patches the two agree
 * where no other patch's points reach a patch's out-of-domain ghost cells.
 * set_physical_boundary validates the boxes (inside the ghost box) and copies
 * them to the device; the other two enqueue one kernel (graph-safe). */
typedef struct
{
    int patch;      /* patch index in the level */
    int codim;      /* 1 .. ndim */
    int location;   /* SAMRAI location index of the codimension */
    int lower[3];
    int upper[3];
} ibamr_bdry_box;
int ibamr_plan_set_physical_boundary(ibamr_plan* plan, const ibamr_bdry_box* boxes, int nboxes, const double* abg);
/* The same boxes with coefficients that vary over each codim-1 box, as a
 * RobinBcCoefStrategy::setBcCoefs fills acoef/bcoef/gcoef over the coefficient
 * box (CartCellRobinPhysBdryOp.cpp:518-531; side data: the normal axis at
 * CartSideRobinPhysBdryOp.cpp:664-677, the tangential axes over the coefficient
 * box extended by one cell along the array's axis, :913-947).  coefs holds, for
 * each codim-1 box in the given order and each array j (as for abg), a[n],
 * b[n], g[n] over the box's n columns, the lower tangential axis fastest (the
 * Fortran's acoef(j,k) order); n is the product of the box's tangential extents,
 * plus one along j for a side array tangential to the box.  ncoefs: the number
 * of doubles in coefs, checked against what the boxes need.  The Dirichlet test
 * |b| < 1e-12 is taken per face, as the reference's kernels take it. */
int ibamr_plan_set_physical_boundary_coefs(ibamr_plan* plan, const ibamr_bdry_box* boxes, int nboxes,
                                           const double* coefs, long ncoefs);
int ibamr_plan_accumulate_physical_boundary(ibamr_plan* plan, int slot, void* stream);
int ibamr_plan_fill_physical_boundary(ibamr_plan* plan, int slot, void* stream);

/* The VecScatter of the LData arrays to the new order (LDataManager.cpp:1769-1840, 1900-1915):
 * row k of dst = row order[k] of src, rows of depth elements of elem_bytes
 * (4 or 8) bytes; dst and src distinct device arrays.  Kernel only (graph-safe). */
int ibamr_permute_rows(void* dst, const void* src, const int* order, int n, int depth, int elem_bytes, void* stream);

/* IBMethod::forwardEulerStep (IBMethod.cpp:713-739): X_new = X + dt U over
 * n values (n = ndim * points), on the stream; X_new may be X.  U is double
 * (u_precision IBAMR_FP64) or float (IBAMR_FP32C).  Kernel only: graph-safe. */
int ibamr_forward_euler_step(double* X_new, const double* X, const void* U, int u_precision, double dt, long n,
                             void* stream);

/* The same update over npoints points of ndim coordinates, each new position
 * then moved back into the domain along the axes with periodic[d] != 0 by one
 * period (x < x_lower: x + L; then x >= x_upper: x - L; L = x_upper - x_lower),
 * as LDataManager::beginDataRedistribution wraps positions at a regrid
 * (LDataManager.cpp:1403-1407), so points stay in the binned domain between
 * regrids.  x_lower, x_upper, periodic: host arrays of ndim.  Kernel only:
 * graph-safe. */
int ibamr_forward_euler_step_periodic(double* X_new, const double* X, const void* U, int u_precision, double dt,
                                      long npoints, int ndim, const double* x_lower, const double* x_upper,
                                      const int* periodic, void* stream);

/* Lagrangian force generation, IBStandardForceGen::computeLagrangianForce
 * (src/IB/IBStandardForceGen.cpp:254-305): springs (created Hookean,
 * default_spring_force, IBSpringForceFunctions.h:104: tension kappa (R - rest);
 * other laws below), beams (three-point
 * bending with rigidity and reference curvature, :1038-1147) and target points
 * (kappa (X0 - X) - eta U, :1202-1299).  Element arrays are host arrays, node
 * indices 0-based into the n points; curvature and X0 are [element][ndim].
 * Per node the forces are summed in the reference's order (springs, beams,
 * targets; element order), so FP64 results are bitwise those of the reference
 * loops. */
typedef struct ibamr_forces ibamr_forces;
int ibamr_forces_create(ibamr_forces** forces, int ndim, int npoints, int nsprings, const int* spring_master,
                        const int* spring_slave, const double* spring_kappa, const double* spring_rest, int nbeams,
                        const int* beam_master, const int* beam_next, const int* beam_prev,
                        const double* beam_rigidity, const double* beam_curvature, int ntargets,
                        const int* target_node, const double* target_kappa, const double* target_eta,
                        const double* target_X0);
/* F (device, [n][ndim], double for IBAMR_FP64 or float for IBAMR_FP32C) += (mode
 * IBAMR_SPREAD_ACCUMULATE) or = (IBAMR_SPREAD_OVERWRITE) the force at positions X
 * (device, double); U (device, [n][ndim], double for u_precision IBAMR_FP64 or
 * float for IBAMR_FP32C, read as double; may be NULL = at rest) feeds the target
 * point damping.  Forces are summed in double either way.  Kernels only
 * (graph-safe). */
int ibamr_forces_compute(ibamr_forces* forces, void* F, int f_precision, const double* X, const void* U,
                         int u_precision, int mode, void* stream);
int ibamr_forces_destroy(ibamr_forces* forces);

/* Spring force functions (IBStandardForceGen::registerSpringForceFunction,
 * src/IB/IBStandardForceGen.cpp:133-140, and the per-spring force_fcn_idx and
 * parameter lists of IBSpringForceSpec / the .spring files,
 * IBStandardInitializer.cpp:397-434).  The reference registers host function
 * pointers T(R, params); on the GPU the registry maps a force function index
 * to one of these built-in laws (params p[0 .. n), n >= 2):
 *   IBAMR_SPRING_HOOKEAN        p[0] (R - p[1])  (default_spring_force,
 *                               IBSpringForceFunctions.h:104; index 0 holds it
 *                               from creation, as in the reference's constructor)
 *   IBAMR_SPRING_LINEAR_STRAIN  p[0] (R / p[1] - 1)  (the linear_spring_force
 *                               of examples/IB/explicit/ex3/example.cpp:58-63)
 *   IBAMR_SPRING_POLYNOMIAL     x (p[0] + x (p[2] + ... + x p[n-1])), x = R - p[1]
 *                               (no reference counterpart: a nonlinear law in
 *                               the extension; with n == 2 it is the Hookean law)
 * ibamr_forces_set_spring_parameters gives spring k its force function index
 * and parameters params[param_offsets[k] .. param_offsets[k+1]) (offsets:
 * [nsprings + 1], from 0) and resolves every index through the registry then,
 * as the reference resolves them when level data is initialised
 * (IBStandardForceGen.cpp:728-790): register first, then set.  Unregistered
 * indices and fewer than two parameters are errors.  Forces use
 * F = (T(R) / R) D per spring as before, bitwise in FP64 with the reference
 * loop evaluating the same law. */
enum
{
    IBAMR_SPRING_HOOKEAN = 0,
    IBAMR_SPRING_LINEAR_STRAIN = 1,
    IBAMR_SPRING_POLYNOMIAL = 2
};
int ibamr_forces_register_spring_function(ibamr_forces* forces, int force_fcn_index, int law);
int ibamr_forces_set_spring_parameters(ibamr_forces* forces, const int* force_fcn_idx, const int* param_offsets,
                                       const double* params);

/* Kernel timing for measurement.  With timing on, HIP events are recorded on
 * the launch stream right around each spread tile kernel and each
 * interpolation kernel (the point permutation before them is not included).
 * ibamr_plan_kernel_times waits for those events and returns the mean
 * duration (ms) and the number of launches since timing was switched on or
 * last read, then starts a new count.  Keep timing off for work captured in
 * a HIP graph. */
int ibamr_plan_set_timing(ibamr_plan* plan, int on);
int ibamr_plan_kernel_times(ibamr_plan* plan, double* spread_ms, int* nspread, double* interp_ms, int* ninterp);

/* Plan facts for reports: number of spread tiles, tile edge, bins. */
int ibamr_plan_info(const ibamr_plan* plan, int* ntiles, int* tile, int* nbins);

/* Spread tile shape.  The spread's work is a list of tiles (boxes of grid
 * values of one array of one patch), one workgroup each; a plan builds the
 * work list of every shape built for its form when it is created and keeps
 * them until it is destroyed, so changing the shape only selects another list
 * (a HIP graph captured earlier keeps launching the shape it captured, which
 * stays valid).  Shapes (x, y, z extents; the first is the default):
 *   2-D: 16 x 16.
 *   3-D FP64: 8 x 8 grid lines times 32 or 16 z values held in registers, or
 *     16 x 4 grid lines times 32.
 *   3-D FP32C: 8 x 8 grid lines times 16, 8 or 24 z values, or 16 x 4 times 16;
 *     and 8 x 8 grid lines times a z segment of 64, 128 or 256 values swept
 *     through a 16-value register window, layer of bins by layer (each point
 *     applied once per 8 x 8 column; built when the kernel has precomputed
 *     stencils, bin + 2 reach <= 16, i.e. drift_cells 0 for IB_6).
 * FP64 results do not depend on the shape (each value is summed in an order
 * fixed by the bins); FP32C results are reproducible for a given shape and
 * differ between shapes within the FP32C tolerance (a shape groups the points
 * into other 64-point chunks).  No
 * counterpart in the reference, whose CPU loops have no tiles: these tune this
 * implementation's work decomposition to the caller's points and hardware.
 *   set_tile_shape / get_tile_shape: shape[3].
 *   set_tile_height / get_tile_height: the first shape with that z extent / the
 *     z extent of the shape in use.
 *   autotune: for each shape one untimed spread of F at X into the arrays
 *     bound to `slot` (IBAMR_SPREAD_OVERWRITE: their values are replaced), then
 *     3 rounds over all shapes of `reps` timed spreads each, and keeps the
 *     shape with the fastest round.  shapes[3*i .. 3*i+2] and ms[i] (capacity
 *     IBAMR_MAX_TILE_SHAPES) receive each candidate and its best round's mean
 *     time per spread call, *ncand their number; synchronises the stream.
 *     Points must be binned. */
#define IBAMR_MAX_TILE_SHAPES 8
int ibamr_plan_set_tile_shape(ibamr_plan* plan, const int* shape);
int ibamr_plan_get_tile_shape(const ibamr_plan* plan, int* shape);
int ibamr_plan_set_tile_height(ibamr_plan* plan, int height);
int ibamr_plan_get_tile_height(const ibamr_plan* plan, int* height);
int ibamr_plan_autotune(ibamr_plan* plan, int slot, const void* F, const double* X, int reps, int* shapes,
                        float* ms, int* ncand, void* stream);

const char* ibamr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
