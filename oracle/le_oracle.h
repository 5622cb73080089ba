/*
 * le_oracle.h -- CPU oracle for the Eulerian-Lagrangian spread/interpolate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of IBAMR's FP64 CPU
 * algorithm (the m4-generated Fortran kernels behind IBTK::LEInteractor) used
 * as the checker for the HIP library.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every function cites the reference file:line it follows.  Arithmetic is
 * evaluated in the same operation order as the Fortran source (left to right,
 * no FMA contraction: build with -ffp-contract=off), NINT is round-half-away
 * (lround), r**n is expanded as gfortran expands small integer powers
 * (r2=r*r, r3=r2*r, r4=r2*r2, r6=r3*r3).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   - interpolate, IB_4/IB_6/BSPLINE_3, cell data: golden vectors from
 *     tests/interpolate/interpolate_01_{2d,3d}.<kernel>.output (regenerated
 *     point sets with mt19937(42)).
 *   - spread geometry on side data (x_lower - dx/2, side box, depth layout):
 *     tests/IBTK/ghost_accumulation_01_3d.side.spread.a.output (PIECEWISE_LINEAR).
 *   - 1-D weights: scalar delta functions (lagrangian_delta.f.m4:174-196,
 *     243-260), partition of unity, moment and sum-of-squares identities.
 *   - spread values: adjoint identity with interpolate, brute-force O(N*G)
 *     evaluation, order independence to round-off.
 *   - side-centred IB_4/IB_6/BSPLINE_3: no reference fixture exists; pinned by
 *     the identities above and by the structural identity "side axis a = cell
 *     kernel on the half-shifted grid".
 */
#ifndef LE_ORACLE_H
#define LE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum orc_kernel
{
    ORC_PIECEWISE_LINEAR = 0,
    ORC_IB_4 = 1,
    ORC_IB_6 = 2,
    ORC_BSPLINE_3 = 3,
    ORC_BSPLINE_4 = 4,
    ORC_BSPLINE_5 = 5,
    ORC_BSPLINE_6 = 6
};

/* Registry (LEInteractor.cpp:2035-2114). Returns -1 for unknown names. */
int orc_kernel_from_name(const char* name);
int orc_stencil_size(int kernel);      /* getStencilSize: PWL 2, IB_4 4, IB_6 6, BSPLINE_3 4, _4 4, _5 6, _6 6 */
int orc_min_ghost_width(int kernel);   /* floor(stencil/2)+1 */

/* Scalar delta functions (lagrangian_delta.f.m4:174-196 and :243-260). */
double orc_ib_4_delta(double r);
/* lagrangian_delta.f.m4:268-354 (BSPLINE_4, _5, _6) and any kernel's scalar delta. */
double orc_bspline_4_delta(double x);
double orc_bspline_5_delta(double x);
double orc_bspline_6_delta(double x);
double orc_delta(int kernel, double r);
double orc_bspline_3_delta(double x);

/* 1-D weights of one coordinate exactly as the kernel computes them.
 * X_o_dx_num is (X+Xshift)-x_lower; returns the first stencil index relative
 * to ilower (unclipped) and writes the stencil weights into w[] (length
 * orc_kernel_width(kernel)). */
int orc_kernel_width(int kernel);      /* PWL 2, IB_4 4, IB_6 6, BSPLINE_3 3, _4 4, _5 5, _6 6 */
int orc_weights_1d(int kernel, double Xs, double x_lower, double dx, double* w);
/* The same with the unshifted X as well: BSPLINE_4 and BSPLINE_6 choose their
 * stencil by comparing the unshifted X(d,s) with the centre of the cell of
 * X(d,s)+Xshift (lagrangian_interaction3d.f.m4:2891, :3469; 2d.f.m4:2708,
 * :3156), which for a periodic image forces the choice by the shift's sign. */
int orc_weights_1d_u(int kernel, double Xs, double Xu, double x_lower, double dx, double* w);

/* Fortran-equivalent kernels.  Arrays follow the Fortran declarations:
 *   u(ilower0-nugc0:iupper0+nugc0, ..., 0:depth-1)  column-major, depth slowest
 *   X(0:ndim-1, 0:*), V(0:depth-1, 0:*), Xshift(0:ndim-1, 0:nindices-1)
 * interp:  lagrangian_<k>_interp{2,3}d  (overwrites V(:,indices(l)))
 * spread:  lagrangian_<k>_spread{2,3}d  (accumulates into u) */
void orc_interp(int kernel, int ndim, const double* dx, const double* x_lower, int depth,
                const int* ilower, const int* iupper, const int* nugc, const double* u,
                const int* indices, const double* Xshift, int nindices, const double* X, double* V);
void orc_spread(int kernel, int ndim, const double* dx, const double* x_lower, int depth,
                const int* indices, const double* Xshift, int nindices, const double* X, const double* V,
                const int* ilower, const int* iupper, const int* nugc, double* u);

/* IndexUtilities::getCellIndex (IndexUtilities-inl.h:52-72). */
void orc_get_cell_index(int ndim, const double* X, const double* x_lower, const double* x_upper,
                        const double* dx, const int* ilower, const int* iupper, int* idx);

/* Batch form of orc_get_cell_index over n points (X is [n][ndim], idx [n][ndim]). */
void orc_get_cell_indices(int ndim, const double* X, int n, const double* x_lower, const double* x_upper,
                          const double* dx, const int* ilower, const int* iupper, int* idx);

/* Raw-array index list (LEInteractor.cpp:6088-6126): points whose cell lies in
 * box [box_lo, box_hi]; returns the count, fills indices in point order. */
int orc_build_local_indices(int ndim, const int* box_lo, const int* box_hi, const double* x_lower,
                            const double* x_upper, const double* dx, const int* ilower, const int* iupper,
                            const double* X, int npoints, int* indices);

/* Side-centred wrappers (LEInteractor.cpp:2402-2489 interp, 3627-3714 spread).
 * q_axis[a] points to the ghosted array of axis a over toSideBox(box, a).
 * Q is AoS [n][ndim]; the kernel runs once per axis with depth 1. */
void orc_interp_side(int kernel, int ndim, const double* dx, const double* x_lower, const int* ilower,
                     const int* iupper, const int* gcw, const double* const* q_axis, const int* indices,
                     const double* Xshift, int nindices, const double* X, double* Q);
void orc_spread_side(int kernel, int ndim, const double* dx, const double* x_lower, const int* ilower,
                     const int* iupper, const int* gcw, double* const* q_axis, const int* indices,
                     const double* Xshift, int nindices, const double* X, const double* Q);

/* Lagrangian index entries of one patch as LIndexSetData::cacheLocalIndices
 * builds them (LIndexSetData.cpp:55-150) for a level whose points have cells
 * cell[ndim*p+d] (level index space): every point whose cell, or whose periodic
 * image cell cell+o*N (o in {-1,0,1} per periodic axis, N = periodic_shift),
 * lies in the patch ghost box [ilower-idx_gcw, iupper+idx_gcw].  Xshift =
 * (double)(o*N)*dx.  interior_only restricts to o=0 and cells in the patch box.
 *
 * Order: the reference iterates its node sets in SAMRAI IndexData order, which
 * the reference sources do not fix; the oracle uses the canonical order
 * (bz,oz,by,oy,bx,ox,point) with b = floor((cell - domain_lower)/bin) the
 * physical bin of the point and o its image offset, which the GPU spread
 * reproduces exactly.  Returns the number of entries (<= max_entries, or the
 * required count if larger). */
int orc_build_patch_entries(int ndim, const int* cell, int npoints, const int* ilower, const int* iupper,
                            int idx_gcw, const int* periodic_shift, const int* domain_lower, const double* dx,
                            const int* bin, int interior_only, int max_entries, int* indices, double* Xshift);

/* Reorders a patch entry list (from orc_build_patch_entries, canonical bin
 * order) into the order in which the GPU 3-D spread sums the contributions to
 * one array: within each (bin, image offset) group the bin's points are cut
 * into consecutive blocks of `chunk` points (bin memory order = point index
 * order), and inside a block the entries are ordered by the first stencil
 * index along z for this array (x_lower is the array's lower corner, i.e. the
 * patch x_lower shifted by -dx/2 along the side axis), then by point.  Any
 * order gives the same sum to round-off; this one makes the FP64 GPU result
 * bitwise comparable.  `cell` holds the level cell of every point (npoints of
 * them: a bin's points are ranked by the cell's z plane inside the bin, then index); perm
 * receives the new order as positions into indices[]. */
void orc_order_entries_z(int kernel, const int* cell, int npoints, const int* indices, const double* Xshift,
                         int nentries, const double* X, const double* x_lower, const double* dx,
                         const int* domain_lower, const int* bin, int chunk, const long* ranks, int* perm);

/* The ranks orc_order_entries_z uses (ranks == NULL computes them per call):
 * rank_of[q] = position of point q among the points of its bin, ordered by the
 * cell's z plane inside the bin, then point index.  They depend only on the
 * cells and the bins, so a level's patches can share one computation. */
void orc_bin_ranks(const int* cell, int npoints, const int* domain_lower, const int* bin, long* rank_of);

/* Brute-force O(N*G) evaluation with the scalar delta functions: value of the
 * spread of point values V (depth 1) at one grid location, periodic images
 * included.  Independent of the kernel loops (used for IB_4 and BSPLINE_3). */
double orc_bruteforce_spread_at(int kernel, int ndim, const double* xg, const double* dx,
                                const double* period, const double* X, const double* V, int npoints);

/* IBStandardForceGen::computeLagrangianForce: F += spring, beam and target
 * point forces in the reference's scatter order (src/IB/IBStandardForceGen.cpp:254-305). */
/* (orc_lagrangian_force below.)  The same with a spring law (0 Hookean, 1 linear strain, 2 polynomial in the
 * extension) and a parameter list par[poff[k] .. poff[k+1]) per spring. */
void orc_lagrangian_force_laws(int ndim, int n, double* F, const double* X, const double* U, int ns, const int* sm,
                               const int* ss, const int* law, const int* poff, const double* par, int nb,
                               const int* bm, const int* bn, const int* bp, const double* bK, const double* bc,
                               int nt, const int* tn, const double* tK, const double* tE, const double* tX0);
/* LDataManager::beginDataRedistribution / computeNodeDistribution for one
 * level: wrap/clamp, periodic offsets, owner patch, new order by (patch, cell
 * in box order, Lagrangian index).  0 ok, 1 escaped, 2 in no patch, 3 lost. */
int orc_redistribute(int ndim, int n, double* X, const int* lag, const double* x_lower, const double* x_upper,
                     const double* dx, const int* dom_lo, const int* dom_hi, const int* periodic, int error_if_leave,
                     int max_wraps, int npatches, const int* plo, const int* phi, int* new_order,
                     int* patch_offsets, int* offset);

/* FEDataManager::spread's quadrature point source (FEDataManager.cpp:1169-1204). */
void orc_fe_quadrature_points(int ndim, int nvars, int nelem, int nbasis, int nqp, const int* elem_nodes,
                              const double* phi, const double* JxW, const double* X_nodes, const double* F_nodes,
                              double* X_qp, double* F_qp);

/* FEDataManager::interpWeighted's right-hand side (FEDataManager.cpp:2009-2061). */
void orc_fe_integrate_rhs(int nvars, int nelem, int nbasis, int nqp, const int* elem_nodes, const double* phi,
                          const double* JxW, const double* U_qp, double* F);

/* Nodal quadrature's diagonal mass (FEProjector.cpp:587-672), the adaptive
 * quadrature order per element (libmesh_utilities.cpp:187-225 with
 * get_max_edge_length, libmesh_utilities.h:1373-1511; element types 0..11 =
 * EDGE2 EDGE3 TRI3 TRI6 QUAD4 QUAD8 QUAD9 TET4 TET10 HEX8 HEX20 HEX27) and the
 * quadrature point source / right-hand side with a rule per element. */
void orc_fe_nodal_mass(int nelem, int nbasis, int nnodes, const int* elem_nodes, int nqp, const double* phi,
                       const double* JxW, double* M);
double orc_fe_max_edge_length(int type, int ndim, const int* n, const double* X);
void orc_fe_adaptive_orders(int type, int ndim, int nelem, int nbasis, const int* elem_nodes, const double* X,
                            int quad_type, double point_density, double dx_min, int* order);
void orc_fe_quadrature_points_rules(int ndim, int nvars, int nelem, int nbasis, const int* elem_nodes,
                                    const int* elem_rule, const int* rule_nqp, const long* rule_phi_off,
                                    const double* phi, const long* qp_off, const double* JxW,
                                    const double* X_nodes, const double* F_nodes, double* X_qp, double* F_qp);
void orc_fe_integrate_rhs_rules(int nvars, int nelem, int nbasis, const int* elem_nodes, const int* elem_rule,
                                const int* rule_nqp, const long* rule_phi_off, const double* phi, const long* qp_off,
                                const double* JxW, const double* U_qp, double* F);

/* CartCellRobinPhysBdryOp / CartSideRobinPhysBdryOp LINEAR ghost operators on one
 * patch: adjoint = accumulateFromPhysicalBoundaryData, else setPhysicalBoundaryConditions.
 * boxes: nboxes x {codim, location, lo[3], hi[3]}; abg: [nboxes][arrays][3], one
 * (a, b, g) per codim-1 box and array. */
void orc_phys_bdry(int ndim, int side, int depth, int gcw, const int* ilo, const int* ihi, const double* dx,
                   double** arrays, int nboxes, const int* boxes, const double* abg, int adjoint);
/* The same with the Robin coefficients varying over each codim-1 box, as
 * RobinBcCoefStrategy::setBcCoefs fills them on the coefficient box
 * (CartCellRobinPhysBdryOp.cpp:518-531, CartSideRobinPhysBdryOp.cpp:664-677, 913):
 * for every codim-1 box in order and every array j (cell data: component; side
 * data: axis), a[n], b[n], g[n] over the box's tangential cells (one more along
 * j for a side array tangential to the box), the lower tangential axis fastest
 * (the Fortran acoef(j,k) order).  orc_phys_bdry_ncoefs: the doubles coefs holds. */
long orc_phys_bdry_ncoefs(int ndim, int side, int depth, int nboxes, const int* boxes);
void orc_phys_bdry_coefs(int ndim, int side, int depth, int gcw, const int* ilo, const int* ihi, const double* dx,
                         double** arrays, int nboxes, const int* boxes, const double* coefs, int adjoint);

void orc_lagrangian_force(int ndim, int n, double* F, const double* X, const double* U, int ns, const int* sm,
                          const int* ss, const double* sk, const double* sr, int nb, const int* bm, const int* bn,
                          const int* bp, const double* bK, const double* bc, int nt, const int* tn, const double* tK,
                          const double* tE, const double* tX0);

#ifdef __cplusplus
}
#endif
#endif
