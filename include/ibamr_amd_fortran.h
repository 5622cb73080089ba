/*
 * ibamr_amd_fortran.h -- link-level drop-in for IBTK's Fortran kernels.
 *
 * IBTK::LEInteractor calls its spreading/interpolation kernels through the
 * extern "C" block of ibtk/src/lagrangian/LEInteractor.cpp:237-1516 (symbols
 * mangled by FortranCInterface, IBTK_FC_FUNC_, CMakeLists.txt:66-83; gfortran
 * gives lower case plus a trailing underscore).  This library exports the same
 * symbols for IB_4, IB_6, BSPLINE_3..6 and PIECEWISE_LINEAR in 2-D and 3-D, with the same argument
 * lists (every scalar by reference: C++ `const int&` == `const int*`) and the
 * same semantics:
 *   - interp overwrites V(0:depth-1, indices(l)) for every listed entry and
 *     leaves the rest of V alone;
 *   - spread accumulates into u and never zeroes it;
 *   - stencil entries outside the ghost box are dropped;
 *   - x_upper is not used.
 * Arrays are host pointers in the Fortran layouts:
 *   u(ilower0-nugc0:iupper0+nugc0, ..., 0:depth-1), X(0:NDIM-1, 0:*),
 *   V(0:depth-1, 0:*), Xshift(0:NDIM-1, 0:nindices-1).
 * Each call copies its data to the GPU, runs there and copies results back.
 * Results are bitwise those of the reference loops: interp per entry in the
 * same operation order; spread sums the contributions to every grid value in
 * list order (contributions are generated per entry, stably sorted by grid
 * index, then summed sequentially).
 *
 * A rebuilt IBTK links this library instead of the Fortran objects of these
 * kernels; on a failure the functions print a message to stderr and abort,
 * as the reference's TBOX_ERROR does.
 */
#ifndef IBAMR_AMD_FORTRAN_H
#define IBAMR_AMD_FORTRAN_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 3-D: LEInteractor.cpp:541-600 (IB_4), and the IB_6 / BSPLINE_3 blocks of the same form */
#define IBAMR_S0_INTERP3D_ARGS                                                                                  \
    const double *dx, const double *x_lower, const double *x_upper, const int *depth, const int *ilower0,      \
        const int *iupper0, const int *ilower1, const int *iupper1, const int *ilower2, const int *iupper2,    \
        const int *nugc0, const int *nugc1, const int *nugc2, const double *u, const int *indices,             \
        const double *Xshift, const int *nindices, const double *X, double *V
#define IBAMR_S0_SPREAD3D_ARGS                                                                                  \
    const double *dx, const double *x_lower, const double *x_upper, const int *depth, const int *indices,      \
        const double *Xshift, const int *nindices, const double *X, const double *V, const int *ilower0,       \
        const int *iupper0, const int *ilower1, const int *iupper1, const int *ilower2, const int *iupper2,    \
        const int *nugc0, const int *nugc1, const int *nugc2, double *u
/* ---- 2-D: the NDIM == 2 branches of the same declarations */
#define IBAMR_S0_INTERP2D_ARGS                                                                                  \
    const double *dx, const double *x_lower, const double *x_upper, const int *depth, const int *ilower0,      \
        const int *iupper0, const int *ilower1, const int *iupper1, const int *nugc0, const int *nugc1,        \
        const double *u, const int *indices, const double *Xshift, const int *nindices, const double *X,       \
        double *V
#define IBAMR_S0_SPREAD2D_ARGS                                                                                  \
    const double *dx, const double *x_lower, const double *x_upper, const int *depth, const int *indices,      \
        const double *Xshift, const int *nindices, const double *X, const double *V, const int *ilower0,       \
        const int *iupper0, const int *ilower1, const int *iupper1, const int *nugc0, const int *nugc1,        \
        double *u

/* lagrangian_interaction3d.f.m4:1203 / :1344 */
void lagrangian_ib_4_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_ib_4_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
/* lagrangian_interaction3d.f.m4:2178 / :2384 */
void lagrangian_ib_6_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_ib_6_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
/* lagrangian_interaction3d.f.m4:2591 / :2703 */
void lagrangian_bspline_3_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_bspline_3_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
/* lagrangian_interaction2d.f.m4:1152 / :1273 */
void lagrangian_ib_4_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_ib_4_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
/* lagrangian_interaction2d.f.m4:1971 / :2136 */
void lagrangian_ib_6_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_ib_6_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
/* lagrangian_interaction2d.f.m4:2300 / :2410 */
void lagrangian_bspline_3_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_bspline_3_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
/* lagrangian_interaction3d.f.m4:493 / :617 */
void lagrangian_piecewise_linear_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_piecewise_linear_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
/* lagrangian_interaction2d.f.m4:469 / :587 */
void lagrangian_piecewise_linear_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_piecewise_linear_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
/* lagrangian_interaction3d.f.m4:2814 / :2933 (BSPLINE_4), :3050 / :3162
 * (BSPLINE_5), :3273 / :3392 (BSPLINE_6) */
void lagrangian_bspline_4_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_bspline_4_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
void lagrangian_bspline_5_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_bspline_5_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
void lagrangian_bspline_6_interp3d_(IBAMR_S0_INTERP3D_ARGS);
void lagrangian_bspline_6_spread3d_(IBAMR_S0_SPREAD3D_ARGS);
/* lagrangian_interaction2d.f.m4:2519 / :2634, :2748 / :2858, :2967 / :3082 */
void lagrangian_bspline_4_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_bspline_4_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
void lagrangian_bspline_5_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_bspline_5_spread2d_(IBAMR_S0_SPREAD2D_ARGS);
void lagrangian_bspline_6_interp2d_(IBAMR_S0_INTERP2D_ARGS);
void lagrangian_bspline_6_spread2d_(IBAMR_S0_SPREAD2D_ARGS);

/* Status-returning forms of the same operations (no abort), for callers that
 * want an error code: kernel is "IB_4" | "IB_6" | "BSPLINE_3".."BSPLINE_6" | "PIECEWISE_LINEAR", ndim 2 or 3,
 * ilower/iupper/nugc arrays of length ndim.  Returns an ibamr_status. */
int ibamr_s0_interp(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                    const int* ilower, const int* iupper, const int* nugc, const double* u, const int* indices,
                    const double* Xshift, int nindices, const double* X, double* V);
int ibamr_s0_spread(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                    const int* indices, const double* Xshift, int nindices, const double* X, const double* V,
                    const int* ilower, const int* iupper, const int* nugc, double* u);

/* Allocation-free, device-pointer form of ibamr_s0_interp.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  dx, x_lower, ilower, iupper and nugc are host arrays as above;
 * u, indices, Xshift, X and V are device pointers; X and V hold npoints points
 * (entries whose index lies outside [0, npoints) are skipped).  The kernels'
 * scratch -- the last list entry of every point, which decides which entry
 * sets V(:, s) when an index repeats -- is a file-scope device array of
 * IBAMR_S0_NOALLOC_MAX_POINTS ints that each GPU the library runs on holds
 * until the process ends; a larger npoints fails with IBAMR_ERR_CAPACITY.
 * Enqueues kernels on `stream` only: no allocation, copy or synchronisation.
 * Every call on one GPU uses that GPU's array, so calls that overlap on one
 * GPU (other streams, other threads) race. */
#define IBAMR_S0_NOALLOC_MAX_POINTS (1 << 22)
int ibamr_s0_interp_noalloc(const char* kernel, int ndim, const double* dx, const double* x_lower, int depth,
                            const int* ilower, const int* iupper, const int* nugc, const double* u,
                            const int* indices, const double* Xshift, int nindices, const double* X, int npoints,
                            double* V, void* stream);

#ifdef __cplusplus
}
#endif
#endif
