/*
 * le_oracle.c -- CPU oracle (test infrastructure only) restating the FP64
 * spread/interpolate kernels of IBAMR.  See le_oracle.h for scope, pinning and
 * the rules on who may use this file.  Build with -ffp-contract=off and no
 * fast-math so that every operation is rounded as written.
 *
 * Reference sources (paths relative to the IBAMR checkout):
 *   ibtk/src/lagrangian/fortran/lagrangian_interaction3d.f.m4
 *     PIECEWISE_LINEAR interp :493-606, spread :617-730
 *     IB_4      interp :1203-1334, spread :1344-1475
 *     IB_6      interp :2178-2375, spread :2384-2582
 *     BSPLINE_3 interp :2591-2693, spread :2703-2805, loop macros :24-86
 *   ibtk/src/lagrangian/fortran/lagrangian_interaction2d.f.m4
 *     PIECEWISE_LINEAR :469-700, IB_4 :1152-1392, IB_6 :1971-2299,
 *     BSPLINE_3 :2300-2518, loop macros :17-86
 *   ibtk/src/lagrangian/fortran/lagrangian_delta.f.m4:174-196, 243-260
 *   ibtk/src/lagrangian/LEInteractor.cpp (registry :2035-2114, side wrappers
 *     :2402-2489 and :3627-3714, raw index list :6088-6126)
 *   ibtk/include/ibtk/private/IndexUtilities-inl.h:52-72
 *   ibtk/src/lagrangian/LIndexSetData.cpp:55-150
 */
#include "le_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ registry */

int orc_kernel_from_name(const char* name)
{
    if (!name) return -1;
    if (strcmp(name, "PIECEWISE_LINEAR") == 0) return ORC_PIECEWISE_LINEAR;
    if (strcmp(name, "IB_4") == 0) return ORC_IB_4;
    if (strcmp(name, "IB_6") == 0) return ORC_IB_6;
    if (strcmp(name, "BSPLINE_3") == 0) return ORC_BSPLINE_3;
    if (strcmp(name, "BSPLINE_4") == 0) return ORC_BSPLINE_4;
    if (strcmp(name, "BSPLINE_5") == 0) return ORC_BSPLINE_5;
    if (strcmp(name, "BSPLINE_6") == 0) return ORC_BSPLINE_6;
    return -1;
}

/* LEInteractor::getStencilSize (LEInteractor.cpp:2035-2105): BSPLINE_3 reports 4. */
int orc_stencil_size(int kernel)
{
    switch (kernel)
    {
    case ORC_PIECEWISE_LINEAR: return 2;
    case ORC_IB_4: return 4;
    case ORC_IB_6: return 6;
    case ORC_BSPLINE_3: return 4;
    case ORC_BSPLINE_4: return 4;
    case ORC_BSPLINE_5: return 6;
    case ORC_BSPLINE_6: return 6;
    default: return -1;
    }
}

/* LEInteractor::getMinimumGhostWidth (LEInteractor.cpp:2108-2114). */
int orc_min_ghost_width(int kernel)
{
    const int s = orc_stencil_size(kernel);
    return s < 0 ? -1 : (int)floor(0.5 * s) + 1;
}

/* Number of weights per coordinate the loops actually use. */
int orc_kernel_width(int kernel)
{
    switch (kernel)
    {
    case ORC_PIECEWISE_LINEAR: return 2;
    case ORC_IB_4: return 4;
    case ORC_IB_6: return 6;
    case ORC_BSPLINE_3: return 3;
    case ORC_BSPLINE_4: return 4;
    case ORC_BSPLINE_5: return 5;
    case ORC_BSPLINE_6: return 6;
    default: return -1;
    }
}

/* ------------------------------------------------------ scalar delta functions */

/* lagrangian_delta.f.m4:174-196 */
double orc_ib_4_delta(double r)
{
    if (r < 0.0) r = -r;
    if (r < 1.0)
    {
        const double t2 = r * r;
        const double t6 = sqrt(-4.0 * t2 + 4.0 * r + 1.0);
        return -r / 4.0 + 3.0 / 8.0 + t6 / 8.0;
    }
    if (r < 2.0)
    {
        const double t2 = r * r;
        const double t6 = sqrt(12.0 * r - 7.0 - 4.0 * t2);
        return -r / 4.0 + 5.0 / 8.0 - t6 / 8.0;
    }
    return 0.0;
}

/* lagrangian_delta.f.m4:243-260 */
double orc_bspline_3_delta(double x)
{
    const double modx = fabs(x);
    const double r = modx + 1.5;
    const double r2 = r * r;
    if (modx <= 0.5) return 0.5 * (-2.0 * r2 + 6.0 * r - 3.0);
    if (modx <= 1.5) return 0.5 * (r2 - 6.0 * r + 9.0);
    return 0.0;
}

/* lagrangian_delta.f.m4:268-288 */
double orc_bspline_4_delta(double x)
{
    const double modx = fabs(x);
    const double r = modx + 2.0;
    const double r2 = r * r;
    const double r3 = r2 * r;
    if (modx <= 1.0) return (1.0 / 6.0) * (((3.0 * r3 - 24.0 * r2) + 60.0 * r) - 44.0);
    if (modx <= 2.0) return (1.0 / 6.0) * (((-r3 + 12.0 * r2) - 48.0 * r) + 64.0);
    return 0.0;
}

/* lagrangian_delta.f.m4:296-320 */
double orc_bspline_5_delta(double x)
{
    const double modx = fabs(x);
    const double r = modx + 2.5;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double r4 = r3 * r;
    if (modx <= 0.5) return (1.0 / 24.0) * ((((6.0 * r4 - 60.0 * r3) + 210.0 * r2) - 300.0 * r) + 155.0);
    if (modx <= 1.5) return (1.0 / 24.0) * ((((-4.0 * r4 + 60.0 * r3) - 330.0 * r2) + 780.0 * r) - 655.0);
    if (modx <= 2.5) return (1.0 / 24.0) * ((((r4 - 20.0 * r3) + 150.0 * r2) - 500.0 * r) + 625.0);
    return 0.0;
}

/* lagrangian_delta.f.m4:328-354 */
double orc_bspline_6_delta(double x)
{
    const double modx = fabs(x);
    const double r = modx + 3.0;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double r4 = r3 * r;
    const double r5 = r4 * r;
    if (modx <= 1.0)
        return (1.0 / 60.0) * (((((2193.0 - 3465.0 * r) + 2130.0 * r2) - 630.0 * r3) + 90.0 * r4) - 5.0 * r5);
    if (modx <= 2.0)
        return (1.0 / 120.0) * (((((-10974.0 + 12270.0 * r) - 5340.0 * r2) + 1140.0 * r3) - 120.0 * r4) + 5.0 * r5);
    if (modx <= 3.0)
        return (1.0 / 120.0) * (((((7776.0 - 6480.0 * r) + 2160.0 * r2) - 360.0 * r3) + 30.0 * r4) - r5);
    return 0.0;
}

double orc_delta(int kernel, double r)
{
    switch (kernel)
    {
    case ORC_IB_4: return orc_ib_4_delta(r);
    case ORC_BSPLINE_3: return orc_bspline_3_delta(r);
    case ORC_BSPLINE_4: return orc_bspline_4_delta(r);
    case ORC_BSPLINE_5: return orc_bspline_5_delta(r);
    case ORC_BSPLINE_6: return orc_bspline_6_delta(r);
    default: return 0.0;
    }
}

/* ------------------------------------------------------------- 1-D weights */

static int nint_(double x) { return (int)lround(x); } /* Fortran NINT: half away from zero */

/* IB_6 weights from r (lagrangian_interaction3d.f.m4:2243-2272).  K is the
 * PARAMETER of :2227, evaluated in double with one rounding per operation,
 * which is what compile-time folding produces. */
static void ib6_weights(double r, double* w)
{
    const double K = (59.0 / 60.0) * (1.0 - sqrt(1.0 - (3220.0 / 3481.0)));
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double r4 = r2 * r2;
    const double r6 = r3 * r3;
    const double alpha = 28.0;
    /* constant sub-expressions, folded as the Fortran compiler folds them */
    const double c_b1 = (22.0 / 3.0) - 7.0 * K;
    const double c_g1 = (((161.0 / 36.0) - (59.0 / 6.0) * K) + 5.0 * (K * K)) * (1.0 / 2.0);
    const double c_g2 = (-(109.0 / 24.0) + 5.0 * K) * (1.0 / 3.0);
    const double c_m2 = (1.0 / 12.0) * (3.0 * K - 1.0);
    const double c_m1 = (1.0 / 6.0) * (4.0 - 3.0 * K);
    const double beta = (((9.0 / 4.0) - (3.0 / 2.0) * (K + r2)) + c_b1 * r) - (7.0 / 3.0) * r3;
    const double gamma = (1.0 / 4.0) * ((c_g1 * r2 + c_g2 * r4) + (5.0 / 18.0) * r6);
    const double discr = beta * beta - (4.0 * alpha) * gamma;
    const double sgn = ((3.0 / 2.0) - K) >= 0.0 ? 1.0 : -1.0; /* sign(1.d0, 3/2-K) */
    const double pm3 = (-beta + sgn * sqrt(discr)) / (2.0 * alpha);
    const double pm2 = ((((-(3.0 * pm3)) - (1.0 / 16.0)) + (1.0 / 8.0) * (K + r2)) + c_m2 * r) + (1.0 / 12.0) * r3;
    const double pm1 = (((2.0 * pm3) + (1.0 / 4.0)) + c_m1 * r) - (1.0 / 6.0) * r3;
    const double p = ((2.0 * pm3) + (5.0 / 8.0)) - (1.0 / 4.0) * (K + r2);
    const double pp1 = (((-(3.0 * pm3)) + (1.0 / 4.0)) - c_m1 * r) + (1.0 / 6.0) * r3;
    const double pp2 = (((pm3 - (1.0 / 16.0)) + (1.0 / 8.0) * (K + r2)) - c_m2 * r) - (1.0 / 12.0) * r3;
    w[0] = pm3;
    w[1] = pm2;
    w[2] = pm1;
    w[3] = p;
    w[4] = pp1;
    w[5] = pp2;
}

/* Returns the first stencil index relative to ilower and fills w.  Xs is the
 * already-shifted coordinate X(d,s)+Xshift(d,l). */
int orc_weights_1d(int kernel, double Xs, double x_lower, double dx, double* w)
{
    return orc_weights_1d_u(kernel, Xs, Xs, x_lower, dx, w);
}

int orc_weights_1d_u(int kernel, double Xs, double Xu, double x_lower, double dx, double* w)
{
    switch (kernel)
    {
    case ORC_IB_4:
    {
        /* lagrangian_interaction3d.f.m4:1265-1273 */
        const double X_o_dx = (Xs - x_lower) / dx;
        const int n = nint_(X_o_dx);
        const int lo = n - 2;                              /* ic_lower - ilower */
        const double r = X_o_dx - ((double)(lo + 1) + 0.5);
        const double q = sqrt(1.0 + (4.0 * r) * (1.0 - r));
        w[0] = 0.125 * ((3.0 - 2.0 * r) - q);
        w[1] = 0.125 * ((3.0 - 2.0 * r) + q);
        w[2] = 0.125 * ((1.0 + 2.0 * r) + q);
        w[3] = 0.125 * ((1.0 + 2.0 * r) - q);
        return lo;
    }
    case ORC_IB_6:
    {
        /* lagrangian_interaction3d.f.m4:2238-2241 */
        const double X_o_dx = (Xs - x_lower) / dx;
        const int n = nint_(X_o_dx);
        const int lo = n - 3;
        const double r = (1.0 - X_o_dx) + ((double)(lo + 2) + 0.5);
        ib6_weights(r, w);
        return lo;
    }
    case ORC_BSPLINE_3:
    {
        /* lagrangian_interaction3d.f.m4:2655-2676 (weights depend only on ic,
         * so computing all three before clipping gives the same values) */
        const int c = (int)floor((Xs - x_lower) / dx);
        const int lo = c - 1;
        for (int k = 0; k < 3; ++k)
        {
            const double X_cell = x_lower + ((double)(lo + k) + 0.5) * dx;
            w[k] = orc_bspline_3_delta((Xs - X_cell) / dx);
        }
        return lo;
    }
    case ORC_BSPLINE_4:
    case ORC_BSPLINE_5:
    case ORC_BSPLINE_6:
    {
        /* lagrangian_interaction3d.f.m4:2882-2908 (BSPLINE_4), :3230-3248
         * (BSPLINE_5: centre -2..+2), :3460-3480 (BSPLINE_6); the even orders
         * pick the side by the unshifted X against the cell centre */
        const int W = orc_kernel_width(kernel);
        const int c = (int)floor((Xs - x_lower) / dx);
        const double X_c = x_lower + ((double)c + 0.5) * dx;
        int lo;
        if (kernel == ORC_BSPLINE_5)
            lo = c - 2;
        else
            lo = (Xu < X_c) ? c - W / 2 : c - W / 2 + 1;
        for (int k = 0; k < W; ++k)
        {
            const double X_cell = x_lower + ((double)(lo + k) + 0.5) * dx;
            w[k] = orc_delta(kernel, (Xs - X_cell) / dx);
        }
        return lo;
    }
    case ORC_PIECEWISE_LINEAR:
    {
        /* lagrangian_interaction3d.f.m4:680-699 */
        const int c = nint_((Xs - x_lower) / dx - 0.5);
        const double X_cell = x_lower + ((double)c + 0.5) * dx;
        if (Xs < X_cell)
        {
            w[0] = (X_cell - Xs) / dx;
            w[1] = 1.0 - w[0];
            return c - 1;
        }
        w[0] = 1.0 + (X_cell - Xs) / dx;
        w[1] = 1.0 - w[0];
        return c;
    }
    default: return 0;
    }
}

/* ------------------------------------------------------------ kernel loops */

static int is_ib_family(int kernel) { return kernel == ORC_IB_4 || kernel == ORC_IB_6; }

typedef struct
{
    int base[3];  /* absolute index of w[d][0] */
    int lo[3];    /* clipped loop bounds (absolute) */
    int hi[3];
    double w[3][8];
} stencil_t;

static void make_stencil(int kernel, int ndim, const double* Xs, const double* Xu, const double* dx,
                         const double* x_lower, const int* ilower, const int* iupper, const int* nugc, stencil_t* st)
{
    const int W = orc_kernel_width(kernel);
    for (int d = 0; d < ndim; ++d)
    {
        const int rel = orc_weights_1d_u(kernel, Xs[d], Xu[d], x_lower[d], dx[d], st->w[d]);
        const int b = ilower[d] + rel;
        const int glo = ilower[d] - nugc[d], ghi = iupper[d] + nugc[d];
        st->base[d] = b;
        st->lo[d] = b > glo ? b : glo;
        st->hi[d] = (b + W - 1) < ghi ? (b + W - 1) : ghi;
    }
}

void orc_interp(int kernel, int ndim, const double* dx, const double* x_lower, int depth, const int* ilower,
                const int* iupper, const int* nugc, const double* u, const int* indices, const double* Xshift,
                int nindices, const double* X, double* V)
{
    long n[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) n[d] = iupper[d] - ilower[d] + 1 + 2 * nugc[d];
    const long ncomp = n[0] * n[1] * n[2];
    const int ib = is_ib_family(kernel);
    for (int l = 0; l < nindices; ++l)
    {
        const int s = indices[l];
        double Xs[3];
        for (int d = 0; d < ndim; ++d) Xs[d] = X[ndim * s + d] + Xshift[ndim * l + d];
        stencil_t st;
        make_stencil(kernel, ndim, Xs, X + ndim * s, dx, x_lower, ilower, iupper, nugc, &st);
        for (int c = 0; c < depth; ++c)
        {
            const double* uc = u + c * ncomp;
            double v = 0.0;
            if (ndim == 3)
            {
                for (int i2 = st.lo[2]; i2 <= st.hi[2]; ++i2)
                    for (int i1 = st.lo[1]; i1 <= st.hi[1]; ++i1)
                        for (int i0 = st.lo[0]; i0 <= st.hi[0]; ++i0)
                        {
                            const double w0 = st.w[0][i0 - st.base[0]];
                            const double w1 = st.w[1][i1 - st.base[1]];
                            const double w2 = st.w[2][i2 - st.base[2]];
                            const double uv = uc[(i0 - (ilower[0] - nugc[0])) +
                                                 n[0] * ((i1 - (ilower[1] - nugc[1])) +
                                                         n[1] * (long)(i2 - (ilower[2] - nugc[2])))];
                            /* IB: w(i0,i1,i2) = w0(i0)*(w1(i1)*w2(i2))  (:1288-1296)
                             * BS: V + w0*w1*w2*u, left to right      (:24-37) */
                            if (ib)
                                v = v + (w0 * (w1 * w2)) * uv;
                            else
                                v = v + ((w0 * w1) * w2) * uv;
                        }
            }
            else
            {
                for (int i1 = st.lo[1]; i1 <= st.hi[1]; ++i1)
                    for (int i0 = st.lo[0]; i0 <= st.hi[0]; ++i0)
                    {
                        const double w0 = st.w[0][i0 - st.base[0]];
                        const double w1 = st.w[1][i1 - st.base[1]];
                        const double uv = uc[(i0 - (ilower[0] - nugc[0])) + n[0] * (long)(i1 - (ilower[1] - nugc[1]))];
                        if (ib)
                            v = v + (w0 * w1) * uv;
                        else
                            v = v + (w0 * w1) * uv;
                    }
            }
            V[depth * s + c] = v;
        }
    }
}

void orc_spread(int kernel, int ndim, const double* dx, const double* x_lower, int depth, const int* indices,
                const double* Xshift, int nindices, const double* X, const double* V, const int* ilower,
                const int* iupper, const int* nugc, double* u)
{
    long n[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) n[d] = iupper[d] - ilower[d] + 1 + 2 * nugc[d];
    const long ncomp = n[0] * n[1] * n[2];
    const int ib = is_ib_family(kernel);
    const double vol = ndim == 3 ? (dx[0] * dx[1]) * dx[2] : dx[0] * dx[1];
    for (int l = 0; l < nindices; ++l)
    {
        const int s = indices[l];
        double Xs[3];
        for (int d = 0; d < ndim; ++d) Xs[d] = X[ndim * s + d] + Xshift[ndim * l + d];
        stencil_t st;
        make_stencil(kernel, ndim, Xs, X + ndim * s, dx, x_lower, ilower, iupper, nugc, &st);
        for (int c = 0; c < depth; ++c)
        {
            double* uc = u + c * ncomp;
            const double Vv = V[depth * s + c];
            if (ndim == 3)
            {
                for (int i2 = st.lo[2]; i2 <= st.hi[2]; ++i2)
                    for (int i1 = st.lo[1]; i1 <= st.hi[1]; ++i1)
                        for (int i0 = st.lo[0]; i0 <= st.hi[0]; ++i0)
                        {
                            const double w0 = st.w[0][i0 - st.base[0]];
                            const double w1 = st.w[1][i1 - st.base[1]];
                            const double w2 = st.w[2][i2 - st.base[2]];
                            double* up = &uc[(i0 - (ilower[0] - nugc[0])) +
                                             n[0] * ((i1 - (ilower[1] - nugc[1])) +
                                                     n[1] * (long)(i2 - (ilower[2] - nugc[2])))];
                            /* IB: wz = w2/(dx0*dx1*dx2); w = w0*(w1*wz); u += w*V  (:1432-1465)
                             * BS: u + (w0*w1*w2*V)/(dx0*dx1*dx2)                 (:51-64) */
                            if (ib)
                                *up = *up + (w0 * (w1 * (w2 / vol))) * Vv;
                            else
                                *up = *up + (((w0 * w1) * w2) * Vv) / vol;
                        }
            }
            else
            {
                for (int i1 = st.lo[1]; i1 <= st.hi[1]; ++i1)
                    for (int i0 = st.lo[0]; i0 <= st.hi[0]; ++i0)
                    {
                        const double w0 = st.w[0][i0 - st.base[0]];
                        const double w1 = st.w[1][i1 - st.base[1]];
                        double* up = &uc[(i0 - (ilower[0] - nugc[0])) + n[0] * (long)(i1 - (ilower[1] - nugc[1]))];
                        if (ib)
                            *up = *up + (w0 * (w1 / vol)) * Vv;
                        else
                            *up = *up + ((w0 * w1) * Vv) / vol;
                    }
            }
        }
    }
}

/* ------------------------------------------------------------------ binning */

void orc_get_cell_index(int ndim, const double* X, const double* x_lower, const double* x_upper, const double* dx,
                        const int* ilower, const int* iupper, int* idx)
{
    for (int d = 0; d < ndim; ++d)
    {
        const double dX_lower = X[d] - x_lower[d], dX_upper = X[d] - x_upper[d];
        if (fabs(dX_lower) <= fabs(dX_upper))
            idx[d] = (int)((double)ilower[d] + floor(dX_lower / dx[d]));
        else
            idx[d] = (int)((double)iupper[d] + floor(dX_upper / dx[d]) + 1.0);
    }
}

void orc_get_cell_indices(int ndim, const double* X, int n, const double* x_lower, const double* x_upper,
                          const double* dx, const int* ilower, const int* iupper, int* idx)
{
    for (int k = 0; k < n; ++k)
        orc_get_cell_index(ndim, &X[(long)ndim * k], x_lower, x_upper, dx, ilower, iupper, &idx[(long)ndim * k]);
}

int orc_build_local_indices(int ndim, const int* box_lo, const int* box_hi, const double* x_lower,
                            const double* x_upper, const double* dx, const int* ilower, const int* iupper,
                            const double* X, int npoints, int* indices)
{
    int n = 0;
    for (int k = 0; k < npoints; ++k)
    {
        int idx[3];
        orc_get_cell_index(ndim, &X[ndim * k], x_lower, x_upper, dx, ilower, iupper, idx);
        int inside = 1;
        for (int d = 0; d < ndim; ++d) inside = inside && idx[d] >= box_lo[d] && idx[d] <= box_hi[d];
        if (inside) indices[n++] = k;
    }
    return n;
}

/* ------------------------------------------------------------ side wrappers */

static void side_geometry(int ndim, int axis, const double* dx, const double* x_lower, const int* iupper,
                          double* x_lower_axis, int* iupper_axis)
{
    for (int d = 0; d < ndim; ++d)
    {
        x_lower_axis[d] = x_lower[d];
        iupper_axis[d] = iupper[d];
    }
    x_lower_axis[axis] -= 0.5 * dx[axis];   /* LEInteractor.cpp:2465 */
    iupper_axis[axis] += 1;                 /* SideGeometry::toSideBox */
}

static int max_index(const int* indices, int n)
{
    int m = -1;
    for (int l = 0; l < n; ++l)
        if (indices[l] > m) m = indices[l];
    return m;
}

void orc_interp_side(int kernel, int ndim, const double* dx, const double* x_lower, const int* ilower,
                     const int* iupper, const int* gcw, const double* const* q_axis, const int* indices,
                     const double* Xshift, int nindices, const double* X, double* Q)
{
    if (nindices == 0) return;
    const int sz = max_index(indices, nindices) + 1;
    double* buf = (double*)calloc((size_t)sz, sizeof(double));
    for (int a = 0; a < ndim; ++a)
    {
        double xl[3];
        int iu[3];
        side_geometry(ndim, a, dx, x_lower, iupper, xl, iu);
        orc_interp(kernel, ndim, dx, xl, 1, ilower, iu, gcw, q_axis[a], indices, Xshift, nindices, X, buf);
        for (int l = 0; l < nindices; ++l) Q[ndim * indices[l] + a] = buf[indices[l]];
    }
    free(buf);
}

void orc_spread_side(int kernel, int ndim, const double* dx, const double* x_lower, const int* ilower,
                     const int* iupper, const int* gcw, double* const* q_axis, const int* indices,
                     const double* Xshift, int nindices, const double* X, const double* Q)
{
    if (nindices == 0) return;
    const int sz = max_index(indices, nindices) + 1;
    double* buf = (double*)calloc((size_t)sz, sizeof(double));
    for (int a = 0; a < ndim; ++a)
    {
        double xl[3];
        int iu[3];
        side_geometry(ndim, a, dx, x_lower, iupper, xl, iu);
        for (int l = 0; l < nindices; ++l) buf[indices[l]] = Q[ndim * indices[l] + a];
        orc_spread(kernel, ndim, dx, xl, 1, indices, Xshift, nindices, X, buf, ilower, iu, gcw, q_axis[a]);
    }
    free(buf);
}

/* ---------------------------------------------------------- patch entries */

typedef struct
{
    int key[8];
    int point;
    int off[3];
} entry_t;

static int cmp_entry(const void* a, const void* b)
{
    const entry_t* x = (const entry_t*)a;
    const entry_t* y = (const entry_t*)b;
    for (int k = 0; k < 8; ++k)
    {
        if (x->key[k] < y->key[k]) return -1;
        if (x->key[k] > y->key[k]) return 1;
    }
    return 0;
}

static int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

/* The cell's plane inside its bin along the last axis (z in 3-D, y in 2-D):
 * the library's binning sorts a bin's points by it, then by index. */
static int plane_in_bin(int ndim, const int* c, const int* domain_lower, const int* bin)
{
    const int d = ndim - 1;
    return c[d] - domain_lower[d] - floor_div(c[d] - domain_lower[d], bin[d]) * bin[d];
}

int orc_build_patch_entries(int ndim, const int* cell, int npoints, const int* ilower, const int* iupper,
                            int idx_gcw, const int* periodic_shift, const int* domain_lower, const double* dx,
                            const int* bin, int interior_only, int max_entries, int* indices, double* Xshift)
{
    int cap = npoints * (ndim == 3 ? 8 : 4) + 1;
    entry_t* e = (entry_t*)malloc(sizeof(entry_t) * (size_t)cap);
    int ne = 0;
    for (int p = 0; p < npoints; ++p)
    {
        const int* c = &cell[ndim * p];
        int olo[3] = {0, 0, 0}, ohi[3] = {0, 0, 0};
        if (!interior_only)
            for (int d = 0; d < ndim; ++d)
                if (periodic_shift[d] > 0)
                {
                    olo[d] = -1;
                    ohi[d] = 1;
                }
        for (int oz = olo[2]; oz <= ohi[2]; ++oz)
            for (int oy = olo[1]; oy <= ohi[1]; ++oy)
                for (int ox = olo[0]; ox <= ohi[0]; ++ox)
                {
                    const int o[3] = {ox, oy, oz};
                    int ok = 1;
                    for (int d = 0; d < ndim; ++d)
                    {
                        const int ci = c[d] + o[d] * periodic_shift[d];
                        const int g = interior_only ? 0 : idx_gcw;
                        ok = ok && ci >= ilower[d] - g && ci <= iupper[d] + g;
                    }
                    if (!ok) continue;
                    entry_t* en = &e[ne++];
                    int b[3] = {0, 0, 0};
                    for (int d = 0; d < ndim; ++d) b[d] = floor_div(c[d] - domain_lower[d], bin[d]);
                    if (ndim == 3)
                    {
                        en->key[0] = b[2];
                        en->key[1] = o[2];
                        en->key[2] = b[1];
                        en->key[3] = o[1];
                        en->key[4] = b[0];
                        en->key[5] = o[0];
                    }
                    else
                    {
                        en->key[0] = 0;
                        en->key[1] = 0;
                        en->key[2] = b[1];
                        en->key[3] = o[1];
                        en->key[4] = b[0];
                        en->key[5] = o[0];
                    }
                    en->key[6] = plane_in_bin(ndim, c, domain_lower, bin);
                    en->key[7] = p;
                    en->point = p;
                    for (int d = 0; d < 3; ++d) en->off[d] = o[d];
                }
    }
    qsort(e, (size_t)ne, sizeof(entry_t), cmp_entry);
    if (ne <= max_entries)
    {
        for (int k = 0; k < ne; ++k)
        {
            indices[k] = e[k].point;
            for (int d = 0; d < ndim; ++d)
                Xshift[ndim * k + d] = (double)(e[k].off[d] * periodic_shift[d]) * dx[d]; /* LIndexSetData.cpp:116 */
        }
    }
    free(e);
    return ne;
}

/* --------------------------------------------------- z-bucketed entry order */

typedef struct
{
    long key[4];
    int pos;
} zkey_t;

static int cmp_zkey(const void* a, const void* b)
{
    const zkey_t* x = (const zkey_t*)a;
    const zkey_t* y = (const zkey_t*)b;
    for (int k = 0; k < 4; ++k)
    {
        if (x->key[k] < y->key[k]) return -1;
        if (x->key[k] > y->key[k]) return 1;
    }
    return 0;
}

typedef struct
{
    long bin;
    int l, q;
} rkey_t;

static int cmp_rkey_bin(const void* a, const void* b)
{
    const rkey_t* x = (const rkey_t*)a;
    const rkey_t* y = (const rkey_t*)b;
    if (x->bin != y->bin) return x->bin < y->bin ? -1 : 1;
    if (x->l != y->l) return x->l < y->l ? -1 : 1;
    return x->q < y->q ? -1 : (x->q > y->q);
}

void orc_bin_ranks(const int* cell, int npoints, const int* domain_lower, const int* bin, long* rank_of)
{
    /* rank of every point among the points of its bin in the library's bin order
     * (cell's z plane inside the bin, then index; the GPU cuts that whole list into chunks) */
    const int maxp = npoints;
    int bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    for (int q = 0; q < maxp; ++q)
        for (int d = 0; d < 3; ++d)
        {
            const int bq = floor_div(cell[3 * q + d] - domain_lower[d], bin[d]);
            if (q == 0 || bq < bmin[d]) bmin[d] = bq;
            if (q == 0 || bq > bmax[d]) bmax[d] = bq;
        }
    const long nbx = bmax[0] - bmin[0] + 1, nby = bmax[1] - bmin[1] + 1, nbz = bmax[2] - bmin[2] + 1;
    rkey_t* rk = (rkey_t*)malloc(sizeof(rkey_t) * (size_t)(maxp > 0 ? maxp : 1));
    for (int q = 0; q < maxp; ++q)
    {
        long id = 0;
        for (int d = 2; d >= 0; --d)
        {
            const long nbd = d == 0 ? nbx : (d == 1 ? nby : nbz);
            id = id * nbd + (floor_div(cell[3 * q + d] - domain_lower[d], bin[d]) - bmin[d]);
        }
        rk[q].bin = id;
        rk[q].l = plane_in_bin(3, cell + 3 * q, domain_lower, bin);
        rk[q].q = q;
    }
    qsort(rk, (size_t)maxp, sizeof(rkey_t), cmp_rkey_bin);
    for (int i = 0, run = 0; i < maxp; ++i)
    {
        run = (i > 0 && rk[i].bin == rk[i - 1].bin) ? run + 1 : 0;
        rank_of[rk[i].q] = run;
    }
    free(rk);
}

void orc_order_entries_z(int kernel, const int* cell, int npoints, const int* indices, const double* Xshift,
                         int nentries, const double* X, const double* x_lower, const double* dx,
                         const int* domain_lower, const int* bin, int chunk, const long* ranks, int* perm)
{
    zkey_t* z = (zkey_t*)malloc(sizeof(zkey_t) * (size_t)(nentries > 0 ? nentries : 1));
    long* own = NULL;
    if (!ranks)
    {
        own = (long*)malloc(sizeof(long) * (size_t)(npoints > 0 ? npoints : 1));
        orc_bin_ranks(cell, npoints, domain_lower, bin, own);
        ranks = own;
    }
    long group = -1;
    int prev_b[3] = {0, 0, 0}, prev_o[3] = {0, 0, 0};
    for (int l = 0; l < nentries; ++l)
    {
        const int p = indices[l];
        int b[3], o[3];
        for (int d = 0; d < 3; ++d)
        {
            b[d] = floor_div(cell[3 * p + d] - domain_lower[d], bin[d]);
            o[d] = (int)lround(Xshift[3 * l + d] / dx[d]);
        }
        /* entries arrive in canonical (bz,oz,by,oy,bx,ox,p) order: a (bin, image)
         * group is a run of consecutive entries */
        int newgroup = l == 0;
        for (int d = 0; d < 3; ++d) newgroup = newgroup || b[d] != prev_b[d] || o[d] != prev_o[d];
        if (newgroup)
        {
            ++group;
            for (int d = 0; d < 3; ++d)
            {
                prev_b[d] = b[d];
                prev_o[d] = o[d];
            }
        }
        const long rank = ranks[p];
        double w[8];
        const double Xs = X[3 * p + 2] + Xshift[3 * l + 2];
        z[l].key[0] = group;
        z[l].key[1] = rank / chunk;
        z[l].key[2] = orc_weights_1d_u(kernel, Xs, X[3 * p + 2], x_lower[2], dx[2], w);
        z[l].key[3] = rank;  /* lane order inside a chunk = bin order */
        z[l].pos = l;
    }
    qsort(z, (size_t)nentries, sizeof(zkey_t), cmp_zkey);
    for (int l = 0; l < nentries; ++l) perm[l] = z[l].pos;
    free(own);
    free(z);
}

/* ------------------------------------------------------- brute force check */

double orc_bruteforce_spread_at(int kernel, int ndim, const double* xg, const double* dx, const double* period,
                                const double* X, const double* V, int npoints)
{
    double vol = 1.0;
    for (int d = 0; d < ndim; ++d) vol *= dx[d];
    double acc = 0.0;
    for (int p = 0; p < npoints; ++p)
    {
        double wprod = 1.0;
        for (int d = 0; d < ndim; ++d)
        {
            double wsum = 0.0;
            const int nimg = period[d] > 0.0 ? 1 : 0;
            for (int o = -nimg; o <= nimg; ++o)
            {
                const double r = (xg[d] - (X[ndim * p + d] + o * period[d])) / dx[d];
                wsum += orc_delta(kernel, r);
            }
            wprod *= wsum;
        }
        acc += wprod * V[p];
    }
    return acc / vol;
}

/* ------------------------------------------------ Lagrangian force generation */

/* IBStandardForceGen::computeLagrangianForce (src/IB/IBStandardForceGen.cpp:254-305):
 * F_ghost = 0; springs (:814-930, default_spring_force kappa (R - rest),
 * IBSpringForceFunctions.h:104), beams (:1038-1147), target points (:1202-1299)
 * scattered into F_ghost in element order; then F += F_ghost. */
/* Spring tension T(R) of the built-in spring laws, selected per spring:
 *   law 0: default_spring_force, kappa (R - rest), params {kappa, rest}
 *          (include/ibamr/IBSpringForceFunctions.h:104-107);
 *   law 1: stiffness (R / rest - 1), params {stiffness, rest}: the
 *          linear_spring_force examples/IB/explicit/ex3/example.cpp:58-63
 *          registers (IBStandardForceGen::registerSpringForceFunction);
 *   law 2: a polynomial in the extension x = R - rest, params {c1, rest, c2, ...,
 *          cM}: T = x (c1 + x (c2 + ... + x cM)) (Horner, innermost last); with
 *          params {c1, rest} it is law 0.  No counterpart among the reference's
 *          functions: a built-in stand-in for a user's nonlinear law. */
static double spring_tension(int law, const double* p, int np, double R)
{
    if (law == 1) return p[0] * (R / p[1] - 1.0);
    if (law == 2)
    {
        const double x = R - p[1];
        double h = np > 2 ? p[np - 1] : p[0];
        for (int m = np - 2; m >= 2; --m) h = p[m] + x * h;
        if (np > 2) h = p[0] + x * h;
        return x * h;
    }
    return p[0] * (R - p[1]);
}

/* IBStandardForceGen::computeLagrangianForce with a spring law and parameter
 * list per spring (the reference's force_fcns[k] and parameters[k],
 * IBStandardForceGen.cpp:728-790, evaluated at :872 as force_fcn(R, params) / R).
 * Spring k's parameters are par[poff[k] .. poff[k+1]). */
void orc_lagrangian_force_laws(int ndim, int n, double* F, const double* X, const double* U, int ns, const int* sm,
                               const int* ss, const int* law, const int* poff, const double* par, int nb,
                               const int* bm, const int* bn, const int* bp, const double* bK, const double* bc,
                               int nt, const int* tn, const double* tK, const double* tE, const double* tX0)
{
    double* G = (double*)calloc((size_t)(n > 0 ? n : 1) * (size_t)ndim, sizeof(double));
    for (int k = 0; k < ns; ++k)
    {
        const int m = ndim * sm[k], s = ndim * ss[k];
        double D[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < ndim; ++d) D[d] = X[s + d] - X[m + d];
        double R2 = D[0] * D[0] + D[1] * D[1];
        if (ndim == 3) R2 = R2 + D[2] * D[2];
        const double R = sqrt(R2);
        if (R < DBL_EPSILON) continue;
        const double T_over_R = spring_tension(law[k], par + poff[k], poff[k + 1] - poff[k], R) / R;
        for (int d = 0; d < ndim; ++d)
        {
            const double f = T_over_R * D[d];
            G[m + d] += f;
            G[s + d] -= f;
        }
    }
    for (int k = 0; k < nb; ++k)
    {
        const int m = ndim * bm[k], nx = ndim * bn[k], pv = ndim * bp[k];
        for (int d = 0; d < ndim; ++d)
        {
            const double f = bK[k] * (((X[nx + d] + X[pv + d]) - 2.0 * X[m + d]) - bc[ndim * k + d]);
            G[m + d] += 2.0 * f;
            G[nx + d] -= f;
            G[pv + d] -= f;
        }
    }
    for (int k = 0; k < nt; ++k)
    {
        const int i = ndim * tn[k];
        for (int d = 0; d < ndim; ++d)
            G[i + d] += tK[k] * (tX0[ndim * k + d] - X[i + d]) - tE[k] * (U ? U[i + d] : 0.0);
    }
    for (int i = 0; i < n * ndim; ++i) F[i] = F[i] + G[i];
    free(G);
}

/* The Hookean springs of IBStandardForceGen's default force function (index 0):
 * law 0 with parameters {kappa[k], rest[k]}. */
void orc_lagrangian_force(int ndim, int n, double* F, const double* X, const double* U, int ns, const int* sm,
                          const int* ss, const double* sk, const double* sr, int nb, const int* bm, const int* bn,
                          const int* bp, const double* bK, const double* bc, int nt, const int* tn, const double* tK,
                          const double* tE, const double* tX0)
{
    int* law = (int*)calloc((size_t)(ns > 0 ? ns : 1), sizeof(int));
    int* poff = (int*)malloc(sizeof(int) * ((size_t)ns + 1));
    double* par = (double*)malloc(sizeof(double) * (size_t)(ns > 0 ? 2 * ns : 1));
    for (int k = 0; k <= ns; ++k) poff[k] = 2 * k;
    for (int k = 0; k < ns; ++k)
    {
        par[2 * k] = sk[k];
        par[2 * k + 1] = sr[k];
    }
    orc_lagrangian_force_laws(ndim, n, F, X, U, ns, sm, ss, law, poff, par, nb, bm, bn, bp, bK, bc, nt, tn, tK, tE,
                              tX0);
    free(law);
    free(poff);
    free(par);
}

/* ------------------------------------------------------------- redistribution */

/* LDataManager::beginDataRedistribution (LDataManager.cpp:1393-1431, 1460-1503)
 * and computeNodeDistribution (:2893-2908) for one level held in one place:
 * positions moved into the domain (periodic wrap by whole lengths, or clamp to
 * [x_lower, x_upper - L sqrt(eps)], or an error when error_if_leave), the
 * periodic offsets round((X_old - X_new)/dx), each point's patch (first patch
 * whose box holds getCellIndex), and the new order: by patch, cell in box order
 * (x fastest), Lagrangian index, then old position.  A point more than
 * max_wraps lengths out (or non-finite) counts as lost.
 * Returns 0, 1 (escaped), 2 (in no patch), 3 (lost). */
typedef struct
{
    long long key;
    int lag;
    int i;
} orc_rkey;

static int cmp_rkey(const void* a, const void* b)
{
    const orc_rkey* x = (const orc_rkey*)a;
    const orc_rkey* y = (const orc_rkey*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->lag != y->lag) return x->lag < y->lag ? -1 : 1;
    return (x->i > y->i) - (x->i < y->i);
}

int orc_redistribute(int ndim, int n, double* X, const int* lag, const double* x_lower, const double* x_upper,
                     const double* dx, const int* dom_lo, const int* dom_hi, const int* periodic, int error_if_leave,
                     int max_wraps, int npatches, const int* plo, const int* phi, int* new_order,
                     int* patch_offsets, int* offset)
{
    const double tol = sqrt(DBL_EPSILON); /* LDataManager.cpp:150 */
    long long* pcell = (long long*)malloc(sizeof(long long) * (size_t)(npatches + 1));
    orc_rkey* keys = (orc_rkey*)malloc(sizeof(orc_rkey) * (size_t)(n > 0 ? n : 1));
    int rc = 0;
    pcell[0] = 0;
    for (int q = 0; q < npatches; ++q)
    {
        long long c = 1;
        for (int d = 0; d < ndim; ++d) c *= (long long)(phi[ndim * q + d] - plo[ndim * q + d] + 1);
        pcell[q + 1] = pcell[q] + c;
    }
    for (int p = 0; p < n && rc == 0; ++p)
    {
        double* x = &X[(long)ndim * p];
        int cell[3] = {0, 0, 0};
        for (int d = 0; d < ndim; ++d)
        {
            const double Xr = x[d];
            double v = Xr;
            if (!isfinite(Xr))
            {
                rc = 3;
                break;
            }
            if (periodic[d])
            {
                const double L = x_upper[d] - x_lower[d];
                int w = 0;
                while (v < x_lower[d] && w < max_wraps)
                {
                    v += L;
                    ++w;
                }
                while (v >= x_upper[d] && w < max_wraps)
                {
                    v -= L;
                    ++w;
                }
                if (!(v >= x_lower[d] && v < x_upper[d])) rc = 3;
            }
            else if (error_if_leave)
            {
                if (v < x_lower[d] || v > x_upper[d]) rc = 1;
            }
            else
            {
                v = v > x_lower[d] ? v : x_lower[d];
                const double top = x_upper[d] - (x_upper[d] - x_lower[d]) * tol;
                v = v < top ? v : top;
            }
            x[d] = v;
            if (offset) offset[(long)ndim * p + d] = (int)round((Xr - v) / dx[d]);
        }
        if (rc) break;
        orc_get_cell_index(ndim, x, x_lower, x_upper, dx, dom_lo, dom_hi, cell);
        int owner = -1;
        for (int q = 0; q < npatches && owner < 0; ++q)
        {
            int in = 1;
            for (int d = 0; d < ndim; ++d) in = in && cell[d] >= plo[ndim * q + d] && cell[d] <= phi[ndim * q + d];
            if (in) owner = q;
        }
        if (owner < 0)
        {
            rc = 2;
            break;
        }
        long long lin = 0;
        for (int d = ndim - 1; d >= 0; --d)
            lin = lin * (phi[ndim * owner + d] - plo[ndim * owner + d] + 1) + (cell[d] - plo[ndim * owner + d]);
        keys[p].key = pcell[owner] + lin;
        keys[p].lag = lag ? lag[p] : p;
        keys[p].i = p;
    }
    if (rc == 0)
    {
        qsort(keys, (size_t)n, sizeof(orc_rkey), cmp_rkey);
        for (int k = 0; k < n; ++k) new_order[k] = keys[k].i;
        for (int q = 0; q <= npatches; ++q)
        {
            int c = 0;
            while (c < n && keys[c].key < pcell[q]) ++c;
            patch_offsets[q] = c;
        }
    }
    free(pcell);
    free(keys);
    return rc;
}

/* ---------------------------------------------------- FE quadrature point source */

/* FEDataManager::spread's point source (FEDataManager.cpp:1169-1204,
 * sum_weighted_elem_solution :825-857): X_qp = sum_k X_node phi[k][q],
 * F_qp = (sum_k F_node phi[k][q]) * JxW[e][q], sums from 0.0 in basis order. */
void orc_fe_quadrature_points(int ndim, int nvars, int nelem, int nbasis, int nqp, const int* elem_nodes,
                              const double* phi, const double* JxW, const double* X_nodes, const double* F_nodes,
                              double* X_qp, double* F_qp)
{
    for (int e = 0; e < nelem; ++e)
        for (int q = 0; q < nqp; ++q)
        {
            const long t = (long)e * nqp + q;
            for (int d = 0; d < ndim; ++d) X_qp[t * ndim + d] = 0.0;
            if (F_nodes)
                for (int i = 0; i < nvars; ++i) F_qp[t * nvars + i] = 0.0;
            for (int k = 0; k < nbasis; ++k)
            {
                const long node = elem_nodes[(long)e * nbasis + k];
                const double p = phi[(long)k * nqp + q];
                for (int d = 0; d < ndim; ++d) X_qp[t * ndim + d] += X_nodes[node * ndim + d] * p;
                if (F_nodes)
                    for (int i = 0; i < nvars; ++i) F_qp[t * nvars + i] += F_nodes[node * nvars + i] * p;
            }
            if (F_nodes && JxW)
                for (int i = 0; i < nvars; ++i) F_qp[t * nvars + i] *= JxW[(long)e * nqp + q];
        }
}

/* FEDataManager::interpWeighted's right-hand side (FEDataManager.cpp:2009-2061,
 * integrate_elem_rhs :1569-1606): per element, rhs[i][k] += U_qp[q][i] *
 * (phi[k][q] * JxW[q]) over q, k, i in the reference loop order, then added to
 * the nodes element by element.  F is accumulated into. */
void orc_fe_integrate_rhs(int nvars, int nelem, int nbasis, int nqp, const int* elem_nodes, const double* phi,
                          const double* JxW, const double* U_qp, double* F)
{
    double* rhs = (double*)malloc(sizeof(double) * (size_t)(nvars * nbasis));
    for (int e = 0; e < nelem; ++e)
    {
        for (int t = 0; t < nvars * nbasis; ++t) rhs[t] = 0.0;
        for (int q = 0; q < nqp; ++q)
        {
            const long idx = (long)nvars * ((long)e * nqp + q);
            for (int k = 0; k < nbasis; ++k)
            {
                const double p = phi[(long)k * nqp + q] * JxW[(long)e * nqp + q];
                for (int i = 0; i < nvars; ++i) rhs[nbasis * i + k] += U_qp[idx + i] * p;
            }
        }
        for (int i = 0; i < nvars; ++i)
            for (int k = 0; k < nbasis; ++k) F[(long)elem_nodes[(long)e * nbasis + k] * nvars + i] += rhs[nbasis * i + k];
    }
    free(rhs);
}

/* FEProjector::buildDiagonalL2MassMatrix (ibtk/src/lagrangian/FEProjector.cpp:587-672),
 * the dX of nodal quadrature (FEDataManager.cpp:1008-1014, 1797-1800): per
 * element M_e[k] = sum_q (phi[k][q] * phi[k][q]) * JxW[e][q] from 0.0, added to
 * the nodes element by element into M (zeroed first).  phi and JxW are those of
 * the nodal rule (build_nodal_qrule) the caller's FE library gives. */
void orc_fe_nodal_mass(int nelem, int nbasis, int nnodes, const int* elem_nodes, int nqp, const double* phi,
                       const double* JxW, double* M)
{
    for (int v = 0; v < nnodes; ++v) M[v] = 0.0;
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < nbasis; ++k)
        {
            double m = 0.0;
            for (int q = 0; q < nqp; ++q)
                m += (phi[(long)k * nqp + q] * phi[(long)k * nqp + q]) * JxW[(long)e * nqp + q];
            M[elem_nodes[(long)e * nbasis + k]] += m;
        }
}

/* Element types of the adaptive quadrature rule (libMesh ElemType names). */
enum
{
    ORC_EDGE2 = 0, ORC_EDGE3, ORC_TRI3, ORC_TRI6, ORC_QUAD4, ORC_QUAD8, ORC_QUAD9,
    ORC_TET4, ORC_TET10, ORC_HEX8, ORC_HEX20, ORC_HEX27
};

static double orc_dist2(int ndim, const double* X, long a, long b)
{
    double s = 0.0;
    for (int d = 0; d < ndim; ++d) s += (X[a * ndim + d] - X[b * ndim + d]) * (X[a * ndim + d] - X[b * ndim + d]);
    return s;
}

/* get_max_edge_length (ibtk/include/ibtk/libmesh_utilities.h:1373-1511) of one
 * element, nodes n[] in libMesh order: TET4 and HEX8 the longest vertex-vertex
 * edge; TET10 and HEX27 the longest edge as the sum of its two segments through
 * the mid-edge node; the other types (the reference's general branch) the
 * longest vertex pair on one edge (1-D: the two vertices; 2-D: the sides; HEX20:
 * the hexahedron's edges).  Squared lengths sum over the ndim coordinates. */
double orc_fe_max_edge_length(int type, int ndim, const int* n, const double* X)
{
    static const int tet4[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    static const int hex8[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                    {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
    static const int tet10[6][3] = {{0, 4, 1}, {0, 6, 2}, {0, 7, 3}, {1, 5, 2}, {1, 8, 3}, {2, 9, 3}};
    static const int hex27[12][3] = {{0, 8, 1},  {0, 11, 3}, {0, 12, 4}, {1, 9, 2},  {1, 13, 5}, {2, 10, 3},
                                     {2, 14, 6}, {3, 15, 7}, {4, 16, 5}, {4, 19, 7}, {5, 17, 6}, {6, 18, 7}};
    static const int tri[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    static const int quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}};
    double m2 = 0.0, m = 0.0;
    int i;
    switch (type)
    {
    case ORC_TET10:
        for (i = 0; i < 6; ++i)
        {
            const double l = sqrt(orc_dist2(ndim, X, n[tet10[i][0]], n[tet10[i][1]])) +
                             sqrt(orc_dist2(ndim, X, n[tet10[i][1]], n[tet10[i][2]]));
            m = l > m ? l : m;
        }
        return m;
    case ORC_HEX27:
        for (i = 0; i < 12; ++i)
        {
            const double l = sqrt(orc_dist2(ndim, X, n[hex27[i][0]], n[hex27[i][1]])) +
                             sqrt(orc_dist2(ndim, X, n[hex27[i][1]], n[hex27[i][2]]));
            m = l > m ? l : m;
        }
        return m;
    case ORC_TET4:
        for (i = 0; i < 6; ++i)
        {
            const double l = orc_dist2(ndim, X, n[tet4[i][0]], n[tet4[i][1]]);
            m2 = l > m2 ? l : m2;
        }
        return sqrt(m2);
    case ORC_HEX8:
    case ORC_HEX20:
        for (i = 0; i < 12; ++i)
        {
            const double l = orc_dist2(ndim, X, n[hex8[i][0]], n[hex8[i][1]]);
            m2 = l > m2 ? l : m2;
        }
        return sqrt(m2);
    case ORC_TRI3:
    case ORC_TRI6:
        for (i = 0; i < 3; ++i)
        {
            const double l = orc_dist2(ndim, X, n[tri[i][0]], n[tri[i][1]]);
            m2 = l > m2 ? l : m2;
        }
        return sqrt(m2);
    case ORC_QUAD4:
    case ORC_QUAD8:
    case ORC_QUAD9:
        for (i = 0; i < 4; ++i)
        {
            const double l = orc_dist2(ndim, X, n[quad[i][0]], n[quad[i][1]]);
            m2 = l > m2 ? l : m2;
        }
        return sqrt(m2);
    default: /* EDGE2, EDGE3 */
        return sqrt(orc_dist2(ndim, X, n[0], n[1]));
    }
}

/* getQuadratureKey's adaptive order (ibtk/src/utilities/libmesh_utilities.cpp:187-225)
 * per element: npts = int(ceil(point_density * hmax / dx_min)); below 3 it is 2
 * for first-order elements, 3 otherwise; QGAUSS (quad_type 0) order
 * min(2 npts - 1, 43) (FORTYTHIRD), QGRID (1) order npts. */
void orc_fe_adaptive_orders(int type, int ndim, int nelem, int nbasis, const int* elem_nodes, const double* X,
                            int quad_type, double point_density, double dx_min, int* order)
{
    const int first = type == ORC_EDGE2 || type == ORC_TRI3 || type == ORC_QUAD4 || type == ORC_TET4 ||
                      type == ORC_HEX8;
    for (int e = 0; e < nelem; ++e)
    {
        const double hmax = orc_fe_max_edge_length(type, ndim, elem_nodes + (long)e * nbasis, X);
        int npts = (int)ceil(point_density * hmax / dx_min);
        if (npts < 3) npts = first ? 2 : 3;
        order[e] = quad_type == 0 ? (2 * npts - 1 < 43 ? 2 * npts - 1 : 43) : npts;
    }
}

/* The quadrature point source and the weighted right-hand side with a rule per
 * element (the reference keys each element's rule separately, FEDataManager.cpp:
 * 1153-1165, 1951-1963): element e uses rule r = elem_rule[e] with rule_nqp[r]
 * points and table phi + rule_phi_off[r] ([nbasis][rule_nqp[r]]); its points and
 * JxW values start at qp_off[e].  Arithmetic per point as in the one-rule forms. */
void orc_fe_quadrature_points_rules(int ndim, int nvars, int nelem, int nbasis, const int* elem_nodes,
                                    const int* elem_rule, const int* rule_nqp, const long* rule_phi_off,
                                    const double* phi, const long* qp_off, const double* JxW,
                                    const double* X_nodes, const double* F_nodes, double* X_qp, double* F_qp)
{
    for (int e = 0; e < nelem; ++e)
    {
        const int r = elem_rule[e];
        orc_fe_quadrature_points(ndim, nvars, 1, nbasis, rule_nqp[r], elem_nodes + (long)e * nbasis,
                                 phi + rule_phi_off[r], JxW ? JxW + qp_off[e] : NULL, X_nodes, F_nodes,
                                 X_qp + qp_off[e] * ndim, F_nodes ? F_qp + qp_off[e] * nvars : NULL);
    }
}

void orc_fe_integrate_rhs_rules(int nvars, int nelem, int nbasis, const int* elem_nodes, const int* elem_rule,
                                const int* rule_nqp, const long* rule_phi_off, const double* phi, const long* qp_off,
                                const double* JxW, const double* U_qp, double* F)
{
    for (int e = 0; e < nelem; ++e)
    {
        const int r = elem_rule[e];
        orc_fe_integrate_rhs(nvars, 1, nbasis, rule_nqp[r], elem_nodes + (long)e * nbasis, phi + rule_phi_off[r],
                             JxW + qp_off[e], U_qp + qp_off[e] * nvars, F);
    }
}

/* ------------------------------------------------- physical boundary (Robin) ops */

/* The LINEAR Robin ghost-cell operators of CartCellRobinPhysBdryOp and
 * CartSideRobinPhysBdryOp (ibtk/src/boundary/physical_boundary/fortran/
 * cartphysbdryop{2,3}d.f.m4: ccrobinphysbdryop1{x,y,z}, ccrobinphysbdryop2/3,
 * scrobinphysbdryop1{x,y,z}, scrobinphysbdryop2/3), each with its adjoint
 * (adjoint_op = 1: accumulateFromPhysicalBoundaryData) and forward form
 * (setPhysicalBoundaryConditions), for one patch.  Arrays are the patch's
 * ghosted arrays (cell: one array holding depth components; side: one array
 * per axis), uniform ghost width gcw. */
typedef struct
{
    int ndim, side, depth, gcw;
    int ilo[3], ihi[3];
    double dx[3];
    double* arr[3];
} orc_bpatch;

static double* bp_at(const orc_bpatch* P, int a, int comp, const int* x)
{
    long off = 0, mul = 1;
    int n[3];
    for (int d = 0; d < P->ndim; ++d)
        n[d] = P->ihi[d] - P->ilo[d] + 1 + 2 * P->gcw + ((P->side && d == a) ? 1 : 0);
    for (int d = 0; d < P->ndim; ++d)
    {
        if (x[d] < P->ilo[d] - P->gcw || x[d] >= P->ilo[d] - P->gcw + n[d])
        {
            fprintf(stderr, "orc_phys_bdry: index %d of dim %d outside the ghost box\n", x[d], d);
            abort();
        }
        off += (long)(x[d] - (P->ilo[d] - P->gcw)) * mul;
        mul *= n[d];
    }
    return P->side ? P->arr[a] + off : P->arr[0] + (long)comp * mul + off;
}

/* the tangential dims of a face normal to dir, in the Fortran loop order (outer first) */
static int bp_tangential(int ndim, int dir, int* t)
{
    int m = 0;
    for (int d = ndim - 1; d >= 0; --d)
        if (d != dir) t[m++] = d;
    return m;
}

/* ccrobinphysbdryop1{x,y,z}: data box [dlo, dhi] (the patch box, or the side box of
 * array a for the transverse use), tangential ranges from the coefficient box */
static void bp_cc_op1(const orc_bpatch* P, int a, int comp, int loc, const int* dlo, const int* dhi, const int* blo,
                      const int* bhi, const double* cf, long ncf, int adjoint)
{
    const int ndim = P->ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = P->dx[loc / ndim]; /* as the Fortran indexes it */
    const int sgn = lower ? -1 : 1;
    const int ig = lower ? dlo[dir] - 1 : dhi[dir] + 1, ii = lower ? dlo[dir] : dhi[dir];
    int t[2];
    const int nt = bp_tangential(ndim, dir, t);
    int x[3] = {0, 0, 0};
    const int o_lo = nt > 1 ? blo[t[0]] : 0, o_hi = nt > 1 ? bhi[t[0]] : 0;
    long col = 0; /* the coefficient arrays' index: the lower tangential axis fastest */
    for (int o = o_lo; o <= o_hi; ++o)
        for (int q = blo[t[nt - 1]]; q <= bhi[t[nt - 1]]; ++q, ++col)
        {
            if (nt > 1) x[t[0]] = o;
            x[t[nt - 1]] = q;
            const double ca = cf[col], cb = cf[ncf + col], cg = cf[2 * ncf + col];
            for (int i = 0; i < P->gcw; ++i)
            {
                const double n = 1.0 + 2.0 * i;
                const double f_i = -(ca * n * h - 2.0 * cb) / (ca * n * h + 2.0 * cb);
                const double f_g = 2.0 * n * h / (ca * n * h + 2.0 * cb);
                int xg[3] = {x[0], x[1], x[2]}, xi[3] = {x[0], x[1], x[2]};
                xg[dir] = ig + sgn * i;
                xi[dir] = ii - sgn * i;
                if (adjoint)
                {
                    const double u_g = *bp_at(P, a, comp, xg);
                    *bp_at(P, a, comp, xi) += f_i * u_g;
                }
                else
                {
                    const double u_i = *bp_at(P, a, comp, xi);
                    *bp_at(P, a, comp, xg) = f_i * u_i + f_g * cg;
                }
            }
        }
}

/* scrobinphysbdryop1{x,y,z}: the normal component on its boundary face */
static void bp_sc_op1(const orc_bpatch* P, int loc, const int* blo, const int* bhi, const double* cf, long ncf,
                      int adjoint)
{
    const int ndim = P->ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = P->dx[loc / ndim];
    const int sgn = lower ? -1 : 1;
    const int ib = lower ? P->ilo[dir] : P->ihi[dir] + 1;
    int t[2];
    const int nt = bp_tangential(ndim, dir, t);
    int x[3] = {0, 0, 0};
    const int o_lo = nt > 1 ? blo[t[0]] : 0, o_hi = nt > 1 ? bhi[t[0]] : 0;
    long col = 0;
    for (int o = o_lo; o <= o_hi; ++o)
        for (int q = blo[t[nt - 1]]; q <= bhi[t[nt - 1]]; ++q, ++col)
        {
            if (nt > 1) x[t[0]] = o;
            x[t[nt - 1]] = q;
            const double ca = cf[col], cb = cf[ncf + col], cg = cf[2 * ncf + col];
            int xb[3] = {x[0], x[1], x[2]};
            xb[dir] = ib;
            double* ub = bp_at(P, dir, 0, xb);
            double u_b;
            const int dirichlet = fabs(cb) < 1.0e-12;
            if (dirichlet)
            {
                u_b = cg / ca;
                *ub = u_b;
            }
            else
                u_b = *ub;
            for (int i = 1; i <= P->gcw; ++i)
            {
                double f_i, f_b, f_g = 0.0;
                if (dirichlet)
                {
                    f_i = -1.0;
                    f_b = 2.0;
                }
                else
                {
                    const double n = 2.0 * i;
                    f_i = 1.0;
                    f_b = -ca * n * h / cb;
                    f_g = n * h / cb;
                }
                int xg[3] = {x[0], x[1], x[2]}, xi[3] = {x[0], x[1], x[2]};
                xg[dir] = ib + sgn * i;
                xi[dir] = ib - sgn * i;
                if (adjoint)
                {
                    const double u_g = *bp_at(P, dir, 0, xg);
                    *bp_at(P, dir, 0, xi) += f_i * u_g;
                    *ub += f_b * u_g;
                }
                else
                {
                    const double u_i = *bp_at(P, dir, 0, xi);
                    *bp_at(P, dir, 0, xg) = dirichlet ? f_i * u_i + f_b * u_b : f_i * u_i + f_b * u_b + f_g * cg;
                }
            }
        }
}

/* boundary index and direction of codim-2 boxes in 3-D: edge axis, the two
 * normal dims (n1 < n2), lower/upper per normal dim (PhysicalBoundaryUtilities
 * location conventions) */
static void bp_edge(int loc, int* e, int* n1, int* n2, int* up1, int* up2)
{
    if (loc < 4)
    {
        *e = 0, *n1 = 1, *n2 = 2, *up1 = loc & 1, *up2 = (loc >> 1) & 1;
    }
    else if (loc < 8)
    {
        *e = 1, *n1 = 0, *n2 = 2, *up1 = (loc - 4) >> 1, *up2 = (loc - 4) & 1;
    }
    else
    {
        *e = 2, *n1 = 0, *n2 = 1, *up1 = (loc - 8) & 1, *up2 = (loc - 8) >> 1;
    }
}

/* ccrobinphysbdryop2 (3-D edges): linear extrapolation and its adjoint */
static void bp_cc_edge(const orc_bpatch* P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bp_edge(loc, &e, &n1, &n2, &u1, &u2);
    const int b1 = u1 ? P->ihi[n1] : P->ilo[n1], b2 = u2 ? P->ihi[n2] : P->ilo[n2];
    const int s1 = u1 ? 1 : -1, s2 = u2 ? 1 : -1;
    /* Fortran loops k, j, i; every update keeps the edge index, so any loop order over
       the edge index with (n2 outer, n1 inner) inside gives the same values */
    for (int c2 = flo[n2]; c2 <= fhi[n2]; ++c2)
        for (int c1 = flo[n1]; c1 <= fhi[n1]; ++c1)
            for (int ce = flo[e]; ce <= fhi[e]; ++ce)
            {
                const int m1 = b1 + (b1 - c1 + s1), m2 = b2 + (b2 - c2 + s2);
                int x[3], xm[3], xa[3], xb[3], xc[3], xd[3];
                x[e] = xm[e] = xa[e] = xb[e] = xc[e] = xd[e] = ce;
                x[n1] = c1, x[n2] = c2;
                xm[n1] = m1, xm[n2] = m2;
                xa[n1] = b1, xa[n2] = c2;
                xb[n1] = c1, xb[n2] = b2;
                xc[n1] = b1, xc[n2] = m2;
                xd[n1] = m1, xd[n2] = b2;
                if (adjoint)
                {
                    const double Ug = *bp_at(P, 0, comp, x);
                    *bp_at(P, 0, comp, xm) += Ug;
                    *bp_at(P, 0, comp, xa) += Ug;
                    *bp_at(P, 0, comp, xb) += Ug;
                    *bp_at(P, 0, comp, xc) -= Ug;
                    *bp_at(P, 0, comp, xd) -= Ug;
                }
                else
                    *bp_at(P, 0, comp, x) = *bp_at(P, 0, comp, xm) + (*bp_at(P, 0, comp, xa) - *bp_at(P, 0, comp, xc)) +
                                            (*bp_at(P, 0, comp, xb) - *bp_at(P, 0, comp, xd));
            }
}

/* ccrobinphysbdryop22d / ccrobinphysbdryop33d (corners) */
static void bp_cc_corner(const orc_bpatch* P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    const int nd = P->ndim;
    int b[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        b[d] = up ? P->ihi[d] : P->ilo[d];
        s[d] = up ? 1 : -1;
    }
    const int zl = nd == 3 ? flo[2] : 0, zh = nd == 3 ? fhi[2] : 0;
    for (int z = zl; z <= zh; ++z)
        for (int y = flo[1]; y <= fhi[1]; ++y)
            for (int xx = flo[0]; xx <= fhi[0]; ++xx)
            {
                const int x[3] = {xx, y, z};
                int m[3] = {0, 0, 0};
                for (int d = 0; d < nd; ++d) m[d] = b[d] + (b[d] - x[d] + s[d]);
                int pk[3][3], mk[3][3];  /* +: keep d, others at b; -: d mirrored, others at b */
                for (int d = 0; d < nd; ++d)
                    for (int o = 0; o < 3; ++o)
                    {
                        pk[d][o] = o == d ? x[o] : b[o];
                        mk[d][o] = o == d ? m[o] : b[o];
                    }
                if (adjoint)
                {
                    const double Ug = *bp_at(P, 0, comp, x);
                    *bp_at(P, 0, comp, m) += Ug;
                    for (int d = 0; d < nd; ++d) *bp_at(P, 0, comp, pk[d]) += Ug;
                    for (int d = 0; d < nd; ++d) *bp_at(P, 0, comp, mk[d]) -= Ug;
                }
                else
                {
                    double v = *bp_at(P, 0, comp, m);
                    for (int d = 0; d < nd; ++d) v = v + (*bp_at(P, 0, comp, pk[d]) - *bp_at(P, 0, comp, mk[d]));
                    *bp_at(P, 0, comp, x) = v;
                }
            }
}

/* scrobinphysbdryop2 (3-D edges, side data): component n1 extrapolated along n2
 * where its index along n1 is outside the side box, then component n2 along n1 */
static void bp_sc_edge(const orc_bpatch* P, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bp_edge(loc, &e, &n1, &n2, &u1, &u2);
    const int b[3] = {0, 0, 0};
    int bd[3] = {b[0], b[1], b[2]}, sh[3] = {0, 0, 0};
    bd[n1] = u1 ? P->ihi[n1] : P->ilo[n1];
    bd[n2] = u2 ? P->ihi[n2] : P->ilo[n2];
    sh[n1] = u1 ? -1 : 1;
    sh[n2] = u2 ? -1 : 1;
    for (int part = 0; part < 2; ++part)
    {
        const int comp = part == 0 ? n1 : n2, other = part == 0 ? n2 : n1;
        int lo[3] = {flo[0], flo[1], flo[2]}, hi[3] = {fhi[0], fhi[1], fhi[2]};
        hi[comp] += 1;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int xx = lo[0]; xx <= hi[0]; ++xx)
                {
                    const int x[3] = {xx, y, z};
                    if (!(x[comp] < P->ilo[comp] || x[comp] > P->ihi[comp] + 1)) continue;
                    const double del = (double)abs(x[other] - bd[other]);
                    int xb[3] = {x[0], x[1], x[2]}, xs[3] = {x[0], x[1], x[2]};
                    xb[other] = bd[other];
                    xs[other] = bd[other] + sh[other];
                    if (adjoint)
                    {
                        const double u_g = *bp_at(P, comp, 0, x);
                        *bp_at(P, comp, 0, xb) += (1.0 + del) * u_g;
                        *bp_at(P, comp, 0, xs) -= del * u_g;
                    }
                    else
                        *bp_at(P, comp, 0, x) = (1.0 + del) * *bp_at(P, comp, 0, xb) - del * *bp_at(P, comp, 0, xs);
                }
    }
}

/* scrobinphysbdryop22d / scrobinphysbdryop33d (corners, side data) */
static void bp_sc_corner(const orc_bpatch* P, int loc, const int* flo, const int* fhi, int adjoint)
{
    const int nd = P->ndim;
    int bd[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        bd[d] = up ? P->ihi[d] : P->ilo[d];
        sh[d] = up ? -1 : 1;
    }
    for (int comp = 0; comp < nd; ++comp)
    {
        int lo[3] = {flo[0], flo[1], nd == 3 ? flo[2] : 0}, hi[3] = {fhi[0], fhi[1], nd == 3 ? fhi[2] : 0};
        hi[comp] += 1;
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int xx = lo[0]; xx <= hi[0]; ++xx)
                {
                    const int x[3] = {xx, y, z};
                    if (!(x[comp] < P->ilo[comp] || x[comp] > P->ihi[comp] + 1)) continue;
                    double dsum = 1.0, del[3] = {0.0, 0.0, 0.0};
                    for (int o = 0; o < nd; ++o)
                        if (o != comp)
                        {
                            del[o] = (double)abs(x[o] - bd[o]);
                            dsum = dsum + del[o];
                        }
                    int xb[3] = {x[0], x[1], x[2]};
                    for (int o = 0; o < nd; ++o)
                        if (o != comp) xb[o] = bd[o];
                    if (adjoint)
                    {
                        const double u_g = *bp_at(P, comp, 0, x);
                        *bp_at(P, comp, 0, xb) += dsum * u_g;
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                *bp_at(P, comp, 0, xs) -= del[o] * u_g;
                            }
                    }
                    else
                    {
                        double v = dsum * *bp_at(P, comp, 0, xb);
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                v = v - del[o] * *bp_at(P, comp, 0, xs);
                            }
                        *bp_at(P, comp, 0, x) = v;
                    }
                }
    }
}

/* One patch: boxes[k] = {codim, location, lo[3], hi[3]} (codim 1: the boundary
 * box, the cell layer outside the face; codim >= 2: the fill box), abg[k][narr][3]
 * the Robin coefficients of codim-1 boxes per array (cell: depth component;
 * side: axis).  adjoint = 1: accumulateFromPhysicalBoundaryData (codim 3, 2,
 * then 1; side: transverse then normal); 0: setPhysicalBoundaryConditions
 * (codim 1 (side: normal, then transverse), 2, 3). */
/* Coefficient values of codim-1 box k (boxes: nboxes x {codim, location, lo[3],
 * hi[3]}) for array j: its tangential cells, one more along j for a side
 * array tangential to the box (the coefficient box's tangential extension). */
static long bp_ncoef(int ndim, int side, const int* bx, int j)
{
    const int dir = bx[1] / 2;
    long m = 1;
    for (int d = 0; d < ndim; ++d)
        if (d != dir) m *= (long)(bx[5 + d] - bx[2 + d] + 1 + (side && d == j ? 1 : 0));
    return m;
}

long orc_phys_bdry_ncoefs(int ndim, int side, int depth, int nboxes, const int* boxes)
{
    const int narr = side ? ndim : depth;
    long n = 0;
    for (int k = 0; k < nboxes; ++k)
        if (boxes[8 * k] == 1)
            for (int j = 0; j < narr; ++j) n += 3 * bp_ncoef(ndim, side, boxes + 8 * k, j);
    return n;
}

void orc_phys_bdry_coefs(int ndim, int side, int depth, int gcw, const int* ilo, const int* ihi, const double* dx,
                         double** arrays, int nboxes, const int* boxes, const double* coefs, int adjoint)
{
    orc_bpatch P;
    P.ndim = ndim;
    P.side = side;
    P.depth = side ? 1 : depth;
    P.gcw = gcw;
    for (int d = 0; d < 3; ++d)
    {
        P.ilo[d] = d < ndim ? ilo[d] : 0;
        P.ihi[d] = d < ndim ? ihi[d] : 0;
        P.dx[d] = d < ndim ? dx[d] : 1.0;
        P.arr[d] = d < (side ? ndim : 1) ? arrays[d] : 0;
    }
    const int narr = side ? ndim : depth;
    /* each codim-1 box's coefficient arrays (a, b, g over its columns) per array, in box order */
    long* off = (long*)malloc(sizeof(long) * (size_t)(nboxes * narr + 1));
    long* ncf = (long*)malloc(sizeof(long) * (size_t)(nboxes * narr + 1));
    long o = 0;
    for (int k = 0; k < nboxes; ++k)
        for (int j = 0; j < narr; ++j)
        {
            off[k * narr + j] = o;
            ncf[k * narr + j] = boxes[8 * k] == 1 ? bp_ncoef(ndim, side, boxes + 8 * k, j) : 0;
            o += 3 * ncf[k * narr + j];
        }
    /* phases: adjoint 3, 2, 1T, 1N; forward 1N, 1T, 2, 3 (cell: 1T is the cell op, 1N empty) */
    const int phases_adj[4] = {3, 2, 10, 11}, phases_fwd[4] = {11, 10, 2, 3};
    for (int ph = 0; ph < 4; ++ph)
    {
        const int phase = adjoint ? phases_adj[ph] : phases_fwd[ph];
        for (int k = 0; k < nboxes; ++k)
        {
            const int* bx = boxes + 8 * k;
            const int codim = bx[0], loc = bx[1];
            const int* lo = bx + 2;
            const int* hi = bx + 5;
            if (phase == 3 && codim == 3 && ndim == 3)
            {
                for (int c = 0; c < (side ? 1 : depth); ++c)
                    side ? bp_sc_corner(&P, loc, lo, hi, adjoint) : bp_cc_corner(&P, c, loc, lo, hi, adjoint);
            }
            else if (phase == 2 && codim == 2)
            {
                for (int c = 0; c < (side ? 1 : depth); ++c)
                {
                    if (ndim == 2)
                        side ? bp_sc_corner(&P, loc, lo, hi, adjoint) : bp_cc_corner(&P, c, loc, lo, hi, adjoint);
                    else
                        side ? bp_sc_edge(&P, loc, lo, hi, adjoint) : bp_cc_edge(&P, c, loc, lo, hi, adjoint);
                }
            }
            else if ((phase == 10 || phase == 11) && codim == 1)
            {
                const int dir = loc / 2;
                if (!side)
                {
                    if (phase == 10)
                        for (int c = 0; c < depth; ++c)
                            bp_cc_op1(&P, 0, c, loc, P.ilo, P.ihi, lo, hi, coefs + off[k * narr + c],
                                      ncf[k * narr + c], adjoint);
                }
                else if (phase == 10)
                {
                    for (int a = 0; a < ndim; ++a)
                    {
                        if (a == dir) continue;
                        int shi[3] = {P.ihi[0], P.ihi[1], P.ihi[2]}, bhi[3] = {hi[0], hi[1], hi[2]};
                        shi[a] += 1;  /* the side box of array a */
                        bhi[a] += 1;  /* tangential extension of the coefficient box */
                        bp_cc_op1(&P, a, 0, loc, P.ilo, shi, lo, bhi, coefs + off[k * narr + a], ncf[k * narr + a],
                                  adjoint);
                    }
                }
                else
                    bp_sc_op1(&P, loc, lo, hi, coefs + off[k * narr + dir], ncf[k * narr + dir], adjoint);
            }
        }
    }
    free(off);
    free(ncf);
}

void orc_phys_bdry(int ndim, int side, int depth, int gcw, const int* ilo, const int* ihi, const double* dx,
                   double** arrays, int nboxes, const int* boxes, const double* abg, int adjoint)
{
    /* uniform coefficients: every column of a box and array gets its (a, b, g) */
    const int narr = side ? ndim : depth;
    const long n = orc_phys_bdry_ncoefs(ndim, side, depth, nboxes, boxes);
    double* coefs = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    long o = 0;
    for (int k = 0; k < nboxes; ++k)
        if (boxes[8 * k] == 1)
            for (int j = 0; j < narr; ++j)
            {
                const long m = bp_ncoef(ndim, side, boxes + 8 * k, j);
                for (int c = 0; c < 3; ++c)
                    for (long i = 0; i < m; ++i) coefs[o + c * m + i] = abg[3 * ((long)k * narr + j) + c];
                o += 3 * m;
            }
    orc_phys_bdry_coefs(ndim, side, depth, gcw, ilo, ihi, dx, arrays, nboxes, boxes, coefs, adjoint);
    free(coefs);
}
