// le_weights.h -- 1-D regularized delta weights on the device (gfx950).
//
// FP64 forms evaluate the same operations in the same order as the reference
// Fortran kernels, so stencil placement and weights are bitwise identical to
// the FP64 CPU path (the library is compiled with -ffp-contract=off; FP64
// division and sqrt are IEEE correctly rounded on gfx950).  FP32 forms keep the
// position-to-index conversion and the stencil selection in FP64 and evaluate
// the weights in FP32.
//
// Reference: ibtk/src/lagrangian/fortran/lagrangian_interaction3d.f.m4
//   IB_4 :1265-1273, IB_6 :2238-2272 (K at :2227), BSPLINE_3 :2655-2676,
//   BSPLINE_4 :2882-2908, BSPLINE_5 :3230-3248, BSPLINE_6 :3460-3480,
//   and lagrangian_delta.f.m4:243-354 (B-spline deltas).
#pragma once

#include <hip/hip_runtime.h>

namespace ibamr
{
enum KernelId : int
{
    K_PWL = 0,
    K_IB4 = 1,
    K_IB6 = 2,
    K_BS3 = 3,
    K_BS4 = 4,
    K_BS5 = 5,
    K_BS6 = 6
};

template <int K>
struct KernelTraits;
template <>
struct KernelTraits<K_IB4>
{
    static constexpr int W = 4;       // weights per coordinate
    static constexpr bool IB = true;  // tensor product w0*(w1*w2) (else ((w0*w1)*w2))
    static constexpr int GHOST = 3;   // getMinimumGhostWidth
};
template <>
struct KernelTraits<K_IB6>
{
    static constexpr int W = 6;
    static constexpr bool IB = true;
    static constexpr int GHOST = 4;
};
template <>
struct KernelTraits<K_BS3>
{
    static constexpr int W = 3;
    static constexpr bool IB = false;
    static constexpr int GHOST = 3;
};
template <>
struct KernelTraits<K_BS4>
{
    static constexpr int W = 4;
    static constexpr bool IB = false;
    static constexpr int GHOST = 3;  // stencil size 4
};
template <>
struct KernelTraits<K_BS5>
{
    static constexpr int W = 5;
    static constexpr bool IB = false;
    static constexpr int GHOST = 4;  // stencil size 6 (LEInteractor.cpp getStencilSize)
};
template <>
struct KernelTraits<K_BS6>
{
    static constexpr int W = 6;
    static constexpr bool IB = false;
    static constexpr int GHOST = 4;
};
// Kernels whose stencil side depends on the unshifted X (weights_1d_u).
template <int K>
constexpr bool side_by_unshifted_x()
{
    return K == K_BS4 || K == K_BS6;
}
template <>
struct KernelTraits<K_PWL>
{
    static constexpr int W = 2;
    static constexpr bool IB = false;  // (((w0*w1)*w2)*V)/vol (f.m4:717-721)
    static constexpr int GHOST = 2;
};

// Fortran NINT: round half away from zero.
__device__ __forceinline__ int nint_d(double x) { return (int)round(x); }

// lagrangian_delta.f.m4:243-260, FP64
__device__ __forceinline__ double bspline3_delta_d(double x)
{
    const double modx = fabs(x);
    const double r = modx + 1.5;
    const double r2 = r * r;
    if (modx <= 0.5) return 0.5 * (-2.0 * r2 + 6.0 * r - 3.0);
    if (modx <= 1.5) return 0.5 * (r2 - 6.0 * r + 9.0);
    return 0.0;
}

__device__ __forceinline__ float bspline3_delta_f(float x)
{
    // Same piecewise quadratic written around the cell centre to avoid the
    // cancellation of the (|x|+1.5) form: 3/4 - x^2 and (3/2-|x|)^2/2.
    const float a = fabsf(x);
    if (a <= 0.5f) return 0.75f - a * a;
    if (a <= 1.5f)
    {
        const float t = 1.5f - a;
        return 0.5f * t * t;
    }
    return 0.0f;
}

// lagrangian_delta.f.m4:268-354, FP64, the reference's operation order: each
// piece a polynomial in r = |x| + n/2, evaluated left to right.
__device__ __forceinline__ double bspline4_delta_d(double x)
{
    const double modx = fabs(x);
    const double r = modx + 2.0;
    const double r2 = r * r;
    const double r3 = r2 * r;
    if (modx <= 1.0) return (1.0 / 6.0) * (((3.0 * r3 - 24.0 * r2) + 60.0 * r) - 44.0);
    if (modx <= 2.0) return (1.0 / 6.0) * (((-r3 + 12.0 * r2) - 48.0 * r) + 64.0);
    return 0.0;
}
__device__ __forceinline__ double bspline5_delta_d(double x)
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
__device__ __forceinline__ double bspline6_delta_d(double x)
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
    if (modx <= 3.0) return (1.0 / 120.0) * (((((7776.0 - 6480.0 * r) + 2160.0 * r2) - 360.0 * r3) + 30.0 * r4) - r5);
    return 0.0;
}

// Shared stencil of BSPLINE_4/5/6: centre cell of X+Xshift, then the odd order
// takes centre -2..+2 and the even orders one side chosen by comparing the
// unshifted X with that cell's centre (Xu; Xu = Xs without a shift).
template <int K>
__device__ __forceinline__ int bspline_hi_weights(double Xs, double Xu, double x_lower, double dx, double* w)
{
    constexpr int W = KernelTraits<K>::W;
    const int c = (int)floor((Xs - x_lower) / dx);
    int lo;
    if (K == K_BS5)
        lo = c - 2;
    else
    {
        const double X_c = x_lower + ((double)c + 0.5) * dx;
        lo = Xu < X_c ? c - W / 2 : c - W / 2 + 1;
    }
#pragma unroll
    for (int k = 0; k < W; ++k)
    {
        const double X_cell = x_lower + ((double)(lo + k) + 0.5) * dx;
        const double t = (Xs - X_cell) / dx;
        w[k] = K == K_BS4 ? bspline4_delta_d(t) : (K == K_BS5 ? bspline5_delta_d(t) : bspline6_delta_d(t));
    }
    return lo;
}

// K of the 6-point kernel (f.m4:2227) and the constant sub-expressions of
// :2243-2260 as the Fortran compiler folds them (one rounding per operation).
struct IB6Const
{
    double K, c_b1, c_g1, c_g2, c_m2, c_m1;
};
__device__ __forceinline__ IB6Const ib6_const()
{
    IB6Const c;
    c.K = (59.0 / 60.0) * (1.0 - sqrt(1.0 - (3220.0 / 3481.0)));
    c.c_b1 = (22.0 / 3.0) - 7.0 * c.K;
    c.c_g1 = (((161.0 / 36.0) - (59.0 / 6.0) * c.K) + 5.0 * (c.K * c.K)) * (1.0 / 2.0);
    c.c_g2 = (-(109.0 / 24.0) + 5.0 * c.K) * (1.0 / 3.0);
    c.c_m2 = (1.0 / 12.0) * (3.0 * c.K - 1.0);
    c.c_m1 = (1.0 / 6.0) * (4.0 - 3.0 * c.K);
    return c;
}

__device__ __forceinline__ void ib6_weights_d(double r, double* w)
{
    const IB6Const c = ib6_const();
    const double K = c.K;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double r4 = r2 * r2;
    const double r6 = r3 * r3;
    const double alpha = 28.0;
    const double beta = (((9.0 / 4.0) - (3.0 / 2.0) * (K + r2)) + c.c_b1 * r) - (7.0 / 3.0) * r3;
    const double gamma = (1.0 / 4.0) * ((c.c_g1 * r2 + c.c_g2 * r4) + (5.0 / 18.0) * r6);
    const double discr = beta * beta - (4.0 * alpha) * gamma;
    const double pm3 = (-beta + 1.0 * sqrt(discr)) / (2.0 * alpha); // sign(1, 3/2-K) = +1
    w[0] = pm3;
    w[1] = ((((-(3.0 * pm3)) - (1.0 / 16.0)) + (1.0 / 8.0) * (K + r2)) + c.c_m2 * r) + (1.0 / 12.0) * r3;
    w[2] = (((2.0 * pm3) + (1.0 / 4.0)) + c.c_m1 * r) - (1.0 / 6.0) * r3;
    w[3] = ((2.0 * pm3) + (5.0 / 8.0)) - (1.0 / 4.0) * (K + r2);
    w[4] = (((-(3.0 * pm3)) + (1.0 / 4.0)) - c.c_m1 * r) + (1.0 / 6.0) * r3;
    w[5] = (((pm3 - (1.0 / 16.0)) + (1.0 / 8.0) * (K + r2)) - c.c_m2 * r) - (1.0 / 12.0) * r3;
}

__device__ __forceinline__ void ib6_weights_f(float r, float* w)
{
    // Constants rounded once from their FP64 values.
    const IB6Const c = ib6_const();
    const float K = (float)c.K;
    const float b0 = (float)(9.0 / 4.0 - 1.5 * c.K);
    const float b1 = (float)c.c_b1;
    const float g1 = (float)c.c_g1, g2 = (float)c.c_g2;
    const float m2 = (float)c.c_m2, m1 = (float)c.c_m1;
    const float Kq = (float)(c.K / 8.0);
    const float r2 = r * r;
    const float r3 = r2 * r;
    const float r4 = r2 * r2;
    const float r6 = r3 * r3;
    const float beta = ((b0 - 1.5f * r2) + b1 * r) - (7.0f / 3.0f) * r3;
    const float gamma = 0.25f * ((g1 * r2 + g2 * r4) + (5.0f / 18.0f) * r6);
    const float discr = beta * beta - 112.0f * gamma;
    // root (-beta + sqrt(discr)) / 56 of 28 p^2 + beta p + gamma = 0, in the
    // form that does not cancel when beta > 0
    const float sq = sqrtf(discr);
    const float pm3 = beta > 0.0f ? (-2.0f * gamma) / (beta + sq) : (sq - beta) * (1.0f / 56.0f);
    const float s = Kq + 0.125f * r2;            // (1/8)(K + r^2)
    w[0] = pm3;
    w[1] = ((-3.0f * pm3 - 0.0625f) + s + m2 * r) + (1.0f / 12.0f) * r3;
    w[2] = ((2.0f * pm3 + 0.25f) + m1 * r) - (1.0f / 6.0f) * r3;
    w[3] = (2.0f * pm3 + 0.625f) - 2.0f * s;
    w[4] = ((-3.0f * pm3 + 0.25f) - m1 * r) + (1.0f / 6.0f) * r3;
    w[5] = ((pm3 - 0.0625f) + s - m2 * r) - (1.0f / 12.0f) * r3;
    (void)K;
}

// FP64 1-D weights.  Xs = X + Xshift (already added, as the Fortran does);
// returns the first stencil index relative to ilower.
template <int K>
__device__ __forceinline__ int weights_1d(double Xs, double x_lower, double dx, double* w);

template <>
__device__ __forceinline__ int weights_1d<K_IB4>(double Xs, double x_lower, double dx, double* w)
{
    const double X_o_dx = (Xs - x_lower) / dx;
    const int lo = nint_d(X_o_dx) - 2;
    const double r = X_o_dx - ((double)(lo + 1) + 0.5);
    const double q = sqrt(1.0 + (4.0 * r) * (1.0 - r));
    w[0] = 0.125 * ((3.0 - 2.0 * r) - q);
    w[1] = 0.125 * ((3.0 - 2.0 * r) + q);
    w[2] = 0.125 * ((1.0 + 2.0 * r) + q);
    w[3] = 0.125 * ((1.0 + 2.0 * r) - q);
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_IB6>(double Xs, double x_lower, double dx, double* w)
{
    const double X_o_dx = (Xs - x_lower) / dx;
    const int lo = nint_d(X_o_dx) - 3;
    const double r = (1.0 - X_o_dx) + ((double)(lo + 2) + 0.5);
    ib6_weights_d(r, w);
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_BS3>(double Xs, double x_lower, double dx, double* w)
{
    const int c = (int)floor((Xs - x_lower) / dx);
    const int lo = c - 1;
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        const double X_cell = x_lower + ((double)(lo + k) + 0.5) * dx;
        w[k] = bspline3_delta_d((Xs - X_cell) / dx);
    }
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_BS4>(double Xs, double x_lower, double dx, double* w)
{
    return bspline_hi_weights<K_BS4>(Xs, Xs, x_lower, dx, w);
}
template <>
__device__ __forceinline__ int weights_1d<K_BS5>(double Xs, double x_lower, double dx, double* w)
{
    return bspline_hi_weights<K_BS5>(Xs, Xs, x_lower, dx, w);
}
template <>
__device__ __forceinline__ int weights_1d<K_BS6>(double Xs, double x_lower, double dx, double* w)
{
    return bspline_hi_weights<K_BS6>(Xs, Xs, x_lower, dx, w);
}

// PIECEWISE_LINEAR, lagrangian_interaction3d.f.m4:680-699 (spread) / 556-575
// (interp): the cell whose centre is nearest, then the linear hat on the two
// cells straddling X.
template <>
__device__ __forceinline__ int weights_1d<K_PWL>(double Xs, double x_lower, double dx, double* w)
{
    const int c = nint_d((Xs - x_lower) / dx - 0.5);
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

// FP32 1-D weights: index selection and the offset r in FP64, weights in FP32.
template <int K>
__device__ __forceinline__ int weights_1d(double Xs, double x_lower, double dx, float* w);

template <>
__device__ __forceinline__ int weights_1d<K_IB4>(double Xs, double x_lower, double dx, float* w)
{
    const double X_o_dx = (Xs - x_lower) / dx;
    const int lo = nint_d(X_o_dx) - 2;
    const float r = (float)(X_o_dx - ((double)(lo + 1) + 0.5));
    const float q = sqrtf(1.0f + (4.0f * r) * (1.0f - r));
    w[0] = 0.125f * ((3.0f - 2.0f * r) - q);
    w[1] = 0.125f * ((3.0f - 2.0f * r) + q);
    w[2] = 0.125f * ((1.0f + 2.0f * r) + q);
    w[3] = 0.125f * ((1.0f + 2.0f * r) - q);
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_IB6>(double Xs, double x_lower, double dx, float* w)
{
    const double X_o_dx = (Xs - x_lower) / dx;
    const int lo = nint_d(X_o_dx) - 3;
    const float r = (float)((1.0 - X_o_dx) + ((double)(lo + 2) + 0.5));
    ib6_weights_f(r, w);
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_BS3>(double Xs, double x_lower, double dx, float* w)
{
    const double X_o_dx = (Xs - x_lower) / dx;
    const int c = (int)floor(X_o_dx);
    const int lo = c - 1;
    // distance from the centre cell in cells, FP64 then rounded once
    const float t = (float)(X_o_dx - ((double)c + 0.5));
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = bspline3_delta_f(t - (float)(k - 1));
    return lo;
}

template <>
__device__ __forceinline__ int weights_1d<K_PWL>(double Xs, double x_lower, double dx, float* w)
{
    double wd[2];
    const int lo = weights_1d<K_PWL>(Xs, x_lower, dx, wd);
    w[0] = (float)wd[0];
    w[1] = (float)wd[1];
    return lo;
}

// BSPLINE_4/5/6 in FP32 forms: the FP64 weights rounded once (their pieces are
// polynomials in |x| + n/2 whose terms cancel, which FP32 would not survive).
#define IBAMR_BS_HI_F(K)                                                                     \
    template <>                                                                              \
    __device__ __forceinline__ int weights_1d<K>(double Xs, double x_lower, double dx, float* w) \
    {                                                                                        \
        double wd[KernelTraits<K>::W];                                                       \
        const int lo = weights_1d<K>(Xs, x_lower, dx, wd);                                   \
        for (int k = 0; k < KernelTraits<K>::W; ++k) w[k] = (float)wd[k];                    \
        return lo;                                                                           \
    }
IBAMR_BS_HI_F(K_BS4)
IBAMR_BS_HI_F(K_BS5)
IBAMR_BS_HI_F(K_BS6)
#undef IBAMR_BS_HI_F

// Weights of X(d,s)+Xshift with the unshifted X(d,s) as well: only BSPLINE_4/6
// use it (their stencil side); for the others this is weights_1d.
template <int K, typename T>
__device__ __forceinline__ int weights_1d_u(double Xs, double Xu, double x_lower, double dx, T* w)
{
    if constexpr (side_by_unshifted_x<K>())
    {
        double wd[KernelTraits<K>::W];
        const int lo = bspline_hi_weights<K>(Xs, Xu, x_lower, dx, wd);
        for (int k = 0; k < KernelTraits<K>::W; ++k) w[k] = (T)wd[k];
        return lo;
    }
    else
    {
        (void)Xu;
        return weights_1d<K>(Xs, x_lower, dx, w);
    }
}

} // namespace ibamr
