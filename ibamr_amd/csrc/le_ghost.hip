// le_ghost.hip -- grid steps around the path: periodic and same-level
// ghost fills (the RefineSchedule fill before interp, LDataManager.cpp:742-745) and
// the physical-boundary Robin operators (CartCell/CartSideRobinPhysBdryOp).
#include "le_plan_internal.h"

namespace ibamr
{
namespace
{
// ----------------------------------------------------------------------------- periodic ghost fill

// Periodic ghost fill of one array of a single-patch, fully periodic level:
// every value outside the core [g, g+N) of each axis takes the value of its
// periodic image, which lies in the core (the ghost width never exceeds the
// period).  Only the ghost shell is enumerated: z slabs (whole planes), then y
// slabs inside the core z range, then x slabs inside the core y and z ranges.
template <typename T>
__global__ void k_fill_periodic(T* __restrict__ arr, int n0, int n1, int n2, int g0, int g1, int g2, int N0, int N1,
                                int N2, long sz, long sy, long sx)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i0, i1, i2;
    if (t < sz)
    {
        const long plane = (long)n0 * n1;
        const int q = (int)(t / plane);                       // which ghost plane
        const long r = t - (long)q * plane;
        i2 = q < g2 ? q : q - g2 + g2 + N2;                   // lower planes, then upper planes
        i1 = (int)(r / n0);
        i0 = (int)(r - (long)i1 * n0);
    }
    else if ((t -= sz) < sy)
    {
        const long slab = (long)(n1 - N1) * n0;              // per core z plane
        const int z = (int)(t / slab);
        const long r = t - (long)z * slab;
        const int q = (int)(r / n0);
        i0 = (int)(r - (long)q * n0);
        i1 = q < g1 ? q : q - g1 + g1 + N1;
        i2 = g2 + z;
    }
    else if ((t -= sy) < sx)
    {
        const int w = n0 - N0;                               // per core (y, z) row
        const long row = t / w;
        const int q = (int)(t - row * w);
        i0 = q < g0 ? q : q - g0 + g0 + N0;
        i1 = g1 + (int)(row % N1);
        i2 = g2 + (int)(row / N1);
    }
    else
        return;
    int j0 = i0 - g0, j1 = i1 - g1, j2 = i2 - g2;
    j0 += j0 < 0 ? N0 : (j0 >= N0 ? -N0 : 0);
    j1 += j1 < 0 ? N1 : (j1 >= N1 ? -N1 : 0);
    j2 += j2 < 0 ? N2 : (j2 >= N2 ? -N2 : 0);
    arr[(long)i0 + (long)n0 * (i1 + (long)n1 * i2)] =
        arr[(long)(j0 + g0) + (long)n0 * ((j1 + g1) + (long)n1 * (j2 + g2))];
}

// Same-level ghost fill (the RefineSchedule::fillData step before interp,
// LDataManager.cpp:742-745, restricted to copies between patches of one level):
// every ghost value of every patch array whose index -- wrapped across periodic
// directions -- lies in another patch's (or, periodically, its own) interior
// side box is copied from there.  Ghost values outside the domain of a
// non-periodic direction, or not covered by the level (coarse-fine boundary),
// are left as they are.
template <typename T>
struct FillArgs
{
    const int4* segs;
    const PatchDev* patches;
    T* const* arrays;
    const int* bin_patch_start;
    const int* bin_patches;
    int ndim, narr, side, depth;
    int gcw[3];
    int dom_lo[3], dom_hi[3];
    int pshift[3];
    int nb[3], nsb[3];
    int bin;
};

template <typename T>
__global__ __launch_bounds__(256) void k_fill_level(FillArgs<T> a)
{
    const int4 sg = a.segs[blockIdx.x];
    if ((int)threadIdx.x >= sg.w) return;
    const int pid = sg.x, arr_i = sg.y;
    const int axis = a.side ? arr_i : -1;
    const PatchDev P = a.patches[pid];
    int n[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
    for (int d = 0; d < a.ndim; ++d)
    {
        lo[d] = P.ilo[d] - a.gcw[d];
        n[d] = P.ihi[d] - P.ilo[d] + 1 + 2 * a.gcw[d] + (d == axis ? 1 : 0);
    }
    const long e = (long)sg.z + threadIdx.x;
    int idx[3] = {(int)(e % n[0]), (int)((e / n[0]) % n[1]), (int)(e / ((long)n[0] * n[1]))};
    bool interior = true;
    int w[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d)
    {
        if (d >= a.ndim)
            continue;
        const int i = lo[d] + idx[d];
        const int top = P.ihi[d] + (d == axis ? 1 : 0);
        interior = interior && i >= P.ilo[d] && i <= top;
        const int N = a.pshift[d];
        const int dtop = a.dom_hi[d] + (d == axis ? 1 : 0);
        if (N > 0)
            w[d] = a.dom_lo[d] + (((i - a.dom_lo[d]) % N) + N) % N;
        else
        {
            if (i < a.dom_lo[d] || i > dtop) return;  // physical boundary ghost: not filled here
            w[d] = i;
        }
    }
    if (interior) return;
    // owner: a patch whose interior side box holds the wrapped index; its cell,
    // or for a side axis also the cell below (the face is that cell's upper face)
    int q = -1;
    for (int tries = 0; tries < 2 && q < 0; ++tries)
    {
        if (tries == 1 && axis < 0) break;
        int c[3] = {0, 0, 0};
        bool ok = true;
        for (int d = 0; d < a.ndim; ++d)
        {
            c[d] = w[d] - (tries == 1 && d == axis ? 1 : 0);
            ok = ok && c[d] >= a.dom_lo[d] && c[d] <= a.dom_hi[d];
        }
        if (!ok) continue;
        int b[3] = {0, 0, 0};
        for (int d = 0; d < a.ndim; ++d) b[d] = (c[d] - a.dom_lo[d]) / a.bin;
        const int key = bin_mem(a.ndim, b[0], b[1], b[2], a.nb, a.nsb);
        for (int j = a.bin_patch_start[key]; j < a.bin_patch_start[key + 1] && q < 0; ++j)
        {
            const PatchDev& Q = a.patches[a.bin_patches[j]];
            bool in = true;
            for (int d = 0; d < a.ndim; ++d)
                in = in && w[d] >= Q.ilo[d] && w[d] <= Q.ihi[d] + (d == axis ? 1 : 0);
            if (in) q = a.bin_patches[j];
        }
    }
    if (q < 0) return;  // not covered by this level
    const PatchDev Q = a.patches[q];
    int qn[3] = {1, 1, 1};
    long src = 0, stride = 1;
    for (int d = 0; d < a.ndim; ++d)
    {
        qn[d] = Q.ihi[d] - Q.ilo[d] + 1 + 2 * a.gcw[d] + (d == axis ? 1 : 0);
        src += stride * (long)(w[d] - (Q.ilo[d] - a.gcw[d]));
        stride *= qn[d];
    }
    const long ptot = (long)n[0] * n[1] * n[2], qtot = stride;
    T* __restrict__ dst = a.arrays[pid * a.narr + arr_i];
    const T* __restrict__ from = a.arrays[q * a.narr + arr_i];
    const int ncomp = a.side ? 1 : a.depth;
    for (int cd = 0; cd < ncomp; ++cd) dst[(long)cd * ptot + e] = from[(long)cd * qtot + src];
}
} // namespace

// ----------------------------------------------------------------------------- physical boundary

namespace
{
// The LINEAR Robin ghost operators of CartCellRobinPhysBdryOp / CartSideRobinPhysBdryOp
// (cartphysbdryop{2,3}d.f.m4), adjoint (accumulateFromPhysicalBoundaryData) and
// forward (setPhysicalBoundaryConditions).  One workgroup per patch walks the
// patch's boundary boxes in order; inside an operator, threads take the index
// sets that the operator's updates keep fixed (the columns normal to a face,
// the index along an edge, ...) and each runs the reference's loop over the
// rest, so every value sees the reference's sequence of operations.  Corners
// (a few gcw^3 values) run on one thread.
template <typename T>
struct BdryArgs
{
    const int4* patches_list;
    const int* boxes;
    // Robin coefficients of codim-1 boxes: for box slot k and coefficient array j
    // (an axis of side data, a component of cell data), a, b and g over the box's
    // columns start at coef + coef_off[k * m + j], m = ndim (side) or depth (cell)
    const double* coef;
    const long* coef_off;
    const PatchDev* patches;
    T* const* arrays;
    int ndim, side, depth, narr, gcw, adjoint;
    double dx[3];
};

template <typename T>
struct BPatch
{
    int ndim, side, gcw;
    int ilo[3], ihi[3];
    long vol;  // values per depth component (cell data)
    T* arr[3];
    __device__ T* at(int a, int comp, const int* x) const
    {
        long off = 0, mul = 1;
        for (int d = 0; d < ndim; ++d)
        {
            const int n = ihi[d] - ilo[d] + 1 + 2 * gcw + ((side && d == a) ? 1 : 0);
            off += (long)(x[d] - (ilo[d] - gcw)) * mul;
            mul *= n;
        }
        return side ? arr[a] + off : arr[0] + (long)comp * vol + off;
    }
};

__device__ inline int bp_tangential_dev(int ndim, int dir, int* t)
{
    int m = 0;
    for (int d = ndim - 1; d >= 0; --d)
        if (d != dir) t[m++] = d;
    return m;
}

// column c of a face: coordinates of the tangential dims
__device__ inline void face_column(int ndim, int dir, const int* blo, const int* bhi, long c, int* x)
{
    int t[2];
    const int nt = bp_tangential_dev(ndim, dir, t);
    const int n_in = bhi[t[nt - 1]] - blo[t[nt - 1]] + 1;
    x[t[nt - 1]] = blo[t[nt - 1]] + (int)(c % n_in);
    if (nt > 1) x[t[0]] = blo[t[0]] + (int)(c / n_in);
}

__device__ inline long face_columns(int ndim, int dir, const int* blo, const int* bhi)
{
    long m = 1;
    for (int d = 0; d < ndim; ++d)
        if (d != dir) m *= (long)(bhi[d] - blo[d] + 1);
    return m;
}

template <typename T>
__device__ void bd_cc_op1(const BPatch<T>& P, int a, int comp, int loc, const int* dhi, const int* blo,
                          const int* bhi, const double* cf, int adjoint, const double* dx)
{
    const int ndim = P.ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = dx[loc / ndim];  // as the Fortran indexes it
    const int sgn = lower ? -1 : 1;
    const int ig = lower ? P.ilo[dir] - 1 : dhi[dir] + 1, ii = lower ? P.ilo[dir] : dhi[dir];
    const long ncol = face_columns(ndim, dir, blo, bhi);
    for (long c = threadIdx.x; c < ncol; c += blockDim.x)
    {
        int x[3] = {0, 0, 0};
        face_column(ndim, dir, blo, bhi, c, x);
        // column c of the coefficient box (the lower tangential axis fastest, as acoef(j,k))
        const double ca = cf[c], cb = cf[ncol + c], cg = cf[2 * ncol + c];
        for (int i = 0; i < P.gcw; ++i)
        {
            const double n = 1.0 + 2.0 * i;
            const double f_i = -(ca * n * h - 2.0 * cb) / (ca * n * h + 2.0 * cb);
            const double f_g = 2.0 * n * h / (ca * n * h + 2.0 * cb);
            int xg[3] = {x[0], x[1], x[2]}, xi[3] = {x[0], x[1], x[2]};
            xg[dir] = ig + sgn * i;
            xi[dir] = ii - sgn * i;
            if (adjoint)
            {
                const double u_g = (double)*P.at(a, comp, xg);
                T* t = P.at(a, comp, xi);
                *t = (T)((double)*t + f_i * u_g);
            }
            else
            {
                const double u_i = (double)*P.at(a, comp, xi);
                *P.at(a, comp, xg) = (T)(f_i * u_i + f_g * cg);
            }
        }
    }
}

template <typename T>
__device__ void bd_sc_op1(const BPatch<T>& P, int loc, const int* blo, const int* bhi, const double* cf, int adjoint,
                          const double* dx)
{
    const int ndim = P.ndim, dir = loc / 2, lower = (loc % 2) == 0;
    const double h = dx[loc / ndim];
    const int sgn = lower ? -1 : 1;
    const int ib = lower ? P.ilo[dir] : P.ihi[dir] + 1;
    const long ncol = face_columns(ndim, dir, blo, bhi);
    for (long c = threadIdx.x; c < ncol; c += blockDim.x)
    {
        int x[3] = {0, 0, 0};
        face_column(ndim, dir, blo, bhi, c, x);
        const double ca = cf[c], cb = cf[ncol + c], cg = cf[2 * ncol + c];
        const bool dirichlet = fabs(cb) < 1.0e-12;  // per face (scrobinphysbdryop1: abs(b) .lt. 1.d-12)
        int xb[3] = {x[0], x[1], x[2]};
        xb[dir] = ib;
        T* ub = P.at(dir, 0, xb);
        double u_b;
        if (dirichlet)
        {
            u_b = cg / ca;
            *ub = (T)u_b;
        }
        else
            u_b = (double)*ub;
        for (int i = 1; i <= P.gcw; ++i)
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
                const double u_g = (double)*P.at(dir, 0, xg);
                T* t = P.at(dir, 0, xi);
                *t = (T)((double)*t + f_i * u_g);
                *ub = (T)((double)*ub + f_b * u_g);
            }
            else
            {
                const double u_i = (double)*P.at(dir, 0, xi);
                *P.at(dir, 0, xg) = (T)(dirichlet ? f_i * u_i + f_b * u_b : f_i * u_i + f_b * u_b + f_g * cg);
            }
        }
    }
}

__device__ inline void bd_edge(int loc, int* e, int* n1, int* n2, int* up1, int* up2)
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

// cell data, 3-D edges: every update keeps the edge index (one thread each)
template <typename T>
__device__ void bd_cc_edge(const BPatch<T>& P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bd_edge(loc, &e, &n1, &n2, &u1, &u2);
    const int b1 = u1 ? P.ihi[n1] : P.ilo[n1], b2 = u2 ? P.ihi[n2] : P.ilo[n2];
    const int s1 = u1 ? 1 : -1, s2 = u2 ? 1 : -1;
    for (int ce = flo[e] + (int)threadIdx.x; ce <= fhi[e]; ce += blockDim.x)
        for (int c2 = flo[n2]; c2 <= fhi[n2]; ++c2)
            for (int c1 = flo[n1]; c1 <= fhi[n1]; ++c1)
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
                    const double Ug = (double)*P.at(0, comp, x);
                    T* t;
                    t = P.at(0, comp, xm), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xa), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xb), *t = (T)((double)*t + Ug);
                    t = P.at(0, comp, xc), *t = (T)((double)*t - Ug);
                    t = P.at(0, comp, xd), *t = (T)((double)*t - Ug);
                }
                else
                {
                    const double v = (double)*P.at(0, comp, xm) +
                                     ((double)*P.at(0, comp, xa) - (double)*P.at(0, comp, xc)) +
                                     ((double)*P.at(0, comp, xb) - (double)*P.at(0, comp, xd));
                    *P.at(0, comp, x) = (T)v;
                }
            }
}

// cell data, corners (codim == ndim): one thread
template <typename T>
__device__ void bd_cc_corner(const BPatch<T>& P, int comp, int loc, const int* flo, const int* fhi, int adjoint)
{
    if (threadIdx.x != 0) return;
    const int nd = P.ndim;
    int b[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        b[d] = up ? P.ihi[d] : P.ilo[d];
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
                int pk[3][3], mk[3][3];
                for (int d = 0; d < 3; ++d)
                    for (int o = 0; o < 3; ++o)
                    {
                        pk[d][o] = o == d ? x[o] : b[o];
                        mk[d][o] = o == d ? m[o] : b[o];
                    }
                if (adjoint)
                {
                    const double Ug = (double)*P.at(0, comp, x);
                    T* t = P.at(0, comp, m);
                    *t = (T)((double)*t + Ug);
                    for (int d = 0; d < nd; ++d) t = P.at(0, comp, pk[d]), *t = (T)((double)*t + Ug);
                    for (int d = 0; d < nd; ++d) t = P.at(0, comp, mk[d]), *t = (T)((double)*t - Ug);
                }
                else
                {
                    double v = (double)*P.at(0, comp, m);
                    for (int d = 0; d < nd; ++d)
                        v = v + ((double)*P.at(0, comp, pk[d]) - (double)*P.at(0, comp, mk[d]));
                    *P.at(0, comp, x) = (T)v;
                }
            }
}

// side data, 3-D edges: component n1 extrapolated along n2 (columns keep the edge
// index and n1), then component n2 along n1 (columns keep the edge index and n2)
template <typename T>
__device__ void bd_sc_edge(const BPatch<T>& P, int loc, const int* flo, const int* fhi, int adjoint)
{
    int e, n1, n2, u1, u2;
    bd_edge(loc, &e, &n1, &n2, &u1, &u2);
    int bd[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
    bd[n1] = u1 ? P.ihi[n1] : P.ilo[n1];
    bd[n2] = u2 ? P.ihi[n2] : P.ilo[n2];
    sh[n1] = u1 ? -1 : 1;
    sh[n2] = u2 ? -1 : 1;
    for (int part = 0; part < 2; ++part)
    {
        const int comp = part == 0 ? n1 : n2, other = part == 0 ? n2 : n1;
        int lo[3] = {flo[0], flo[1], flo[2]}, hi[3] = {fhi[0], fhi[1], fhi[2]};
        hi[comp] += 1;
        const int ne = hi[e] - lo[e] + 1, nc = hi[comp] - lo[comp] + 1;
        for (int c = threadIdx.x; c < ne * nc; c += blockDim.x)
        {
            int x[3];
            x[e] = lo[e] + c % ne;
            x[comp] = lo[comp] + c / ne;
            if (!(x[comp] < P.ilo[comp] || x[comp] > P.ihi[comp] + 1)) continue;
            for (int o = lo[other]; o <= hi[other]; ++o)
            {
                x[other] = o;
                const double del = (double)abs(o - bd[other]);
                int xb[3] = {x[0], x[1], x[2]}, xs[3] = {x[0], x[1], x[2]};
                xb[other] = bd[other];
                xs[other] = bd[other] + sh[other];
                if (adjoint)
                {
                    const double u_g = (double)*P.at(comp, 0, x);
                    T* t = P.at(comp, 0, xb);
                    *t = (T)((double)*t + (1.0 + del) * u_g);
                    t = P.at(comp, 0, xs);
                    *t = (T)((double)*t - del * u_g);
                }
                else
                    *P.at(comp, 0, x) =
                        (T)((1.0 + del) * (double)*P.at(comp, 0, xb) - del * (double)*P.at(comp, 0, xs));
            }
        }
        __syncthreads();
    }
}

// side data, corners: one thread
template <typename T>
__device__ void bd_sc_corner(const BPatch<T>& P, int loc, const int* flo, const int* fhi, int adjoint)
{
    if (threadIdx.x != 0) return;
    const int nd = P.ndim;
    int bd[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d)
    {
        const int up = (loc >> d) & 1;
        bd[d] = up ? P.ihi[d] : P.ilo[d];
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
                    if (!(x[comp] < P.ilo[comp] || x[comp] > P.ihi[comp] + 1)) continue;
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
                        const double u_g = (double)*P.at(comp, 0, x);
                        T* t = P.at(comp, 0, xb);
                        *t = (T)((double)*t + dsum * u_g);
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                t = P.at(comp, 0, xs);
                                *t = (T)((double)*t - del[o] * u_g);
                            }
                    }
                    else
                    {
                        double v = dsum * (double)*P.at(comp, 0, xb);
                        for (int o = 0; o < nd; ++o)
                            if (o != comp)
                            {
                                int xs[3] = {xb[0], xb[1], xb[2]};
                                xs[o] += sh[o];
                                v = v - del[o] * (double)*P.at(comp, 0, xs);
                            }
                        *P.at(comp, 0, x) = (T)v;
                    }
                }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_phys_bdry(BdryArgs<T> a)
{
    const int4 pl = a.patches_list[blockIdx.x];
    const int pid = pl.x;
    const PatchDev& Q = a.patches[pid];
    BPatch<T> P;
    P.ndim = a.ndim;
    P.side = a.side;
    P.gcw = a.gcw;
    long vol = 1;
    for (int d = 0; d < 3; ++d)
    {
        P.ilo[d] = d < a.ndim ? Q.ilo[d] : 0;
        P.ihi[d] = d < a.ndim ? Q.ihi[d] : 0;
        if (d < a.ndim) vol *= (long)(Q.ihi[d] - Q.ilo[d] + 1 + 2 * a.gcw);
    }
    P.vol = vol;
    for (int d = 0; d < 3; ++d) P.arr[d] = d < a.narr ? a.arrays[pid * a.narr + d] : nullptr;
    if (!a.side) P.arr[0] = a.arrays[pid];
    const int phases_adj[4] = {3, 2, 10, 11}, phases_fwd[4] = {11, 10, 2, 3};
    for (int ph = 0; ph < 4; ++ph)
    {
        const int phase = a.adjoint ? phases_adj[ph] : phases_fwd[ph];
        for (int k = pl.y; k < pl.y + pl.z; ++k)
        {
            const int* bx = a.boxes + 8 * k;
            const int codim = bx[0], loc = bx[1];
            const int lo[3] = {bx[2], bx[3], bx[4]}, hi[3] = {bx[5], bx[6], bx[7]};
            const int ncomp = a.side ? 1 : a.depth;
            bool did = false;
            if (phase == 3 && codim == 3 && a.ndim == 3)
            {
                for (int c = 0; c < ncomp; ++c)
                {
                    if (a.side)
                        bd_sc_corner(P, loc, lo, hi, a.adjoint);
                    else
                        bd_cc_corner(P, c, loc, lo, hi, a.adjoint);
                    __syncthreads();
                }
                did = true;
            }
            else if (phase == 2 && codim == 2)
            {
                for (int c = 0; c < ncomp; ++c)
                {
                    if (a.ndim == 2)
                    {
                        if (a.side)
                            bd_sc_corner(P, loc, lo, hi, a.adjoint);
                        else
                            bd_cc_corner(P, c, loc, lo, hi, a.adjoint);
                    }
                    else if (a.side)
                        bd_sc_edge(P, loc, lo, hi, a.adjoint);
                    else
                        bd_cc_edge(P, c, loc, lo, hi, a.adjoint);
                    __syncthreads();
                }
                did = true;
            }
            else if ((phase == 10 || phase == 11) && codim == 1)
            {
                const int dir = loc / 2;
                const double* q = a.coef;
                // coefficient arrays per box: one per axis (side data) or per component (cell data)
                const long* qo = a.coef_off + (long)k * (a.side ? a.ndim : a.depth);
                if (!a.side)
                {
                    if (phase == 10)
                        for (int c = 0; c < a.depth; ++c)
                        {
                            bd_cc_op1(P, 0, c, loc, P.ihi, lo, hi, q + qo[c], a.adjoint, a.dx);
                            __syncthreads();
                        }
                }
                else if (phase == 10)
                {
                    for (int ax = 0; ax < a.ndim; ++ax)
                    {
                        if (ax == dir) continue;
                        int shi[3] = {P.ihi[0], P.ihi[1], P.ihi[2]}, bhi[3] = {hi[0], hi[1], hi[2]};
                        shi[ax] += 1;
                        bhi[ax] += 1;
                        bd_cc_op1(P, ax, 0, loc, shi, lo, bhi, q + qo[ax], a.adjoint, a.dx);
                        __syncthreads();
                    }
                }
                else
                {
                    bd_sc_op1(P, loc, lo, hi, q + qo[dir], a.adjoint, a.dx);
                    __syncthreads();
                }
                did = true;
            }
            (void)did;
        }
    }
}
} // namespace
} // namespace ibamr

extern "C" {

int ibamr_plan_fill_periodic_ghosts(ibamr_plan* P, int slot, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->npatches != 1) return fail(IBAMR_ERR_INVALID, "periodic ghost fill needs a single-patch level");
    const PatchDev& pd = P->patches_h[0];
    for (int d = 0; d < P->ndim; ++d)
    {
        if (P->pshift[d] != pd.ihi[d] - pd.ilo[d] + 1 || pd.ilo[d] != P->dom_lo[d])
            return fail(IBAMR_ERR_INVALID, "periodic ghost fill needs a fully periodic domain covered by the patch");
    }
    // host copy of the pointer table: no device-to-host copy, so this can be captured
    const std::vector<void*>& ptrs = P->arrays_h[slot];
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ncomps = P->centering == IBAMR_SIDE ? 1 : P->depth;
    for (int a = 0; a < P->narr; ++a)
    {
        const int axis = P->centering == IBAMR_SIDE ? a : -1;
        int n[3] = {1, 1, 1}, g[3] = {0, 0, 0}, N[3] = {1, 1, 1};
        for (int d = 0; d < P->ndim; ++d)
        {
            n[d] = pd.ihi[d] - pd.ilo[d] + 1 + 2 * P->data_gcw[d] + (d == axis ? 1 : 0);
            g[d] = P->data_gcw[d];
            N[d] = pd.ihi[d] - pd.ilo[d] + 1;
        }
        const long total = (long)n[0] * n[1] * n[2];
        // the core along a side axis is the N faces [g, g+N): face N is the image of face 0
        const long sz = (long)(n[2] - N[2]) * n[0] * n[1];
        const long sy = (long)N[2] * (n[1] - N[1]) * n[0];
        const long sx = (long)N[2] * N[1] * (n[0] - N[0]);
        const long shell = sz + sy + sx;
        if (shell == 0) continue;
        for (int c = 0; c < ncomps; ++c)
        {
            const int B = 256;
            const long G = (shell + B - 1) / B;
            if (P->precision == IBAMR_FP64)
                hipLaunchKernelGGL(k_fill_periodic<double>, dim3((unsigned)G), dim3(B), 0, s,
                                   static_cast<double*>(ptrs[a]) + c * total, n[0], n[1], n[2], g[0], g[1], g[2],
                                   N[0], N[1], N[2], sz, sy, sx);
            else
                hipLaunchKernelGGL(k_fill_periodic<float>, dim3((unsigned)G), dim3(B), 0, s,
                                   static_cast<float*>(ptrs[a]) + c * total, n[0], n[1], n[2], g[0], g[1], g[2],
                                   N[0], N[1], N[2], sz, sy, sx);
            HIP_OK(hipGetLastError());
        }
    }
    return IBAMR_OK;
}

int ibamr_plan_fill_ghosts(ibamr_plan* P, int slot, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->nfill_segs == 0) return IBAMR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        FillArgs<T> a{};
        a.segs = P->fill_segs_d;
        a.patches = P->patches_d;
        a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
        a.bin_patch_start = P->bin_patch_start_d;
        a.bin_patches = P->bin_patches_d;
        a.ndim = P->ndim;
        a.narr = P->narr;
        a.side = P->centering == IBAMR_SIDE;
        a.depth = P->depth;
        for (int d = 0; d < 3; ++d)
        {
            a.gcw[d] = P->data_gcw[d];
            a.dom_lo[d] = P->dom_lo[d];
            a.dom_hi[d] = P->dom_hi[d];
            a.pshift[d] = P->pshift[d];
            a.nb[d] = P->nb[d];
            a.nsb[d] = P->nsb[d];
        }
        a.bin = P->bin;
        hipLaunchKernelGGL(k_fill_level<T>, dim3((unsigned)P->nfill_segs), dim3(256), 0, s, a);
        HIP_OK(hipGetLastError());
        return IBAMR_OK;
    };
    return P->precision == IBAMR_FP64 ? go((double)0) : go((float)0);
}

// Coefficient values of a codim-1 box for array j: its tangential cells, one more
// along j for a side array tangential to the box (the coefficient box's
// tangential extension, CartSideRobinPhysBdryOp.cpp:913).
static long bdry_ncoef(const ibamr_plan* P, const ibamr_bdry_box& b, int j)
{
    const int dir = b.location / 2;
    long m = 1;
    for (int d = 0; d < P->ndim; ++d)
        if (d != dir) m *= (long)(b.upper[d] - b.lower[d] + 1 + (P->centering == IBAMR_SIDE && d == j ? 1 : 0));
    return m;
}

// Boxes with uniform coefficients (abg: [nboxes][arrays][3]) or coefficients over
// each codim-1 box (coefs: per codim-1 box in the given order and per array, a, b
// and g over the box's columns); exactly one of abg and coefs is given.
static int set_physical_boundary(ibamr_plan* P, const ibamr_bdry_box* boxes, int nboxes, const double* abg,
                                 const double* coefs, long ncoefs, bool varying)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (nboxes < 0 || (nboxes > 0 && !boxes)) return fail(IBAMR_ERR_INVALID, "null boundary boxes");
    if (nboxes > 0 && !(varying ? (const void*)coefs : (const void*)abg))
        return fail(IBAMR_ERR_INVALID, "null boundary coefficients");
    for (int d = 1; d < P->ndim; ++d)
        if (P->data_gcw[d] != P->data_gcw[0])
            return fail(IBAMR_ERR_GHOST_WIDTH, "physical boundary operators need a uniform ghost width");
    if (P->centering == IBAMR_SIDE && P->depth != 1)
        return fail(IBAMR_ERR_INVALID, "physical boundary operators on side data of depth 1 only");
    const int g = P->data_gcw[0];
    // group by patch, keeping the given order inside each patch
    std::vector<std::vector<int>> per(P->npatches);
    for (int k = 0; k < nboxes; ++k)
    {
        const ibamr_bdry_box& b = boxes[k];
        if (b.patch < 0 || b.patch >= P->npatches)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad patch index");
        if (b.codim < 1 || b.codim > P->ndim)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad codimension");
        const int nloc = b.codim == 1 ? 2 * P->ndim : (b.codim == P->ndim ? (1 << P->ndim) : 12);
        if (b.location < 0 || b.location >= nloc)
            return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": bad location index");
        // every cell an operator touches must lie in the patch's ghost box
        const PatchDev& Q = P->patches_h[b.patch];
        for (int d = 0; d < P->ndim; ++d)
        {
            if (b.lower[d] > b.upper[d] || b.lower[d] < Q.ilo[d] - g || b.upper[d] > Q.ihi[d] + g)
                return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + " is not inside the ghost box");
        }
        if (b.codim == 1)
        {
            const int dir = b.location / 2;
            const int want = (b.location % 2) ? Q.ihi[dir] + 1 : Q.ilo[dir] - 1;
            if (b.lower[dir] != want || b.upper[dir] != want)
                return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": a codim-1 box is the cell "
                                               "layer just outside its face");
            for (int d = 0; d < P->ndim; ++d)
                if (d != dir && (b.lower[d] < Q.ilo[d] || b.upper[d] > Q.ihi[d]))
                    return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) + ": a codim-1 box spans at "
                                                   "most its face");
        }
        else
        {
            // codim 2 / ndim: outside the patch, within the ghost width, on the sides the location names
            int up[3] = {-1, -1, -1};  // -1: along the box (edge direction)
            if (b.codim == P->ndim)
                for (int d = 0; d < P->ndim; ++d) up[d] = (b.location >> d) & 1;
            else
            {
                const int l = b.location;
                if (l < 4)
                    up[1] = l & 1, up[2] = (l >> 1) & 1;
                else if (l < 8)
                    up[0] = (l - 4) >> 1, up[2] = (l - 4) & 1;
                else
                    up[0] = (l - 8) & 1, up[1] = (l - 8) >> 1;
            }
            for (int d = 0; d < P->ndim; ++d)
            {
                if (up[d] < 0) continue;
                const bool ok = up[d] ? b.lower[d] > Q.ihi[d] : b.upper[d] < Q.ilo[d];
                if (!ok)
                    return fail(IBAMR_ERR_INVALID, "boundary box " + std::to_string(k) +
                                                       ": an edge or corner box lies outside the patch on its sides");
            }
        }
        per[b.patch].push_back(k);
    }
    const int narr = P->centering == IBAMR_SIDE ? P->ndim : P->depth;
    // where box k's coefficients for array j start in the caller's coefs
    std::vector<long> uoff((size_t)nboxes * narr, 0);
    long need = 0;
    for (int k = 0; k < nboxes; ++k)
        if (boxes[k].codim == 1)
            for (int j = 0; j < narr; ++j)
            {
                uoff[(size_t)k * narr + j] = need;
                need += 3 * bdry_ncoef(P, boxes[k], j);
            }
    if (varying && ncoefs != need)
        return fail(IBAMR_ERR_INVALID, "boundary coefficients: " + std::to_string(ncoefs) + " values given, the boxes "
                                       "need " + std::to_string(need));
    std::vector<int4> pl;
    std::vector<int> bx;
    std::vector<double> cf;
    std::vector<long> cfo;
    for (int p = 0; p < P->npatches; ++p)
    {
        if (per[p].empty()) continue;
        pl.push_back(make_int4(p, (int)(bx.size() / 8), (int)per[p].size(), 0));
        for (int k : per[p])
        {
            const ibamr_bdry_box& b = boxes[k];
            const int v[8] = {b.codim, b.location, b.lower[0], b.lower[1], P->ndim == 3 ? b.lower[2] : 0,
                              b.upper[0], b.upper[1], P->ndim == 3 ? b.upper[2] : 0};
            bx.insert(bx.end(), v, v + 8);
            for (int j = 0; j < narr; ++j)
            {
                cfo.push_back((long)cf.size());
                if (b.codim != 1) continue;
                const long m = bdry_ncoef(P, b, j);
                if (varying)
                {
                    const double* u = coefs + uoff[(size_t)k * narr + j];
                    cf.insert(cf.end(), u, u + 3 * m);
                }
                else
                    for (int c = 0; c < 3; ++c) cf.insert(cf.end(), (size_t)m, abg[3 * ((long)k * narr + j) + c]);
            }
        }
    }
    dfree(P->bdry_patch_d);
    dfree(P->bdry_box_d);
    dfree(P->bdry_coef_d);
    dfree(P->bdry_coef_off_d);
    P->bdry_patch_d = nullptr;
    P->bdry_box_d = nullptr;
    P->bdry_coef_d = nullptr;
    P->bdry_coef_off_d = nullptr;
    P->nbdry_patches = (int)pl.size();
    if (pl.empty()) return IBAMR_OK;
    if (cf.empty()) cf.push_back(0.0);
    HIP_OK(hipMalloc(&P->bdry_patch_d, sizeof(int4) * pl.size()));
    HIP_OK(hipMalloc(&P->bdry_box_d, sizeof(int) * bx.size()));
    HIP_OK(hipMalloc(&P->bdry_coef_d, sizeof(double) * cf.size()));
    HIP_OK(hipMalloc(&P->bdry_coef_off_d, sizeof(long) * cfo.size()));
    HIP_OK(hipMemcpy(P->bdry_patch_d, pl.data(), sizeof(int4) * pl.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(P->bdry_box_d, bx.data(), sizeof(int) * bx.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(P->bdry_coef_d, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(P->bdry_coef_off_d, cfo.data(), sizeof(long) * cfo.size(), hipMemcpyHostToDevice));
    return IBAMR_OK;
}

int ibamr_plan_set_physical_boundary(ibamr_plan* P, const ibamr_bdry_box* boxes, int nboxes, const double* abg)
{
    return set_physical_boundary(P, boxes, nboxes, abg, nullptr, 0, false);
}

int ibamr_plan_set_physical_boundary_coefs(ibamr_plan* P, const ibamr_bdry_box* boxes, int nboxes,
                                           const double* coefs, long ncoefs)
{
    return set_physical_boundary(P, boxes, nboxes, nullptr, coefs, ncoefs, true);
}

static int phys_bdry_launch(ibamr_plan* P, int slot, int adjoint, void* stream)
{
    if (!P) return fail(IBAMR_ERR_INVALID, "null plan");
    if (slot < 0 || slot >= MAX_SLOTS || !P->arrays_d[slot]) return fail(IBAMR_ERR_INVALID, "slot not bound");
    if (P->nbdry_patches == 0) return IBAMR_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto go = [&](auto tag) -> int {
        using T = decltype(tag);
        BdryArgs<T> a{};
        a.patches_list = P->bdry_patch_d;
        a.boxes = P->bdry_box_d;
        a.coef = P->bdry_coef_d;
        a.coef_off = P->bdry_coef_off_d;
        a.patches = P->patches_d;
        a.arrays = reinterpret_cast<T* const*>(P->arrays_d[slot]);
        a.ndim = P->ndim;
        a.side = P->centering == IBAMR_SIDE;
        a.depth = P->depth;
        a.narr = P->narr;
        a.gcw = P->data_gcw[0];
        a.adjoint = adjoint;
        for (int d = 0; d < 3; ++d) a.dx[d] = P->dx[d];
        hipLaunchKernelGGL(k_phys_bdry<T>, dim3((unsigned)P->nbdry_patches), dim3(256), 0, s, a);
        HIP_OK(hipGetLastError());
        return IBAMR_OK;
    };
    return P->precision == IBAMR_FP64 ? go((double)0) : go((float)0);
}

int ibamr_plan_accumulate_physical_boundary(ibamr_plan* P, int slot, void* stream)
{
    return phys_bdry_launch(P, slot, 1, stream);
}

int ibamr_plan_fill_physical_boundary(ibamr_plan* P, int slot, void* stream)
{
    return phys_bdry_launch(P, slot, 0, stream);
}

} // extern "C"
