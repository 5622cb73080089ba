// le_spread3d.h -- 3-D spread kernel for gfx950 (included by le_spread.hip,
// which defines SpreadArgs, PatchDev and TileDesc).
//
// Spread (k_spread3): owner-computes, register-resident.  One wavefront owns a
// column tile of one array: TX x TY (x, y) grid lines (8 x 8 or 16 x 4), one
// per lane, times ZC consecutive z values held in registers.  Points that can reach the
// tile are staged 64 at a time: each lane takes one point's stencil (FP64
// computes the reference's weights; FP32C loads the starts and weights the
// set-up pass computed once per point) and writes the x and y weights, aligned
// to the tile's lanes and zero-padded, plus the z weights to LDS.  The staged
// points are then applied grouped by their first z index: a ballot per z start
// yields the points with that start, so every register index is a compile-time
// constant (no dynamic VGPR indexing).  Per point and lane the work is two LDS
// reads, one product and W fused multiply-adds.  No atomics: each grid value
// lives in one lane's register and receives its contributions in a fixed
// order, so results are bitwise reproducible.
//   FP32C: each chunk's contributions go to FP32 partial sums, which are
//          folded into FP64 totals after the chunk (a chunk holds at most 64
//          points, so the FP32 sums stay short).
//   FP64:  the reference operation order per contribution
//          (IB: (w0*(w1*(w2/vol)))*F, BSPLINE: (((w0*w1)*w2)*F)/vol) and
//          sequential adds, over chunks cut per bin, so the result equals the
//          reference loop run over the oracle's z-bucketed entry order.
#pragma once

namespace
{
// staged weight rows: the tile's TX (x) or TY (y) lanes, padded to an odd
// stride (no bank conflicts between the rows of different points)
template <int N>
constexpr int s3_pad()
{
    return N + 1;
}
constexpr int S3_MAXCOMB = 64;

template <int KER, typename T, bool MERGE, int ZC, int TX>
__global__ __launch_bounds__(64) void k_spread3(SpreadArgs<T> a)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr bool FP64 = sizeof(T) == 8;
    // FP32C stages the set-up pass's stencils, except for kernels whose
    // stencil side depends on the unshifted X (an image's differs)
    constexpr bool PRECOMP = !FP64 && !side_by_unshifted_x<KER>();
    constexpr int TY = 64 / TX;  // the 64 lanes: TX x TY grid lines
    constexpr int PX = s3_pad<TX>(), PY = s3_pad<TY>();
    static_assert(TX * TY == 64, "one lane per grid line");
    __shared__ T s_rx[64 * PX];
    __shared__ T s_ry[64 * PY];
    __shared__ T s_wz[64 * 8];
    __shared__ T s_F[64];
    __shared__ int s_dx0[64];
    __shared__ int s_dy0[64];
    __shared__ int s_lb[3][16];
    __shared__ int s_lo[3][16];
    __shared__ int s_ln[3];
    __shared__ int s_cs[S3_MAXCOMB];    // per combination: first sorted point
    __shared__ int s_cl[S3_MAXCOMB];    //                  number of points
    __shared__ int s_co[S3_MAXCOMB];    //                  image offset code
    __shared__ int s_cp[S3_MAXCOMB + 1];// prefix of lengths

    const int lane = threadIdx.x;
    const int x = lane % TX, y = lane / TX;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;
    const int tsz[3] = {TX, TY, ZC};
    int alo[3], ahi[3], an[3], flo[3], fhi[3];
    double xl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        alo[d] = P.ilo[d] - a.gcw[d];
        ahi[d] = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
        an[d] = ahi[d] - alo[d] + 1;
        flo[d] = alo[d] + td.t[d] * tsz[d];
        fhi[d] = min(flo[d] + tsz[d] - 1, ahi[d]);
        if (a.mode == IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        {
            flo[d] = max(flo[d], P.ilo[d]);
            fhi[d] = min(fhi[d], P.ihi[d] + (d == axis ? 1 : 0));
        }
        xl[d] = P.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);  // LEInteractor.cpp:3689
    }
    if (flo[0] > fhi[0] || flo[1] > fhi[1] || flo[2] > fhi[2]) return;  // tile wholly in the ghost region
    T* __restrict__ arr = a.arrays[td.patch * a.narr + (a.side ? td.comp : 0)];
    const size_t plane = (size_t)an[0] * an[1];
    if (!a.side) arr += (size_t)td.comp * plane * an[2];
    const int fcomp = td.comp;
    const int gx = flo[0] + x, gy = flo[1] + y;
    const bool col_in = gx <= fhi[0] && gy <= fhi[1];
    const size_t col_off = (size_t)(gx - alo[0]) + (size_t)an[0] * (size_t)(gy - alo[1]);

    // cells whose points can reach the tile, inside the patch's Lagrangian ghost box
    int clo[3], chi[3];
    bool any = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        clo[d] = max(flo[d] - a.reach, P.ilo[d] - a.idx_gcw);
        chi[d] = min(fhi[d] + a.reach, P.ihi[d] + a.idx_gcw);
        any = any && clo[d] <= chi[d];
    }
    // per-axis (bin, image offset) lists in increasing (bin, offset) order
    if (lane < 3)
    {
        const int d = lane;
        int n = 0;
        if (any)
        {
            const int N = a.pshift[d];
            const int omax = N > 0 ? 1 : 0;
            for (int o = -omax; o <= omax; ++o)
            {
                const int pl = max(clo[d] - o * N, a.dom_lo[d]);
                const int ph = min(chi[d] - o * N, a.dom_hi[d]);
                if (pl > ph) continue;
                const int bl = floor_div(pl - a.dom_lo[d], a.bin);
                const int bh = floor_div(ph - a.dom_lo[d], a.bin);
                for (int b = bl; b <= bh && n < 16; ++b)
                {
                    int k = n++;
                    while (k > 0 && (s_lb[d][k - 1] > b || (s_lb[d][k - 1] == b && s_lo[d][k - 1] > o)))
                    {
                        s_lb[d][k] = s_lb[d][k - 1];
                        s_lo[d][k] = s_lo[d][k - 1];
                        --k;
                    }
                    s_lb[d][k] = b;
                    s_lo[d][k] = o;
                }
            }
        }
        s_ln[d] = n;
    }
    __syncthreads();
    const int nx = s_ln[0], ny = s_ln[1], nz = s_ln[2];
    const int ncomb = nx * ny * nz;

    // registers: tile values, accumulated in FP64 (FP32C: plus the chunk's
    // FP32 partial sums)
    double acc[ZC];
    T part[FP64 ? 1 : ZC];
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;
#pragma unroll
    for (int z = 0; z < ZC; ++z)
    {
        acc[z] = 0.0;
        if (!FP64) part[z] = T(0);
    }

    // total number of points in the visited bins; an empty tile keeps its values.
    // The lane's combination of the first group is kept for the chunk loop.
    int total = 0, g0_st = 0, g0_ln = 0, g0_oc = 0;
    for (int c = lane; c < ncomb; c += 64)
    {
        const int kx = c % nx, ky = (c / nx) % ny, kz = c / (nx * ny);
        const int m = bin_mem(3, s_lb[0][kx], s_lb[1][ky], s_lb[2][kz], a.nb, a.nsb);
        const int st = a.bin_start[m], ln = a.bin_start[m + 1] - st;
        total += ln;
        if (c == lane)
        {
            g0_st = st;
            g0_ln = ln;
            g0_oc = (s_lo[0][kx] + 1) + 3 * (s_lo[1][ky] + 1) + 9 * (s_lo[2][kz] + 1);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
    if (total == 0 && !overwrite) return;
    // LDataManager::spread (interior mode) spreads into zeroed data and adds the
    // old values afterwards (LDataManager.cpp:594, 661-663); the raw
    // LEInteractor::spread (accumulate mode) adds each contribution to the
    // existing value (f.m4:1463)
    const bool from_old = a.mode == IBAMR_SPREAD_ACCUMULATE;
    if (from_old && col_in)
    {
#pragma unroll
        for (int z = 0; z < ZC; ++z)
            if (flo[2] + z <= fhi[2]) acc[z] = (double)arr[col_off + plane * (size_t)(flo[2] + z - alo[2])];
    }

    const T inv_vol = T(1.0 / a.vol);
    const T vol = T(a.vol);
    for (int cb = 0; cb < ncomb && total > 0; cb += S3_MAXCOMB)
    {
        // ---- this group of combinations: ranges and their prefix (in canonical order)
        {
            const int c = cb + lane;
            int st = cb == 0 ? g0_st : 0, ln = cb == 0 ? g0_ln : 0, oc = cb == 0 ? g0_oc : 0;
            if (cb > 0 && lane < S3_MAXCOMB && c < ncomb)
            {
                const int kx = c % nx, ky = (c / nx) % ny, kz = c / (nx * ny);
                const int m = bin_mem(3, s_lb[0][kx], s_lb[1][ky], s_lb[2][kz], a.nb, a.nsb);
                st = a.bin_start[m];
                ln = a.bin_start[m + 1] - st;
                oc = (s_lo[0][kx] + 1) + 3 * (s_lo[1][ky] + 1) + 9 * (s_lo[2][kz] + 1);
            }
            // inclusive scan of ln across the wavefront
            int inc = ln;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1)
            {
                const int t = __shfl_up(inc, off);
                if (lane >= off) inc += t;
            }
            if (lane < S3_MAXCOMB)
            {
                s_cs[lane] = st;
                s_cl[lane] = ln;
                s_co[lane] = oc;
                s_cp[lane + 1] = inc;
            }
            if (lane == 0) s_cp[0] = 0;
        }
        __syncthreads();
        const int ngc = min(S3_MAXCOMB, ncomb - cb);
        const int gtotal = s_cp[ngc];

        // chunks: MERGE packs any 64 consecutive points of the group; otherwise a
        // chunk never crosses a combination (64-point blocks of each bin).  The
        // next chunk's point data is loaded while the current one is applied.
        int cq = 0, cr = 0, crp = 0;
        auto normalize = [&]() {
            if (!MERGE)
                while (cr < ngc && crp >= s_cl[cr])
                {
                    ++cr;
                    crp = 0;
                }
        };
        auto lane_point = [&](int& k, int& oc) {
            k = -1;
            oc = 13;
            if (MERGE)
            {
                const int q = cq + lane;
                if (q < gtotal)
                {
                    int lo = 0, hi = ngc - 1;  // last combination with prefix <= q
                    while (lo < hi)
                    {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_cp[mid] <= q)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    k = s_cs[lo] + (q - s_cp[lo]);
                    oc = s_co[lo];
                }
            }
            else if (cr < ngc && crp + lane < s_cl[cr])
            {
                k = s_cs[cr] + crp + lane;
                oc = s_co[cr];
            }
        };
        normalize();
        bool cvalid = MERGE ? cq < gtotal : cr < ngc;
        int nk, noc;
        int4 nc4 = make_int4(0, 0, 0, 0);
        double nX[3] = {0.0, 0.0, 0.0};
        T nF = T(0);
        int nst[3] = {0, 0, 0};   // FP32C: precomputed stencil starts and weights
        float nw[3][W];
        // the lane's point of the current chunk: position (FP64) or precomputed
        // starts and weights (FP32C, rows of the point's axis variants), F, cell
        auto load_point = [&]() {
            if (nk < 0) return;
            nc4 = a.cells[nk];
            if (!PRECOMP)
            {
#pragma unroll
                for (int d = 0; d < 3; ++d) nX[d] = a.X[3 * (size_t)nk + d];
            }
            else
            {
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    const int v = (a.nvar == 2 && d == axis) ? 1 : 0;
                    const size_t row = (size_t)(d * a.nvar + v) * a.npts + nk;
                    nst[d] = a.st[row];
#pragma unroll
                    for (int j = 0; j < W; ++j) nw[d][j] = a.wt[row * W + j];
                }
            }
            nF = a.F[(size_t)a.fdepth * nk + fcomp];
        };
        lane_point(nk, noc);
        load_point();
        while (cvalid)
        {
            const int k = nk, oc = noc;
            const int4 c4 = nc4;
            const T Fv = nF;

            int my_zz = 1 << 20;
            if (k >= 0)
            {
                const int o[3] = {oc % 3 - 1, (oc / 3) % 3 - 1, oc / 9 - 1};
                const int cc[3] = {c4.x, c4.y, c4.z};
                bool ok = true;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    const int ci = cc[d] + o[d] * a.pshift[d];
                    ok = ok && ci >= clo[d] && ci <= chi[d];
                }
                if (ok)
                {
                    T w[3][W];
                    int bse[3];
                    if (!PRECOMP)
                    {
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                        {
                            // X + Xshift, Xshift = (double)(offset) * dx  (LIndexSetData.cpp:116)
                            const double Xs = nX[d] + (double)(o[d] * a.pshift[d]) * a.dx[d];
                            bse[d] = P.ilo[d] + weights_1d_u<KER>(Xs, nX[d], xl[d], a.dx[d], w[d]);
                        }
                    }
                    else
                    {
                        // precomputed in the level's frame (k_point_weights); an image
                        // moves the stencil by whole periods
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                        {
                            bse[d] = nst[d] + o[d] * a.pshift[d];
#pragma unroll
                            for (int j = 0; j < W; ++j) w[d][j] = (T)nw[d][j];
                        }
                    }
                    const int dx0 = bse[0] - flo[0], dy0 = bse[1] - flo[1], zz = bse[2] - flo[2];
                    if (dx0 > -W && dx0 < TX && dy0 > -W && dy0 < TY && zz > -W && zz < ZC)
                    {
                        // rows hold the weight of each of the tile's 8 x (y) lanes, zero outside the stencil
                        T* rx = s_rx + lane * PX;
                        T* ry = s_ry + lane * PY;
#pragma unroll
                        for (int i = 0; i < TX; ++i) rx[i] = T(0);
#pragma unroll
                        for (int i = 0; i < TY; ++i) ry[i] = T(0);
                        const T fx = FP64 ? T(1) : (KT::IB ? Fv : Fv * inv_vol);
#pragma unroll
                        for (int j = 0; j < W; ++j)
                        {
                            if ((unsigned)(dx0 + j) < (unsigned)TX) rx[dx0 + j] = FP64 ? w[0][j] : w[0][j] * fx;
                            if ((unsigned)(dy0 + j) < (unsigned)TY) ry[dy0 + j] = w[1][j];
                            // FP64 IB folds 1/vol into w2 exactly as the reference (f.m4:1435)
                            s_wz[lane * 8 + j] = (FP64 && KT::IB) ? w[2][j] / vol
                                                                  : ((!FP64 && KT::IB) ? w[2][j] * inv_vol : w[2][j]);
                        }
                        s_F[lane] = Fv;
                        s_dx0[lane] = dx0;
                        s_dy0[lane] = dy0;
                        my_zz = zz;
                    }
                }
            }
            __syncthreads();
            // advance, and load the next chunk's points while this one is applied
            if (MERGE)
                cq += 64;
            else
            {
                crp += 64;
                normalize();
            }
            cvalid = MERGE ? cq < gtotal : cr < ngc;
            nk = -1;
            if (cvalid) lane_point(nk, noc);
            load_point();
            const bool chunk_any = __ballot(my_zz < (1 << 20)) != 0;
            if (chunk_any)
            {
#pragma unroll
                for (int zz = -(W - 1); zz < ZC; ++zz)
                {
                    unsigned long long m = __ballot(my_zz == zz);
                    while (m)
                    {
                        // two points per step so their LDS reads overlap
                        const int j = __builtin_ctzll(m);
                        m &= m - 1;
                        const int j2 = m ? __builtin_ctzll(m) : j;
                        const bool two = m != 0;
                        if (two) m &= m - 1;
                        const T wx = s_rx[j * PX + x];
                        const T wy = s_ry[j * PY + y];
                        const T wx2 = s_rx[j2 * PX + x];
                        const T wy2 = s_ry[j2 * PY + y];
                        T wz[W], wz2[W];
#pragma unroll
                        for (int i = 0; i < W; ++i)
                        {
                            wz[i] = s_wz[j * 8 + i];
                            wz2[i] = s_wz[j2 * 8 + i];
                        }
                        if (!FP64)
                        {
                            const T b = wx * wy;
                            const T b2 = two ? wx2 * wy2 : T(0);
#pragma unroll
                            for (int i = 0; i < W; ++i)
                                if (zz + i >= 0 && zz + i < ZC)
                                    part[zz + i] = fmaf(b2, wz2[i], fmaf(b, wz[i], part[zz + i]));
                        }
                        else
                        {
                            const bool inside = (unsigned)(x - s_dx0[j]) < (unsigned)W &&
                                                (unsigned)(y - s_dy0[j]) < (unsigned)W;
                            const bool inside2 = two && (unsigned)(x - s_dx0[j2]) < (unsigned)W &&
                                                 (unsigned)(y - s_dy0[j2]) < (unsigned)W;
                            const T Fj = s_F[j], Fj2 = s_F[j2];
                            if (inside)
                            {
#pragma unroll
                                for (int i = 0; i < W; ++i)
                                    if (zz + i >= 0 && zz + i < ZC)
                                    {
                                        const T c = KT::IB ? (wx * (wy * wz[i])) * Fj : (((wx * wy) * wz[i]) * Fj) / vol;
                                        acc[zz + i] = acc[zz + i] + c;
                                    }
                            }
                            if (inside2)
                            {
#pragma unroll
                                for (int i = 0; i < W; ++i)
                                    if (zz + i >= 0 && zz + i < ZC)
                                    {
                                        const T c = KT::IB ? (wx2 * (wy2 * wz2[i])) * Fj2
                                                           : (((wx2 * wy2) * wz2[i]) * Fj2) / vol;
                                        acc[zz + i] = acc[zz + i] + c;
                                    }
                            }
                        }
                    }
                }
                if (!FP64)
                {
                    // fold the chunk's partial sums into the FP64 totals
#pragma unroll
                    for (int z = 0; z < ZC; ++z)
                    {
                        acc[z] = acc[z] + (double)part[z];
                        part[z] = T(0);
                    }
                }
            }
            __syncthreads();
        }
    }

    if (col_in)
    {
#pragma unroll
        for (int z = 0; z < ZC; ++z)
            if (flo[2] + z <= fhi[2])
            {
                T* v = arr + col_off + plane * (size_t)(flo[2] + z - alo[2]);
                *v = (overwrite || from_old) ? (T)acc[z] : (T)((double)*v + acc[z]);
            }
    }
}


// ----------------------------------------------------------------------------- FP32C: z sweep

// FP32C spread swept along z (k_spread3s): a wavefront owns TX x TY grid lines
// of one array over a long z segment and takes the z layers of bins (with
// their periodic images) in increasing z.  The segment's values pass through
// a window of S3S_ZW z values held in registers: before a layer is applied the
// window is placed at the layer's first reachable cell minus the stencil
// reach, so every stencil of the layer lies in the window whole (no z
// clipping; its first z value is one of S3S_ZW - W + 1 window positions), and
// the values below the new window position, which no later layer reaches, are
// written out.  A point is then applied once per xy tile instead of once per
// (xy tile, z tile), and staged once per column segment.  The layers' points
// form one stream of 64-point chunks (a chunk never spans two layers); the
// next chunk's point data and the next layer's bin ranges are loaded while the
// current chunk is applied.  Staging, ballots and the FP32 partials folded into
// FP64 totals are those of k_spread3; each value's summation order is fixed by
// the bins, the points' order and the tile shape.
constexpr int S3S_ZW = 16;    // z values in registers
constexpr int S3S_MAXZ = 64;  // z layers (bin, image) of one segment

template <int KER, int TX>
__global__ __launch_bounds__(64) void k_spread3s(SpreadArgs<float> a, int seg)
{
    using T = float;
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr int TY = 64 / TX;
    constexpr int PX = s3_pad<TX>(), PY = s3_pad<TY>();
    constexpr int ZW = S3S_ZW;
    constexpr int NZZ = ZW - W + 1;  // window positions of a stencil's first z value
    static_assert(TX * TY == 64, "one lane per grid line");
    // staged rows of the chunk's 64 points, plus row 64 of zeros (the partner
    // of the last point of a z start with an odd count)
    __shared__ T s_rx[65 * PX];
    __shared__ T s_ry[65 * PY];
    __shared__ __align__(16) T s_wz[65 * 8];
    __shared__ int s_lb[2][16];
    __shared__ int s_lo[2][16];
    __shared__ int s_ln[3];
    __shared__ int s_zb[S3S_MAXZ];  // z layers: bin, image offset; increasing shifted z
    __shared__ int s_zo[S3S_MAXZ];
    __shared__ int s_cd[64];        // per (x, y) combination of the layer: first sorted point - first position
    __shared__ int s_sx[64];        //                                      image shift in x (cells)
    __shared__ int s_sy[64];        //                                      image shift in y

    const int lane = threadIdx.x;
    const int x = lane % TX, y = lane / TX;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    if (lane < PX) s_rx[64 * PX + lane] = T(0);
    if (lane < PY) s_ry[64 * PY + lane] = T(0);
    if (lane < 8) s_wz[64 * 8 + lane] = T(0);
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;
    const int tsz[3] = {TX, TY, seg};
    int alo[3], ahi[3], an[3], flo[3], fhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        alo[d] = P.ilo[d] - a.gcw[d];
        ahi[d] = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
        an[d] = ahi[d] - alo[d] + 1;
        flo[d] = alo[d] + td.t[d] * tsz[d];
        fhi[d] = min(flo[d] + tsz[d] - 1, ahi[d]);
        if (a.mode == IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        {
            flo[d] = max(flo[d], P.ilo[d]);
            fhi[d] = min(fhi[d], P.ihi[d] + (d == axis ? 1 : 0));
        }
    }
    if (flo[0] > fhi[0] || flo[1] > fhi[1] || flo[2] > fhi[2]) return;  // tile wholly in the ghost region
    T* __restrict__ arr = a.arrays[td.patch * a.narr + (a.side ? td.comp : 0)];
    const size_t plane = (size_t)an[0] * an[1];
    if (!a.side) arr += (size_t)td.comp * plane * an[2];
    const int fcomp = td.comp;
    const int gx = flo[0] + x, gy = flo[1] + y;
    const bool col_in = gx <= fhi[0] && gy <= fhi[1];
    const size_t col_off = (size_t)(gx - alo[0]) + (size_t)an[0] * (size_t)(gy - alo[1]);

    // cells whose points can reach the tile, inside the patch's Lagrangian ghost box
    int clo[3], chi[3];
    bool any = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        clo[d] = max(flo[d] - a.reach, P.ilo[d] - a.idx_gcw);
        chi[d] = min(fhi[d] + a.reach, P.ihi[d] + a.idx_gcw);
        any = any && clo[d] <= chi[d];
    }
    // x and y: (bin, image offset) lists in increasing (bin, offset) order; z:
    // layers with the offset outermost, so their shifted cells increase
    if (lane < 3)
    {
        const int d = lane;
        const int cap = d < 2 ? 16 : S3S_MAXZ;
        int n = 0;
        if (any)
        {
            const int N = a.pshift[d];
            const int omax = N > 0 ? 1 : 0;
            for (int o = -omax; o <= omax; ++o)
            {
                const int pl = max(clo[d] - o * N, a.dom_lo[d]);
                const int ph = min(chi[d] - o * N, a.dom_hi[d]);
                if (pl > ph) continue;
                const int bl = floor_div(pl - a.dom_lo[d], a.bin);
                const int bh = floor_div(ph - a.dom_lo[d], a.bin);
                for (int b = bl; b <= bh && n < cap; ++b)
                {
                    if (d == 2)
                    {
                        s_zb[n] = b;
                        s_zo[n] = o;
                        ++n;
                        continue;
                    }
                    int k = n++;
                    while (k > 0 && (s_lb[d][k - 1] > b || (s_lb[d][k - 1] == b && s_lo[d][k - 1] > o)))
                    {
                        s_lb[d][k] = s_lb[d][k - 1];
                        s_lo[d][k] = s_lo[d][k - 1];
                        --k;
                    }
                    s_lb[d][k] = b;
                    s_lo[d][k] = o;
                }
            }
        }
        s_ln[d] = n;
    }
    __syncthreads();
    const int nx = s_ln[0], ny = s_ln[1], nz = s_ln[2];
    const int nxy = nx * ny;  // at most 64 (the host admits only such shapes)

    double acc[ZW];
    T part[ZW];
#pragma unroll
    for (int z = 0; z < ZW; ++z)
    {
        acc[z] = 0.0;
        part[z] = T(0);
    }
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;
    // the raw LEInteractor::spread adds each contribution to the existing value
    // (f.m4:1463): the totals start from the old values; LDataManager::spread
    // (interior mode) spreads into zeroed data and adds the old values afterwards
    // (LDataManager.cpp:594, 661-663); overwrite replaces them
    const bool from_old = a.mode == IBAMR_SPREAD_ACCUMULATE;
    auto owned = [&](int z) { return col_in && z >= flo[2] && z <= fhi[2]; };
    auto at = [&](int z) { return arr + col_off + plane * (size_t)(z - alo[2]); };
    // the total a value starts from when it enters the window
    auto start = [&](int z) { return from_old && owned(z) ? (double)*at(z) : 0.0; };
    auto put = [&](int z, double v) {
        T* p = at(z);
        *p = (overwrite || from_old) ? (T)v : (T)((double)*p + v);
    };
    int wb = 0, zf = flo[2];  // window base; first value of the segment not yet written
    bool placed = false;
    // overwrite: values of the segment below z0 that were never in the window are zero
    auto zero_fill = [&](int z0) {
        if (overwrite && col_in)
            for (int z = zf; z < min(z0, fhi[2] + 1); ++z) *at(z) = T(0);
        zf = max(zf, z0);
    };
    auto load_window = [&]() {
#pragma unroll
        for (int i = 0; i < ZW; ++i) acc[i] = start(wb + i);
    };
    // moves the window base to nb (>= wb), writing the values that leave the window
    auto shift_to = [&](int nb) {
        if (!placed)
        {
            zero_fill(nb);
            wb = nb;
            placed = true;
            load_window();
            return;
        }
        const int sh = nb - wb;
        if (sh <= 0) return;
        if (sh >= ZW)
        {
#pragma unroll
            for (int i = 0; i < ZW; ++i)
                if (owned(wb + i) && wb + i >= zf) put(wb + i, acc[i]);
            zf = max(zf, wb + ZW);
            zero_fill(nb);
            wb = nb;
            load_window();
            return;
        }
        if (sh == 8)
        {
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (owned(wb + i) && wb + i >= zf) put(wb + i, acc[i]);
#pragma unroll
            for (int i = 0; i < ZW; ++i) acc[i] = i + 8 < ZW ? acc[i + 8] : start(nb + i);
        }
        else
            for (int k = 0; k < sh; ++k)
            {
                if (owned(wb + k) && wb + k >= zf) put(wb + k, acc[0]);
#pragma unroll
                for (int i = 0; i < ZW - 1; ++i) acc[i] = acc[i + 1];
                acc[ZW - 1] = start(wb + k + ZW);
            }
        zf = max(zf, nb);
        wb = nb;
    };

    // ---- the chunk stream
    // bin_start values of the lane's (x, y) combination in layer ez (loads in flight until used)
    auto layer_range = [&](int ez, int& st, int& en) {
        st = 0;
        en = 0;
        if (ez < nz && lane < nxy)
        {
            const int m = bin_mem(3, s_lb[0][lane % nx], s_lb[1][lane / nx], s_zb[ez], a.nb, a.nsb);
            st = a.bin_start[m];
            en = a.bin_start[m + 1];
        }
    };
    int ez = -1, gtotal = 0;
    int lex = 0;   // lane c: first position of combination c in the layer's candidates
    int cf = 0;    // the combination holding the chunk's first position
    int pst, pen;  // the following layer's ranges
    layer_range(0, pst, pen);
    // moves to the next layer with points: its combinations' offsets into LDS
    auto next_layer = [&]() -> bool {
        while (++ez < nz)
        {
            const int st = pst, ln = pen - pst;
            layer_range(ez + 1, pst, pen);
            int inc = ln;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1)
            {
                const int t = __shfl_up(inc, off);
                if (lane >= off) inc += t;
            }
            lex = inc - ln;
            s_cd[lane] = st - lex;
            s_sx[lane] = lane < nxy ? s_lo[0][lane % nx] * a.pshift[0] : 0;
            s_sy[lane] = lane < nxy ? s_lo[1][lane / nx] * a.pshift[1] : 0;
            __syncthreads();
            gtotal = __builtin_amdgcn_readlane(inc, 63);
            cf = 0;
            if (gtotal > 0) return true;
        }
        return false;
    };
    // the lane's point of the chunk at cq: its index, image code, cell, stencil and F
    int cq = 0;
    int nk = -1, nsx = 0, nsy = 0;
    int4 nc4 = make_int4(0, 0, 0, 0);
    int nst[3] = {0, 0, 0};
    T nw[3][W];
    T nF = T(0);
    auto fetch = [&]() {
        // the lane's combination: cf plus the combinations after cf that start at
        // or before its position (offsets increase with the combination)
        const int q = cq + lane;
        int comb = cf, ncf = cf;
        for (int c = cf + 1; c < nxy; ++c)
        {
            const int e = __builtin_amdgcn_readlane(lex, c);
            if (e > cq + 64) break;
            ncf = c;
            comb += e <= q ? 1 : 0;
        }
        cf = ncf;  // holds position cq + 64, the next chunk's first
        nk = -1;
        if (q >= gtotal) return;
        nk = q + s_cd[comb];
        nsx = s_sx[comb];
        nsy = s_sy[comb];
        nc4 = a.cells[nk];
        // precomputed in the level's frame (k_point_weights)
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const int v = (a.nvar == 2 && d == axis) ? 1 : 0;
            const size_t row = (size_t)(d * a.nvar + v) * a.npts + nk;
            nst[d] = a.st[row];
#pragma unroll
            for (int j = 0; j < W; ++j) nw[d][j] = a.wt[row * W + j];
        }
        nF = a.F[(size_t)a.fdepth * nk + fcomp];
    };
    bool have = next_layer();
    if (have) fetch();
    const T iv = T(1.0 / a.vol);
    int zsh = 0;  // the layer's image shift in z (cells)
    while (have)
    {
        // a layer's first chunk: place the window for it
        if (cq == 0)
        {
            zsh = s_zo[ez] * a.pshift[2];
            shift_to(max(a.dom_lo[2] + s_zb[ez] * a.bin + zsh, clo[2]) - a.reach);
        }
        // ---- stage this chunk
        int my_zz = 1 << 20;
        if (nk >= 0)
        {
            const int sh[3] = {nsx, nsy, zsh};
            const int cc[3] = {nc4.x, nc4.y, nc4.z};
            bool ok = true;
            int bse[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                const int ci = cc[d] + sh[d];
                ok = ok && ci >= clo[d] && ci <= chi[d];
                bse[d] = nst[d] + sh[d];  // an image moves the stencil by whole periods
            }
            const int dx0 = bse[0] - flo[0], dy0 = bse[1] - flo[1], zz = bse[2] - wb;
            if (ok && dx0 > -W && dx0 < TX && dy0 > -W && dy0 < TY && zz >= 0 && zz < NZZ)
            {
                // rows hold the weight of each of the tile's x (y) lanes, zero outside the stencil
                T* rx = s_rx + lane * PX;
                T* ry = s_ry + lane * PY;
#pragma unroll
                for (int i = 0; i < TX; ++i) rx[i] = T(0);
#pragma unroll
                for (int i = 0; i < TY; ++i) ry[i] = T(0);
                const T fx = KT::IB ? nF : nF * iv;
                T wz[8];
#pragma unroll
                for (int j = 0; j < W; ++j)
                {
                    if ((unsigned)(dx0 + j) < (unsigned)TX) rx[dx0 + j] = nw[0][j] * fx;
                    if ((unsigned)(dy0 + j) < (unsigned)TY) ry[dy0 + j] = nw[1][j];
                    wz[j] = KT::IB ? nw[2][j] * iv : nw[2][j];
                }
#pragma unroll
                for (int j = W; j < 8; ++j) wz[j] = T(0);
                float4* wq = reinterpret_cast<float4*>(s_wz + lane * 8);
                wq[0] = make_float4(wz[0], wz[1], wz[2], wz[3]);
                if (W > 4) wq[1] = make_float4(wz[4], wz[5], wz[6], wz[7]);
                my_zz = zz;
            }
        }
        __syncthreads();
        // ---- the next chunk (its layer's ranges are already loaded): loads in flight
        cq += 64;
        if (cq >= gtotal)
        {
            cq = 0;
            have = next_layer();
        }
        if (have) fetch();
        // ---- apply this chunk
        if (__ballot(my_zz < (1 << 20)))
        {
#pragma unroll
            for (int zz = 0; zz < NZZ; ++zz)
            {
                unsigned long long m = __ballot(my_zz == zz);
                while (m)
                {
                    // two points per step so their LDS reads overlap
                    const int j = __builtin_ctzll(m);
                    m &= m - 1;
                    const int j2 = m ? __builtin_ctzll(m) : 64;
                    m &= m - 1;
                    const T wx = s_rx[j * PX + x];
                    const T wy = s_ry[j * PY + y];
                    const T wx2 = s_rx[j2 * PX + x];
                    const T wy2 = s_ry[j2 * PY + y];
                    T wz[W], wz2[W];
#pragma unroll
                    for (int i = 0; i < W; ++i)
                    {
                        wz[i] = s_wz[j * 8 + i];
                        wz2[i] = s_wz[j2 * 8 + i];
                    }
                    const T b = wx * wy;
                    const T b2 = wx2 * wy2;
#pragma unroll
                    for (int i = 0; i < W; ++i) part[zz + i] = fmaf(b2, wz2[i], fmaf(b, wz[i], part[zz + i]));
                }
            }
            // fold the chunk's partial sums into the FP64 totals
#pragma unroll
            for (int z = 0; z < ZW; ++z)
            {
                acc[z] = acc[z] + (double)part[z];
                part[z] = T(0);
            }
        }
        __syncthreads();
    }
    // the rest of the segment: the window's values, then zeros above it (overwrite)
    if (placed)
    {
#pragma unroll
        for (int i = 0; i < ZW; ++i)
            if (owned(wb + i) && wb + i >= zf) put(wb + i, acc[i]);
        zf = max(zf, wb + ZW);
    }
    zero_fill(fhi[2] + 1);
}

} // namespace
