// le_spread3d.h -- 3-D spread kernels for gfx950 (included by le_spread.hip,
// which defines SpreadArgs, PatchDev and TileDesc).
//
// FP64 (k_spread3): owner-computes, register-resident.  One wavefront owns a
// column tile of one array: 8 x 8 (x, y) grid lines, one per lane, times ZC
// consecutive z values held in registers.  Points that can reach the tile are
// staged 64 at a time: each lane computes one point's weights (the reference's
// FP64 arithmetic) and writes the x and y weights, aligned to the tile's lanes
// and zero-padded, plus the z weights to LDS.  The staged points are then
// applied grouped by their first z index: a ballot per z start yields the
// points with that start, so every register index is a compile-time constant
// (no dynamic VGPR indexing).  No atomics: each grid value lives in one lane's
// register and receives its contributions in a fixed order -- the reference
// operation order per contribution (IB: (w0*(w1*(w2/vol)))*F, BSPLINE:
// (((w0*w1)*w2)*F)/vol) and sequential adds over chunks cut per bin -- so the
// result equals the reference loop run over the oracle's z-bucketed entry
// order, bitwise.
//
// FP32C (k_spread3_fx, below): point-major accumulation into an LDS tile of
// 64-bit fixed-point integers.
#pragma once

namespace
{
constexpr int S3_PAD = 9;    // staged weight row: the 8 tile lanes (odd stride: no bank conflicts)
constexpr int S3_MAXCOMB = 64;

template <int KER, int ZC>
__global__ __launch_bounds__(64) void k_spread3(SpreadArgs<double> a)
{
    using T = double;
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    __shared__ T s_rx[64 * S3_PAD];
    __shared__ T s_ry[64 * S3_PAD];
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
    const int x = lane & 7, y = lane >> 3;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;
    const int tsz[3] = {8, 8, ZC};
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

    // registers: the tile's values
    double acc[ZC];
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;
    // LDataManager::spread (interior mode) spreads into zeroed data and adds the
    // old values afterwards (LDataManager.cpp:594, 661-663); the raw
    // LEInteractor::spread (accumulate mode) adds each contribution to the
    // existing value (f.m4:1463)
    const bool from_old = a.mode == IBAMR_SPREAD_ACCUMULATE;
#pragma unroll
    for (int z = 0; z < ZC; ++z) acc[z] = 0.0;

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
    if (from_old && col_in)
    {
#pragma unroll
        for (int z = 0; z < ZC; ++z)
            if (flo[2] + z <= fhi[2]) acc[z] = arr[col_off + plane * (size_t)(flo[2] + z - alo[2])];
    }

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

        // chunks: 64-point blocks of each bin (a chunk never crosses a
        // combination).  The next chunk's point data is loaded while the
        // current one is applied.
        int cr = 0, crp = 0;
        auto normalize = [&]() {
            while (cr < ngc && crp >= s_cl[cr])
            {
                ++cr;
                crp = 0;
            }
        };
        auto lane_point = [&](int& k, int& oc) {
            k = -1;
            oc = 13;
            if (cr < ngc && crp + lane < s_cl[cr])
            {
                k = s_cs[cr] + crp + lane;
                oc = s_co[cr];
            }
        };
        normalize();
        bool cvalid = cr < ngc;
        int nk, noc;
        int4 nc4 = make_int4(0, 0, 0, 0);
        double nX[3] = {0.0, 0.0, 0.0};
        T nF = T(0);
        // the lane's point of the current chunk: position, F, cell
        auto load_point = [&]() {
            if (nk < 0) return;
            nc4 = a.cells[nk];
#pragma unroll
            for (int d = 0; d < 3; ++d) nX[d] = a.X[3 * (size_t)nk + d];
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
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        // X + Xshift, Xshift = (double)(offset) * dx  (LIndexSetData.cpp:116)
                        const double Xs = nX[d] + (double)(o[d] * a.pshift[d]) * a.dx[d];
                        bse[d] = P.ilo[d] + weights_1d_u<KER>(Xs, nX[d], xl[d], a.dx[d], w[d]);
                    }
                    const int dx0 = bse[0] - flo[0], dy0 = bse[1] - flo[1], zz = bse[2] - flo[2];
                    if (dx0 > -W && dx0 < 8 && dy0 > -W && dy0 < 8 && zz > -W && zz < ZC)
                    {
                        // rows hold the weight of each of the tile's 8 x (y) lanes, zero outside the stencil
                        T* rx = s_rx + lane * S3_PAD;
                        T* ry = s_ry + lane * S3_PAD;
#pragma unroll
                        for (int i = 0; i < 8; ++i)
                        {
                            rx[i] = T(0);
                            ry[i] = T(0);
                        }
#pragma unroll
                        for (int j = 0; j < W; ++j)
                        {
                            if ((unsigned)(dx0 + j) < 8u) rx[dx0 + j] = w[0][j];
                            if ((unsigned)(dy0 + j) < 8u) ry[dy0 + j] = w[1][j];
                            // IB kernels fold 1/vol into w2 exactly as the reference (f.m4:1435)
                            s_wz[lane * 8 + j] = KT::IB ? w[2][j] / vol : w[2][j];
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
            crp += 64;
            normalize();
            cvalid = cr < ngc;
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
                        const T wx = s_rx[j * S3_PAD + x];
                        const T wy = s_ry[j * S3_PAD + y];
                        const T wx2 = s_rx[j2 * S3_PAD + x];
                        const T wy2 = s_ry[j2 * S3_PAD + y];
                        T wz[W], wz2[W];
#pragma unroll
                        for (int i = 0; i < W; ++i)
                        {
                            wz[i] = s_wz[j * 8 + i];
                            wz2[i] = s_wz[j2 * 8 + i];
                        }
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
                *v = (overwrite || from_old) ? acc[z] : *v + acc[z];
            }
    }
}

// ----------------------------------------------------------------------------- FP32C: fixed-point spread

// FP32C spread (k_spread3_fx): point-major accumulation into an LDS tile of
// 64-bit integers.  A workgroup of FX_THREADS threads owns a TX x TY x TZ box
// of one array (side axis or depth component) of one patch.  It reads the
// points of the bins (with periodic images) that can reach the box, keeps
// those whose cell is within reach in a queue, and then each thread takes one
// queued point: FP64 position-to-index conversion and stencil selection (the
// reference's arithmetic, so placement is that of the FP64 form), FP32
// weights, and one LDS integer add per stencil value inside the box.
//
// Fixed point: a contribution c = wx*wy*wz*F/vol becomes the integer
// round(c * 2^Q) with one FP64 fused multiply-add against 1.5*2^52: the sum
// lands in [2^52, 2^53), where doubles are spaced by 1, so its bit pattern is
// that of 1.5*2^52 plus the rounded integer.  Those bit patterns are added
// with 64-bit integer atomics; 1.5*2^52 has no bit below bit 51, so the low 51
// bits of the total are the integer sum modulo 2^51, which is exact while
// |sum| < 2^50.  Q is one global scale per force component chosen so that
// (points that can reach one value) * max|F| * 2^Q / vol < 2^49 (every weight
// has magnitude <= 1).  Integer addition is associative: the totals do not
// depend on the order of the adds, the queue order or the tile shape, so
// results are bitwise reproducible without floating-point atomics.
constexpr int FX_THREADS = 256;
constexpr double FX_MAGIC = 6755399441055744.0;  // 1.5 * 2^52
constexpr int FX_SUM_BITS = 51;                  // bits of the sum kept below the magic's lowest set bit
constexpr int FX_BOUND_LOG2 = 49;                // bound of |sum| * 2^Q (one bit below 2^50)

// 2^Q / vol for force component c, from the plan's bound buffer: fx[0] = the
// largest number of points in one bin, fx[1 + c] = max |F_c| (float bits).
// npts_factor bounds the points that can reach one value by bins of that
// size.  Returns false when the force holds a non-finite value.
__device__ __forceinline__ bool fx_scale(const unsigned* __restrict__ fx, int c, double npts_factor, double vol,
                                         double* q_over_vol, double* inv_q)
{
    const unsigned fb = fx[1 + c];
    if (fb >= 0x7f800000u)
    {
        *q_over_vol = 0.0;
        *inv_q = 0.0;
        return false;
    }
    const double bound = (double)fx[0] * npts_factor * (double)__uint_as_float(fb) / vol;
    int e = 0;
    if (bound > 0.0) (void)frexp(bound, &e);  // bound < 2^e
    const int Q = FX_BOUND_LOG2 - e;
    *q_over_vol = ldexp(1.0 / vol, Q);
    *inv_q = ldexp(1.0, -Q);
    return true;
}

template <int KER, int TX, int TY, int TZ>
__global__ __launch_bounds__(FX_THREADS) void k_spread3_fx(SpreadArgs<float> a)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr int NT = FX_THREADS;
    constexpr int TV = TX * TY * TZ;
    __shared__ unsigned long long s_acc[TV];
    __shared__ int2 s_q[2 * NT];  // queued candidates: (sorted point, image offset code)
    __shared__ int s_qn;
    __shared__ int s_lb[3][16], s_lo[3][16], s_ln[3];
    __shared__ int s_cs[S3_MAXCOMB], s_co[S3_MAXCOMB], s_cp[S3_MAXCOMB + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;
    const int tsz[3] = {TX, TY, TZ};
    int alo[3], an[3], flo[3], fhi[3], ext[3];
    double xl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        alo[d] = P.ilo[d] - a.gcw[d];
        const int ahi = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
        an[d] = ahi - alo[d] + 1;
        flo[d] = alo[d] + td.t[d] * tsz[d];
        fhi[d] = min(flo[d] + tsz[d] - 1, ahi);
        if (a.mode == IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        {
            flo[d] = max(flo[d], P.ilo[d]);
            fhi[d] = min(fhi[d], P.ihi[d] + (d == axis ? 1 : 0));
        }
        ext[d] = fhi[d] - flo[d] + 1;
        xl[d] = P.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);  // LEInteractor.cpp:3689
    }
    if (ext[0] <= 0 || ext[1] <= 0 || ext[2] <= 0) return;  // tile wholly in the ghost region
    float* __restrict__ arr = a.arrays[td.patch * a.narr + (a.side ? td.comp : 0)];
    const size_t plane = (size_t)an[0] * an[1];
    if (!a.side) arr += (size_t)td.comp * plane * an[2];
    const int fcomp = td.comp;
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;

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
    // per-axis (bin, image offset) lists
    if (tid < 3)
    {
        const int d = tid;
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
                    s_lb[d][n] = b;
                    s_lo[d][n] = o;
                    ++n;
                }
            }
        }
        s_ln[d] = n;
    }
    if (tid == 0) s_qn = 0;
    for (int v = tid; v < TV; v += NT) s_acc[v] = 0ull;
    __syncthreads();
    const int nx = s_ln[0], ny = s_ln[1], nz = s_ln[2];
    const int ncomb = nx * ny * nz;

    double q_over_vol = 0.0, inv_q = 0.0;
    const bool finite = fx_scale(a.fx, fcomp, a.fx_factor, a.vol, &q_over_vol, &inv_q);

    // one queued point: its stencil, clipped to the tile, added into s_acc
    auto apply = [&](int k, int oc) {
        const int o[3] = {oc % 3 - 1, (oc / 3) % 3 - 1, oc / 9 - 1};
        float w[3][W];
        int r[3], lo[3], hi[3];
        bool ok = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double Xu = a.X[3 * (size_t)k + d];
            // X + Xshift, Xshift = (double)(offset) * dx  (LIndexSetData.cpp:116)
            const double Xs = Xu + (double)(o[d] * a.pshift[d]) * a.dx[d];
            r[d] = P.ilo[d] + weights_1d_u<KER>(Xs, Xu, xl[d], a.dx[d], w[d]) - flo[d];
            lo[d] = max(0, -r[d]);
            hi[d] = min(W, ext[d] - r[d]);
            ok = ok && lo[d] < hi[d];
        }
        if (!ok) return;
        const double Fq = (double)a.F[(size_t)a.fdepth * k + fcomp] * q_over_vol;
        double wzs[W], wyd[W];
#pragma unroll
        for (int j = 0; j < W; ++j)
        {
            wzs[j] = (double)w[2][j] * Fq;
            wyd[j] = (double)w[1][j];
        }
        unsigned long long* __restrict__ b = s_acc + ((r[2] * TY + r[1]) * TX + r[0]);
#pragma unroll
        for (int j = 0; j < W; ++j)
        {
            if (j < lo[1] || j >= hi[1]) continue;
#pragma unroll
            for (int i = 0; i < W; ++i)
            {
                if (i < lo[0] || i >= hi[0]) continue;
                const double A = (double)w[0][i] * wyd[j];  // exact: a product of two floats
#pragma unroll
                for (int z = 0; z < W; ++z)
                {
                    if (z < lo[2] || z >= hi[2]) continue;
                    const double t = __builtin_fma(A, wzs[z], FX_MAGIC);
                    atomicAdd(b + (z * TY + j) * TX + i, (unsigned long long)__double_as_longlong(t));
                }
            }
        }
    };

    for (int cb = 0; cb < ncomb; cb += S3_MAXCOMB)
    {
        __syncthreads();  // every wave is done with the previous group's prefix
        // ---- this group of combinations: ranges and their prefix
        {
            const int c = cb + lane;
            int st = 0, ln = 0, oc = 0;
            if (tid < 64 && c < ncomb)
            {
                const int kx = c % nx, ky = (c / nx) % ny, kz = c / (nx * ny);
                const int m = bin_mem(3, s_lb[0][kx], s_lb[1][ky], s_lb[2][kz], a.nb, a.nsb);
                st = a.bin_start[m];
                ln = a.bin_start[m + 1] - st;
                oc = (s_lo[0][kx] + 1) + 3 * (s_lo[1][ky] + 1) + 9 * (s_lo[2][kz] + 1);
            }
            int inc = ln;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1)
            {
                const int t = __shfl_up(inc, off);
                if (lane >= off) inc += t;
            }
            if (tid < S3_MAXCOMB)
            {
                s_cs[tid] = st;
                s_co[tid] = oc;
                s_cp[tid + 1] = inc;
            }
            if (tid == 0) s_cp[0] = 0;
        }
        __syncthreads();
        const int ngc = min(S3_MAXCOMB, ncomb - cb);
        const int gtotal = s_cp[ngc];
        for (int cq = 0; cq < gtotal; cq += NT)
        {
            // ---- filter: candidates whose cell (with its image offset) is within reach
            const int q = cq + tid;
            bool keep = false;
            int k = 0, oc = 13;
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
                const int4 c4 = a.cells[k];
                const int o[3] = {oc % 3 - 1, (oc / 3) % 3 - 1, oc / 9 - 1};
                const int cc[3] = {c4.x, c4.y, c4.z};
                keep = true;
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    const int ci = cc[d] + o[d] * a.pshift[d];
                    keep = keep && ci >= clo[d] && ci <= chi[d];
                }
            }
            const unsigned long long m = __ballot(keep);
            int base = 0;
            if (lane == 0 && m) base = atomicAdd(&s_qn, __popcll(m));
            base = __shfl(base, 0);
            if (keep) s_q[base + __popcll(m & ((1ull << lane) - 1ull))] = make_int2(k, oc);
            __syncthreads();
            // ---- a full queue: every thread applies one point
            const int qn = s_qn;
            __syncthreads();  // every wave has the count before a later pass adds to it
            if (qn >= NT)
            {
                // each thread takes entry tid and moves entry NT + tid down (the
                // only reader of slot tid is thread tid itself)
                const int2 e = s_q[tid];
                if (tid < qn - NT) s_q[tid] = s_q[NT + tid];
                if (tid == 0) s_qn = qn - NT;
                __syncthreads();  // the shifted queue and count before the next pass appends
                apply(e.x, e.y);
            }
        }
    }
    {
        const int qn = s_qn;
        if (tid < qn)
        {
            const int2 e = s_q[tid];
            apply(e.x, e.y);
        }
    }
    __syncthreads();

    // ---- write the tile: total = sign-extended low bits, times 2^-Q
    for (int v = tid; v < TV; v += NT)
    {
        const int x = v % TX, y = (v / TX) % TY, z = v / (TX * TY);
        if (x >= ext[0] || y >= ext[1] || z >= ext[2]) continue;
        const long long iv = (long long)(s_acc[v] << (64 - FX_SUM_BITS)) >> (64 - FX_SUM_BITS);
        double val = finite ? (double)iv * inv_q : __builtin_nan("");
        if (!finite && ncomb == 0) val = 0.0;
        float* dst = arr + (size_t)(flo[0] + x - alo[0]) + (size_t)an[0] * (size_t)(flo[1] + y - alo[1]) +
                     plane * (size_t)(flo[2] + z - alo[2]);
        if (overwrite)
            *dst = (float)val;
        else if (ncomb > 0)
            *dst = (float)((double)*dst + val);
    }
}


// 2.5-D form (k_spread3_fxz): a workgroup owns TX x TY grid lines of one array
// over a segment of z values and sweeps z through an LDS window of NZW planes.
// The z bins (with their periodic images) are taken in increasing z; the
// points of one bin layer reach at most bin + 2 reach consecutive planes, so
// every stencil fits the window whole and no point is clipped in z (only x and
// y are tested, once per stencil row).  Before the next layer, the planes no
// later point can reach are written out and the rest move down the window.
template <int KER, int TX, int TY, int NZW>
__global__ __launch_bounds__(FX_THREADS) void k_spread3_fxz(SpreadArgs<float> a, int seg)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr int NT = FX_THREADS;
    constexpr int PL = TX * TY;  // values per plane
    __shared__ unsigned long long s_acc[NZW * PL];
    __shared__ int2 s_q[2 * NT];  // queued candidates: (sorted point, image offset code)
    __shared__ int s_qn;
    __shared__ int s_lb[3][16], s_lo[3][16], s_ln[3];
    __shared__ int s_cs[S3_MAXCOMB], s_co[S3_MAXCOMB], s_cp[S3_MAXCOMB + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const TileDesc td = a.tiles[blockIdx.x];
    if (td.patch < 0) return;  // padding of the XCD-ordered work list
    const PatchDev P = a.patches[td.patch];
    const int axis = a.side ? td.comp : -1;
    const int tsz[3] = {TX, TY, seg};
    int alo[3], an[3], flo[3], fhi[3], ext[3];
    double xl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        alo[d] = P.ilo[d] - a.gcw[d];
        const int ahi = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0);
        an[d] = ahi - alo[d] + 1;
        flo[d] = alo[d] + td.t[d] * tsz[d];
        fhi[d] = min(flo[d] + tsz[d] - 1, ahi);
        if (a.mode == IBAMR_SPREAD_ACCUMULATE_INTERIOR)
        {
            flo[d] = max(flo[d], P.ilo[d]);
            fhi[d] = min(fhi[d], P.ihi[d] + (d == axis ? 1 : 0));
        }
        ext[d] = fhi[d] - flo[d] + 1;
        xl[d] = P.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);  // LEInteractor.cpp:3689
    }
    if (ext[0] <= 0 || ext[1] <= 0 || ext[2] <= 0) return;  // tile wholly in the ghost region
    float* __restrict__ arr = a.arrays[td.patch * a.narr + (a.side ? td.comp : 0)];
    const size_t plane = (size_t)an[0] * an[1];
    if (!a.side) arr += (size_t)td.comp * plane * an[2];
    const int fcomp = td.comp;
    const bool overwrite = a.mode == IBAMR_SPREAD_OVERWRITE;

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
    // per-axis (bin, image offset) lists; built with the offset outermost and the
    // bin increasing, so the z list runs in increasing image-shifted z
    if (tid < 3)
    {
        const int d = tid;
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
                    s_lb[d][n] = b;
                    s_lo[d][n] = o;
                    ++n;
                }
            }
        }
        s_ln[d] = n;
    }
    if (tid == 0) s_qn = 0;
    for (int v = tid; v < NZW * PL; v += NT) s_acc[v] = 0ull;
    __syncthreads();
    const int nx = s_ln[0], ny = s_ln[1], nz = s_ln[2];
    const int nxy = nx * ny;

    double q_over_vol = 0.0, inv_q = 0.0;
    const bool finite = fx_scale(a.fx, fcomp, a.fx_factor, a.vol, &q_over_vol, &inv_q);
    const bool touched = nx * ny * nz > 0;

    // window: plane 0 holds z = wb; planes below zf are written
    int wb = flo[2] - a.reach, zf = flo[2];
    // write planes [zf, upto) of the owned range, then move the window's base to nb
    auto advance = [&](int upto, int nb) {
        upto = min(upto, fhi[2] + 1);
        for (int v = tid; v < (upto - zf) * PL; v += NT)
        {
            const int z = zf + v / PL, xy = v % PL;
            const int x = xy % TX, y = xy / TX;
            if (x >= ext[0] || y >= ext[1]) continue;
            long long iv = 0;
            if (z >= wb && z < wb + NZW)
                iv = (long long)(s_acc[(z - wb) * PL + xy] << (64 - FX_SUM_BITS)) >> (64 - FX_SUM_BITS);
            double val = finite ? (double)iv * inv_q : __builtin_nan("");
            if (!finite && !touched) val = 0.0;
            float* dst = arr + (size_t)(flo[0] + x - alo[0]) + (size_t)an[0] * (size_t)(flo[1] + y - alo[1]) +
                         plane * (size_t)(z - alo[2]);
            if (overwrite)
                *dst = (float)val;
            else if (touched)
                *dst = (float)((double)*dst + val);
        }
        zf = max(zf, upto);
        const int sh = nb - wb;
        if (sh > 0)
        {
            __syncthreads();  // the writes above have read the planes that move
            // move planes [sh, NZW) down by sh (values read before any is written), zero the rest
            constexpr int PER = (NZW * PL + NT - 1) / NT;
            unsigned long long t[PER];
#pragma unroll
            for (int r = 0; r < PER; ++r)
            {
                const int v = tid + r * NT;
                t[r] = (v < NZW * PL && v + sh * PL < NZW * PL) ? s_acc[v + sh * PL] : 0ull;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < PER; ++r)
            {
                const int v = tid + r * NT;
                if (v < NZW * PL) s_acc[v] = t[r];
            }
            wb = nb;
        }
        __syncthreads();
    };

    // one queued point: its stencil, clipped in x and y to the tile, added into s_acc
    auto apply = [&](int k, int oc) {
        const int o[3] = {oc % 3 - 1, (oc / 3) % 3 - 1, oc / 9 - 1};
        float w[3][W];
        int r[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double Xu = a.X[3 * (size_t)k + d];
            // X + Xshift, Xshift = (double)(offset) * dx  (LIndexSetData.cpp:116)
            const double Xs = Xu + (double)(o[d] * a.pshift[d]) * a.dx[d];
            r[d] = P.ilo[d] + weights_1d_u<KER>(Xs, Xu, xl[d], a.dx[d], w[d]);
        }
        r[0] -= flo[0];
        r[1] -= flo[1];
        r[2] -= wb;
        // the window holds the whole z stencil of every point of the current layer
        if (r[2] < 0 || r[2] + W > NZW) return;
        const int lo0 = max(0, -r[0]), hi0 = min(W, ext[0] - r[0]);
        const int lo1 = max(0, -r[1]), hi1 = min(W, ext[1] - r[1]);
        if (lo0 >= hi0 || lo1 >= hi1) return;
        const double Fq = (double)a.F[(size_t)a.fdepth * k + fcomp] * q_over_vol;
        double wzs[W], wyd[W];
#pragma unroll
        for (int j = 0; j < W; ++j)
        {
            wzs[j] = (double)w[2][j] * Fq;
            wyd[j] = (double)w[1][j];
        }
        unsigned long long* __restrict__ b = s_acc + ((r[2] * TY + r[1]) * TX + r[0]);
#pragma unroll
        for (int j = 0; j < W; ++j)
        {
#pragma unroll
            for (int i = 0; i < W; ++i)
            {
                if (j < lo1 || j >= hi1 || i < lo0 || i >= hi0) continue;
                const double A = (double)w[0][i] * wyd[j];  // exact: a product of two floats
#pragma unroll
                for (int z = 0; z < W; ++z)
                {
                    const double t = __builtin_fma(A, wzs[z], FX_MAGIC);
                    atomicAdd(b + (z * TY + j) * TX + i, (unsigned long long)__double_as_longlong(t));
                }
            }
        }
    };

    for (int ez = 0; ez < nz; ++ez)
    {
        // this z layer's lowest reachable (image-shifted) cell; its points reach
        // planes from there - reach on, so the planes below are final
        const int vlo = max(a.dom_lo[2] + s_lb[2][ez] * a.bin + s_lo[2][ez] * a.pshift[2], clo[2]);
        if (ez == 0)
            wb = vlo - a.reach;  // the window is still empty: just place it
        else
            advance(vlo - a.reach, vlo - a.reach);
        for (int cb = 0; cb < nxy; cb += S3_MAXCOMB)
        {
            __syncthreads();  // every wave is done with the previous group's prefix
            // ---- this group of (x, y) combinations of the layer: ranges and their prefix
            {
                const int c = cb + lane;
                int st = 0, ln = 0, oc = 0;
                if (tid < 64 && c < nxy)
                {
                    const int kx = c % nx, ky = c / nx;
                    const int m = bin_mem(3, s_lb[0][kx], s_lb[1][ky], s_lb[2][ez], a.nb, a.nsb);
                    st = a.bin_start[m];
                    ln = a.bin_start[m + 1] - st;
                    oc = (s_lo[0][kx] + 1) + 3 * (s_lo[1][ky] + 1) + 9 * (s_lo[2][ez] + 1);
                }
                int inc = ln;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1)
                {
                    const int t = __shfl_up(inc, off);
                    if (lane >= off) inc += t;
                }
                if (tid < S3_MAXCOMB)
                {
                    s_cs[tid] = st;
                    s_co[tid] = oc;
                    s_cp[tid + 1] = inc;
                }
                if (tid == 0) s_cp[0] = 0;
            }
            __syncthreads();
            const int ngc = min(S3_MAXCOMB, nxy - cb);
            const int gtotal = s_cp[ngc];
            for (int cq = 0; cq < gtotal; cq += NT)
            {
                // ---- filter: candidates whose cell (with its image offset) is within reach
                const int q = cq + tid;
                bool keep = false;
                int k = 0, oc = 13;
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
                    const int4 c4 = a.cells[k];
                    const int o[3] = {oc % 3 - 1, (oc / 3) % 3 - 1, oc / 9 - 1};
                    const int cc[3] = {c4.x, c4.y, c4.z};
                    keep = true;
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        const int ci = cc[d] + o[d] * a.pshift[d];
                        keep = keep && ci >= clo[d] && ci <= chi[d];
                    }
                }
                const unsigned long long m = __ballot(keep);
                int base = 0;
                if (lane == 0 && m) base = atomicAdd(&s_qn, __popcll(m));
                base = __shfl(base, 0);
                if (keep) s_q[base + __popcll(m & ((1ull << lane) - 1ull))] = make_int2(k, oc);
                __syncthreads();
                // ---- a full queue: every thread applies one point
                const int qn = s_qn;
                __syncthreads();  // every wave has the count before a later pass adds to it
                if (qn >= NT)
                {
                    // each thread takes entry tid and moves entry NT + tid down (the
                    // only reader of slot tid is thread tid itself)
                    const int2 e = s_q[tid];
                    if (tid < qn - NT) s_q[tid] = s_q[NT + tid];
                    if (tid == 0) s_qn = qn - NT;
                    __syncthreads();  // the shifted queue and count before the next pass appends
                    apply(e.x, e.y);
                }
            }
        }
        // ---- the layer's remaining queued points
        {
            const int qn = s_qn;
            if (tid < qn)
            {
                const int2 e = s_q[tid];
                apply(e.x, e.y);
            }
            __syncthreads();
            if (tid == 0) s_qn = 0;
        }
    }
    advance(fhi[2] + 1, wb);
}

} // namespace
