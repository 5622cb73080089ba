// le_interp3d.h -- 3-D interpolation kernel for gfx950 (included by
// le_interp.hip, which defines InterpArgs and Sum).
//
// Interpolation (k_interp3): one IBAMR_I3_THREADS-thread (512) workgroup per
// superbin (2x2x2 bins = 16^3 cells).  For each array it finds the exact stencil extent of its
// points, stages that block of the ghosted array into LDS with coalesced
// loads, then each thread sums the W^3 stencil of its points from LDS.
#pragma once

namespace
{
// ----------------------------------------------------------------------------- interpolation

#ifndef IBAMR_I3_THREADS
#define IBAMR_I3_THREADS 512
#endif
constexpr int I3_THREADS = IBAMR_I3_THREADS;  // threads per interpolation workgroup (one superbin)

// Largest staged block edge: a superbin's 16 cells plus the stencil reach on
// both sides (cell- and side-centred starts included).
template <int KER>
constexpr int interp3_edge()
{
    return 16 + KernelTraits<KER>::W;
}

template <int KER, typename T>
constexpr int interp3_lds_elems()
{
    return interp3_edge<KER>() * interp3_edge<KER>() * interp3_edge<KER>();
}

// Owning patch of a sorted point: the patch whose selection box holds its
// cell (LDataManager.cpp:748-753), or -1.
template <typename T>
__device__ __forceinline__ int interp_owner(const InterpArgs<T>& a, int k, const int4& c4)
{
    if (a.npatches == 1)
    {
        const PatchDev& Q = a.patches[0];
        return (c4.x >= Q.slo[0] && c4.x <= Q.shi[0] && c4.y >= Q.slo[1] && c4.y <= Q.shi[1] && c4.z >= Q.slo[2] &&
                c4.z <= Q.shi[2])
                   ? 0
                   : -1;
    }
    const int key = a.keys_sorted[k] >> a.ksub;
    for (int j = a.bin_patch_start[key]; j < a.bin_patch_start[key + 1]; ++j)
    {
        const PatchDev& Q = a.patches[a.bin_patches[j]];
        if (c4.x >= Q.slo[0] && c4.x <= Q.shi[0] && c4.y >= Q.slo[1] && c4.y <= Q.shi[1] && c4.z >= Q.slo[2] &&
            c4.z <= Q.shi[2])
            return a.bin_patches[j];
    }
    return -1;
}

template <int KER, typename T, bool LD>
__global__ __launch_bounds__(I3_THREADS) void k_interp3(InterpArgs<T> a, int lds_elems, int nsb, int nsx)
{
    using KT = KernelTraits<KER>;
    constexpr int W = KT::W;
    constexpr bool FP64 = sizeof(T) == 8;
    constexpr int UNIT = sizeof(T) / 4;  // dwords per value
    // FP32C IB kernels place stencils in the level's frame, as the spread set-up
    // does, and read its starts and weights when it computed them for these
    // positions (their weights, with square roots and divisions, cost more than
    // the loads).  BSPLINE and FP64 evaluate them here in the patch frame.
    constexpr bool PRE = !FP64 && KT::IB;
    constexpr bool pre_ld = PRE && LD;
    extern __shared__ __align__(16) unsigned char s_raw[];
    T* s_u = reinterpret_cast<T*>(s_raw);
    unsigned int* s_dw = reinterpret_cast<unsigned int*>(s_raw);
    // extents: [0] bases with the patch x_lower, [1] with x_lower - dx/2 (the side axis)
    __shared__ int s_lo[2][3], s_hi[2][3], s_pid, s_flag;

    // XCD-aware: workgroup i runs on XCD i % 8.  The rows of superbins along x
    // are dealt to the XCDs in turn (row r to XCD r % 8), each XCD taking its
    // rows in order: a row's superbins share their x halo in one L2, and every
    // XCD gets a share of every region of the point cloud.
    const int xi = (int)(blockIdx.x % 8), qi = (int)(blockIdx.x / 8);
    const int sb = (xi + 8 * (qi / nsx)) * nsx + qi % nsx;
    if (sb >= nsb) return;
    const int s0 = a.bin_start[sb * 8], s1 = a.bin_start[sb * 8 + 8];
    if (s0 == s1) return;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;

    const int owner = a.sb_owner ? a.sb_owner[sb] : -1;
    if (owner >= 0)
    {
        // a static superbin (one owning patch, blocks inside its ghosted arrays):
        // the blocks from the superbin's cells and the kernel's start offsets
        if (tid < 6)
        {
            const int v = tid / 3, d = tid % 3;
            const int sc = d == 0 ? sb % nsx : (d == 1 ? (sb / nsx) % a.nsy : sb / nsx / a.nsy);
            const int clo = a.lev_lo[d] + a.sbc * sc, chi = min(clo + a.sbc - 1, a.dom_hi[d]);
            s_lo[v][d] = clo + a.sb_off[v][0];
            s_hi[v][d] = chi + a.sb_off[v][1];
        }
        if (tid == 0)
        {
            s_pid = owner;
            s_flag = 0;
        }
    }
    else
    {
        if (tid < 6)
        {
            s_lo[tid / 3][tid % 3] = 1 << 30;
            s_hi[tid / 3][tid % 3] = -(1 << 30);
        }
        if (tid == 0)
        {
            s_pid = -1;
            s_flag = 0;
        }
        __syncthreads();
        // pass 1 (all arrays at once): owning patch and stencil bases of every
        // point -> block extents; flags: bit 0 several patches, bit 1+a array a
        // needs clipping (thread-local min/max, then one LDS atomic per wavefront)
        {
            int lmin[2][3], lmax[2][3];
#pragma unroll
            for (int v = 0; v < 2; ++v)
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    lmin[v][d] = 1 << 30;
                    lmax[v][d] = -(1 << 30);
                }
            int lpid = -1, lflag = 0;
            for (int k = s0 + tid; k < s1; k += I3_THREADS)
            {
                const int4 c4 = a.cells[k];
                const int pid = interp_owner(a, k, c4);
                if (pid < 0) continue;
                if (lpid >= 0 && lpid != pid) lflag |= 1;
                lpid = pid;
                const PatchDev P = a.patches[pid];
                int bc[3], bs[3];
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    if (!PRE)
                    {
                        T w[W];
                        const double X = a.X[3 * (size_t)k + d];
                        bc[d] = P.ilo[d] + weights_1d<KER>(X, P.xlo[d], a.dx[d], w);
                        bs[d] = a.side ? P.ilo[d] + weights_1d<KER>(X, P.xlo[d] - 0.5 * a.dx[d], a.dx[d], w) : bc[d];
                    }
                    else if (pre_ld)
                    {
                        // the starts in the level's frame (k_point_weights)
                        bc[d] = a.st[(size_t)(d * a.nvar) * a.n + k];
                        bs[d] = a.side ? a.st[(size_t)(d * a.nvar + 1) * a.n + k] : bc[d];
                    }
                    else
                    {
                        T w[W];
                        const double X = a.X[3 * (size_t)k + d];
                        bc[d] = a.lev_lo[d] + weights_1d<KER>(X, a.lev_xlo[d], a.dx[d], w);
                        bs[d] = a.side ? a.lev_lo[d] + weights_1d<KER>(X, a.lev_xlo[d] - 0.5 * a.dx[d], a.dx[d], w)
                                       : bc[d];
                    }
                    lmin[0][d] = min(lmin[0][d], bc[d]);
                    lmax[0][d] = max(lmax[0][d], bc[d]);
                    lmin[1][d] = min(lmin[1][d], bs[d]);
                    lmax[1][d] = max(lmax[1][d], bs[d]);
                }
                const int narr = a.side ? 3 : 1;
                for (int ax = 0; ax < narr; ++ax)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        const bool sh = a.side && d == ax;
                        const int b = sh ? bs[d] : bc[d];
                        const int alo = P.ilo[d] - a.gcw[d];
                        const int ahi = P.ihi[d] + a.gcw[d] + (sh ? 1 : 0);
                        if (b < alo || b + W - 1 > ahi) lflag |= 2 << ax;
                    }
            }
            int pmin = lpid < 0 ? (1 << 30) : lpid, pmax = lpid;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
            {
#pragma unroll
                for (int v = 0; v < 2; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        lmin[v][d] = min(lmin[v][d], __shfl_xor(lmin[v][d], off));
                        lmax[v][d] = max(lmax[v][d], __shfl_xor(lmax[v][d], off));
                    }
                pmin = min(pmin, __shfl_xor(pmin, off));
                pmax = max(pmax, __shfl_xor(pmax, off));
                lflag |= __shfl_xor(lflag, off);
            }
            if (lane == 0)
            {
#pragma unroll
                for (int v = 0; v < 2; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        atomicMin(&s_lo[v][d], lmin[v][d]);
                        atomicMax(&s_hi[v][d], lmax[v][d]);
                    }
                if (pmax >= 0)
                {
                    if (pmin != pmax) lflag |= 1;
                    const int prev = atomicCAS(&s_pid, -1, pmin);
                    if (prev != -1 && prev != pmin) lflag |= 1;
                }
                if (lflag) atomicOr(&s_flag, lflag);
            }
        }
    }
    __syncthreads();
    const int pid = s_pid;
    if (pid < 0) return;  // no point of this superbin lies in a patch
    const int flag = s_flag;
    const PatchDev P = a.patches[pid];

    const int narrays = a.side ? 3 : 1;
    for (int ax = 0; ax < narrays; ++ax)
    {
        const int axis = a.side ? ax : -1;
        int alo[3], an[3], blo[3], e[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const int v = d == axis ? 1 : 0;
            alo[d] = P.ilo[d] - a.gcw[d];
            an[d] = P.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0) - alo[d] + 1;
            blo[d] = s_lo[v][d];
            e[d] = s_hi[v][d] - s_lo[v][d] + W;
        }
        const int total = e[0] * e[1] * e[2];
        const size_t plane = (size_t)an[0] * an[1];
        // staged: one patch, no clipping, the block fits the LDS, and planes below
        // 2^24 values (the block's offsets then fit 24-bit products and 32 bits)
        const bool staged = (flag & (1 | (2 << ax))) == 0 && total <= lds_elems && plane < ((size_t)1 << 24);
        const int ncomps = a.side ? 1 : a.depth;
        for (int cd = 0; cd < ncomps; ++cd)
        {
            const T* __restrict__ u = a.arrays[pid * a.narr + (a.side ? ax : 0)] + (size_t)cd * plane * an[2];
            if (staged)
            {
                // Stage the block straight from global memory into LDS (global_load_lds:
                // no registers in between, every load in flight at once).  Dword q of
                // the block goes to s_dw[q]; each wave covers 64 consecutive dwords per
                // instruction, so the LDS destination is wave-uniform.  Value v of the
                // block, (v % e0, (v / e0) % e1, v / (e0 e1)), lies at
                // v + (an0 - e0) (v / e0) + (plane - e1 an0) (v / (e0 e1)) from the
                // block's first value.  The quotients are floor((v + 1/2) * fl(1/d)) in
                // FP32: (v + 1/2) / d is at least 1/(2d) from an integer and the
                // computed product is within (v / d) 2^-23 of it, so the floor is exact
                // for v < 2^22 (the block holds at most lds_elems < 2^14 values).
                const int ndw = total * UNIT;
                const float r0 = 1.0f / (float)e[0], r1 = 1.0f / (float)(e[0] * e[1]);
                const unsigned c0 = (unsigned)(an[0] - e[0]);
                const unsigned c1 = (unsigned)(plane - (size_t)e[1] * an[0]);
                const unsigned int* ub = reinterpret_cast<const unsigned int*>(
                    u + (size_t)(blo[0] - alo[0]) + (size_t)an[0] * (blo[1] - alo[1]) + plane * (size_t)(blo[2] - alo[2]));
                const int w0 = __builtin_amdgcn_readfirstlane(wv) * 64;
                for (int base = w0; base < ndw; base += I3_THREADS)
                {
                    const unsigned q = (unsigned)(base + lane);
                    if (q < (unsigned)ndw)
                    {
                        const unsigned v = q / UNIT;
                        const float fv = (float)v + 0.5f;
                        const unsigned off = v + __umul24(c0, (unsigned)(fv * r0)) + __umul24(c1, (unsigned)(fv * r1));
                        __builtin_amdgcn_global_load_lds(ub + (off * UNIT + q % UNIT), s_dw + base, 4, 0, 0);
                    }
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
            for (int k = s0 + tid; k < s1; k += I3_THREADS)
            {
                const int4 c4 = a.cells[k];
                // the point's own patch (the superbin may straddle patches: then no staging)
                int qid;
                if (staged)  // one owning patch for the whole superbin; points outside it are skipped
                    qid = (c4.x >= P.slo[0] && c4.x <= P.shi[0] && c4.y >= P.slo[1] && c4.y <= P.shi[1] &&
                           c4.z >= P.slo[2] && c4.z <= P.shi[2])
                              ? pid
                              : -1;
                else
                    qid = interp_owner(a, k, c4);
                if (qid < 0) continue;
                const PatchDev Pq = a.patches[qid];
                const int p = a.perm[k];
                T w[3][W];
                int bse[3];
#pragma unroll
                for (int d = 0; d < 3; ++d)
                {
                    if (!PRE)
                    {
                        const double xl = Pq.xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);
                        bse[d] = Pq.ilo[d] + weights_1d<KER>(a.X[3 * (size_t)k + d], xl, a.dx[d], w[d]);
                    }
                    else if (pre_ld)
                    {
                        // precomputed in the level's frame, as the spread uses them
                        const size_t row = (size_t)(d * a.nvar + (a.side && d == axis ? 1 : 0)) * a.n + k;
                        bse[d] = a.st[row];
#pragma unroll
                        for (int j = 0; j < W; ++j) w[d][j] = (T)a.wt[row * W + j];
                    }
                    else
                    {
                        const double xl = a.lev_xlo[d] - (d == axis ? 0.5 * a.dx[d] : 0.0);
                        bse[d] = a.lev_lo[d] + weights_1d<KER>(a.X[3 * (size_t)k + d], xl, a.dx[d], w[d]);
                    }
                }
                Sum<T> acc;
                // a static superbin's block holds the stencils of positions
                // inside its cells up to rounding at cell boundaries: others are
                // read from the array
                bool in_block = staged;
#pragma unroll
                for (int d = 0; d < 3; ++d) in_block = in_block && (unsigned)(bse[d] - blo[d]) <= (unsigned)(e[d] - W);
                if (in_block)
                {
                    const T* base = s_u + (bse[0] - blo[0]) + e[0] * ((bse[1] - blo[1]) + e[1] * (bse[2] - blo[2]));
#pragma unroll
                    for (int i2 = 0; i2 < W; ++i2)
                    {
                        // FP32C: rows as FMA chains, a plain sum over the plane's rows,
                        // one compensated add per plane
                        T pl = T(0);
#pragma unroll
                        for (int i1 = 0; i1 < W; ++i1)
                        {
                            const T* row = base + e[0] * (i1 + e[1] * i2);
                            if (FP64)
                            {
#pragma unroll
                                for (int i0 = 0; i0 < W; ++i0)
                                    acc.add(KT::IB ? (w[0][i0] * (w[1][i1] * w[2][i2])) * row[i0]
                                                   : ((w[0][i0] * w[1][i1]) * w[2][i2]) * row[i0]);
                            }
                            else
                            {
                                T sum = T(0);
#pragma unroll
                                for (int i0 = 0; i0 < W; ++i0) sum = fmaf(w[0][i0], row[i0], sum);
                                pl = fmaf(w[1][i1], sum, pl);
                            }
                        }
                        if (!FP64) acc.add(w[2][i2] * pl);
                    }
                }
                else
                {
                    // clipped stencil straight from the point's ghosted array (f.m4:1309-1314)
                    int lo[3], hi[3], qlo[3], qn[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                    {
                        qlo[d] = Pq.ilo[d] - a.gcw[d];
                        qn[d] = Pq.ihi[d] + a.gcw[d] + (d == axis ? 1 : 0) - qlo[d] + 1;
                        lo[d] = max(bse[d], qlo[d]) - bse[d];
                        hi[d] = min(bse[d] + W - 1, qlo[d] + qn[d] - 1) - bse[d];
                    }
                    const size_t qplane = (size_t)qn[0] * qn[1];
                    const T* __restrict__ uq =
                        a.arrays[qid * a.narr + (a.side ? ax : 0)] + (size_t)cd * qplane * qn[2];
#pragma unroll
                    for (int i2 = 0; i2 < W; ++i2)
                    {
                        if (i2 < lo[2] || i2 > hi[2]) continue;
#pragma unroll
                        for (int i1 = 0; i1 < W; ++i1)
                        {
                            if (i1 < lo[1] || i1 > hi[1]) continue;
                            const T* row = uq + (size_t)(bse[0] - qlo[0]) + (size_t)qn[0] * (bse[1] + i1 - qlo[1]) +
                                           qplane * (size_t)(bse[2] + i2 - qlo[2]);
                            if (FP64)
                            {
#pragma unroll
                                for (int i0 = 0; i0 < W; ++i0)
                                    if (i0 >= lo[0] && i0 <= hi[0])
                                        acc.add(KT::IB ? (w[0][i0] * (w[1][i1] * w[2][i2])) * row[i0]
                                                       : ((w[0][i0] * w[1][i1]) * w[2][i2]) * row[i0]);
                            }
                            else
                            {
                                T sum = T(0);
#pragma unroll
                                for (int i0 = 0; i0 < W; ++i0)
                                    if (i0 >= lo[0] && i0 <= hi[0]) sum = fmaf(w[0][i0], row[i0], sum);
                                acc.add((w[1][i1] * w[2][i2]) * sum);
                            }
                        }
                    }
                }
                a.U[(size_t)a.fdepth * p + (a.side ? ax : cd)] = acc.get();
            }
            __syncthreads();
        }
    }
}
} // namespace
