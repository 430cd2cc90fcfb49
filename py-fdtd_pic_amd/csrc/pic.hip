// Particle-in-cell kernels for gfx950: relativistic Boris push with linear
// gather, then a deterministic stable re-sort by cell fused with the CIC
// current deposit (segmented FP64 sums, no floating-point atomics).
//
// The reference has no particle code; the current enters the grid through
// V.Jx in ADE_ExUpdate (BaseFDTD11.py:663-669) as Ex += (dHy - Jx)*UpExMat.
// Units are the normalised ones of the field kernel: E, H' = Z0*Hy, and the
// deposit writes J' = Z0*Jx (J_x*dz in the reference's units).
//
// Step n, before the field launch of step n (which consumes J^{n+1/2}):
//   pic_push          one thread per particle (sorted order): gather E^n at
//                     x^n and B^n = -(H^{n-1/2} + H^{n+1/2})/(2c) at x^n - 1/2,
//                     Boris, x^{n+1} = x^n + Sc u_z/gamma; writes the new state
//                     and the CIC weights of J' at x^{n+1/2} to records at the
//                     particle's old index.
//   pic_sort_deposit  one wave per G cells: counts of particles moving down /
//                     staying / moving up per cell (|d| <= 1 since Sc*beta < 1),
//                     stable ranks, the new index
//                        d=-1: start[c] - up(c-1) + r
//                        d= 0: start[c] + down(c) + r
//                        d=+1: start[c+1] - up(c) + down(c+1) + r
//                     (no global scan: only neighbour cells enter), the new
//                     start[] of its cells, and J'[m] = S+2(m-2) + S+1(m-1) +
//                     S0(m) + S-1(m+1) from per-cell FP64 sums in a fixed
//                     order.  The wave reads cells c0-2 .. c0+G so every value
//                     it needs is its own.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

__device__ __forceinline__ unsigned long long lanes_le(int lane) {
    return lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
}
__device__ __forceinline__ unsigned long long lanes_lt(int lane) { return (1ull << lane) - 1ull; }

template <typename R>
__global__ __launch_bounds__(256) void pic_push(const SimDev<R>* __restrict__ sims,
                                                const BeamDev<R>* __restrict__ beams,
                                                const int2* __restrict__ blocks, long long launch_idx) {
    const int2 bl = blocks[blockIdx.x];
    const BeamDev<R>& b = beams[bl.x];
    const int i = bl.y + (int)threadIdx.x;
    if (i >= b.np) return;
    const SimDev<R>& sd = sims[b.sim];
    const int par = (int)((launch_idx - sd.launch_begin) & 1);
    const int N = b.N;
    int c = b.cell[par][i];
    R f = b.frac[par][i];
    R ux = b.ux[par][i], uy = b.uy[par][i], uz = b.uz[par][i];
    int code = 1 | (1 << 2);            // d = 0, node offset 0
    R a0 = R(0), a1 = R(0);
    if (c >= 2 && c <= N - 3) {
        const R* E = sd.E[par];
        const R* Hc = sd.H[par];
        const R* Hp = sd.H[par ^ 1];
        const R Eg = (R(1) - f) * E[c] + f * E[c + 1];
        int j;
        R g;
        if (f >= R(0.5)) { j = c; g = f - R(0.5); } else { j = c - 1; g = f + R(0.5); }
        const R h0 = R(0.5) * (Hp[j] + Hc[j]);
        const R h1 = R(0.5) * (Hp[j + 1] + Hc[j + 1]);
        const R Hg = (R(1) - g) * h0 + g * h1;
        // Boris: E = (Eg, 0, 0), B = (0, -Hg/c, 0)
        const R aE = b.a * Eg;
        const R umx = ux + aE, umy = uy, umz = uz;
        const R gm = sqrt(R(1) + umx * umx + umy * umy + umz * umz);
        const R ty = -b.a * Hg / gm;
        const R sy = R(2) * ty / (R(1) + ty * ty);
        const R upx = umx - umz * ty;
        const R upz = umz + umx * ty;
        ux = umx - upz * sy + aE;
        uz = umz + upx * sy;
        uy = umy;
        const R gam = sqrt(R(1) + ux * ux + uy * uy + uz * uz);
        const R bx = ux / gam, bz = uz / gam;
        const R dx = b.Sc * bz;
        // CIC of J' at x^{n+1/2}, relative to cell c: xh in (-1/4, 5/4)
        const R xh = f + R(0.5) * dx;
        const int m0 = xh < R(0) ? -1 : (xh >= R(1) ? 1 : 0);
        const R w1 = xh - (R)m0;
        const R q = b.jscale * bx;
        a0 = q * (R(1) - w1);
        a1 = q * w1;
        // new position
        R nf = f + dx;
        int d = nf < R(0) ? -1 : (nf >= R(1) ? 1 : 0);
        nf -= (R)d;
        int nc = c + d;
        if (nc < 2) { nc = 1; nf = R(0.5); }
        if (nc > N - 3) { nc = N - 2; nf = R(0.5); }
        c = nc;
        f = nf;
        code = (d + 1) | ((m0 + 1) << 2);
    }
    b.tcell[i] = c;
    b.tfrac[i] = f;
    b.tux[i] = ux;
    b.tuy[i] = uy;
    b.tuz[i] = uz;
    b.ta0[i] = a0;
    b.ta1[i] = a1;
    b.tcode[i] = (uint8_t)code;
}

// 64-bit DPP move (two 32-bit halves); lanes the control does not feed read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inclusive scan of four FP64 values per lane, segmented at fl (the first lane
// of the lane's segment): row_shr 1,2,4,8 inside 16-lane rows, then
// row_bcast:15 (rows 1 and 3 take the last lane of the row below) and
// row_bcast:31 (rows 2 and 3 take lane 31).  A step adds its source only when
// the source lane lies in the lane's own segment, so every partial stays a
// contiguous run of the segment.
__device__ __forceinline__ void seg_scan4(double (&v)[4], int lane, int fl) {
#define FDTD_SEG_STEP(CTRL, RM, COND)                                      \
    {                                                                      \
        const bool ok = (COND);                                            \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) {                    \
            const double t = dpp_f64<CTRL, RM>(v[q]);                      \
            v[q] = ok ? v[q] + t : v[q];                                   \
        }                                                                  \
    }
    FDTD_SEG_STEP(0x111, 0xf, (lane & 15) >= 1 && lane - 1 >= fl)
    FDTD_SEG_STEP(0x112, 0xf, (lane & 15) >= 2 && lane - 2 >= fl)
    FDTD_SEG_STEP(0x114, 0xf, (lane & 15) >= 4 && lane - 4 >= fl)
    FDTD_SEG_STEP(0x118, 0xf, (lane & 15) >= 8 && lane - 8 >= fl)
    FDTD_SEG_STEP(0x142, 0xa, ((lane >> 4) & 1) && (lane & ~15) - 1 >= fl)
    FDTD_SEG_STEP(0x143, 0xc, lane >= 32 && 31 >= fl)
#undef FDTD_SEG_STEP
}

// The sort/deposit workgroup is a single wave: LDS accesses then execute in
// program order, so a wavefront-scope fence (a compiler barrier, no wait on
// outstanding global loads) orders the cross-lane LDS traffic; a workgroup
// barrier would also drain the prefetched loads of the next chunk.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Segment bookkeeping of one 64-particle chunk of a cell-sorted range.
struct Chunk {
    bool valid, last;               // last: final lane of its cell in this chunk
    int fl;                         // first lane of this lane's cell in the chunk
};

// c: this lane's old cell (any value when !valid); i: particle index, hi: end of range.
__device__ __forceinline__ Chunk chunk_info(int c, int i, int hi, int lane) {
    Chunk k;
    k.valid = i < hi;
    const int cp = dpp_shr1_i(c);       // lane-1 (lane 0: unused)
    const int cn = dpp_shl1_i(c);       // lane+1 (lane 63: unused)
    const unsigned long long bnd = __ballot(k.valid && (lane == 0 || cp != c));
    const unsigned long long m = bnd & lanes_le(lane);
    k.fl = m ? 63 - __clzll((long long)m) : 0;
    k.last = k.valid && (lane == 63 || i + 1 >= hi || cn != c);
    return k;
}

constexpr int kBatch = 8;           // chunks fetched per memory round trip in the sort kernel

// Per-particle inputs of pass 1.
template <typename R>
struct P1 {
    int c, code;
    R a0, a1;
};
template <typename R>
__device__ __forceinline__ P1<R> load_p1(const BeamDev<R>& b, const int32_t* cell, int i, int hi) {
    P1<R> p;
    const bool v = i < hi;
    p.c = v ? cell[i] : 0x7fffffff;
    p.code = v ? (int)b.tcode[i] : 5;
    p.a0 = v ? b.ta0[i] : R(0);
    p.a1 = v ? b.ta1[i] : R(0);
    return p;
}

template <typename R>
__global__ __launch_bounds__(64) void pic_sort_deposit(const SimDev<R>* __restrict__ sims,
                                                       const BeamDev<R>* __restrict__ beams,
                                                       const int2* __restrict__ items, long long launch_idx) {
    constexpr int L = kPicGMax + 3;
    __shared__ int cnt[L][3];       // down, stay, up per local cell
    __shared__ int rk[L][3];        // running ranks (second pass)
    __shared__ double S[L][4];      // J' sums to nodes c-1, c, c+1, c+2
    const int2 it = items[blockIdx.x];
    const BeamDev<R>& b = beams[it.x];
    const SimDev<R>& sd = sims[b.sim];
    const int par = (int)((launch_idx - sd.launch_begin) & 1);
    const int po = par ^ 1;
    const int lane = threadIdx.x;
    const int N = b.N, np = b.np;
    const int c0 = it.y;
    const int G = min(b.G, kPicGMax);
    const int cend = min(c0 + G, N + 1);           // own cells [c0, cend)
    const int kb = c0 - 2;                          // local index k = c - kb
    for (int k = lane; k < L; k += 64) {
        cnt[k][0] = cnt[k][1] = cnt[k][2] = 0;
        rk[k][0] = rk[k][1] = rk[k][2] = 0;
        S[k][0] = S[k][1] = S[k][2] = S[k][3] = 0.0;
    }
    wave_sync();
    const int32_t* st = b.start[par];
    const int32_t* cell = b.cell[par];
    auto clampi = [&](int v) { return min(max(v, 0), np); };
    const int lo = clampi(st[max(c0 - 2, 0)]);
    const int hi = max(lo, clampi(st[min(c0 + G + 1, N + 1)]));
    const int olo = clampi(st[c0]);
    const int ohi = max(olo, clampi(st[cend]));

    // pass 1: direction counts and segmented J' sums over cells c0-2 .. c0+G.
    // Chunks are fetched kBatch at a time (one memory round trip per batch)
    // and then processed in order from registers.
    for (int bb = lo; bb < hi; bb += 64 * kBatch) {
        P1<R> pk[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) pk[q] = load_p1(b, cell, bb + 64 * q + lane, hi);
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int base = bb + 64 * q;
            if (base >= hi) break;
            const P1<R>& cur = pk[q];
            const int i = base + lane;
            const Chunk k = chunk_info(cur.c, i, hi, lane);
            const int d = (cur.code & 3) - 1;
            const int off = (cur.code >> 2) & 3;
            const double a0 = (double)cur.a0, a1 = (double)cur.a1;
            const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
            const unsigned long long bd0 = __ballot(k.valid && d == -1);
            const unsigned long long bd1 = __ballot(k.valid && d == 0);
            const unsigned long long bd2 = __ballot(k.valid && d == 1);
            double v[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) v[w] = (w == off ? a0 : 0.0) + (w == off + 1 ? a1 : 0.0);
            seg_scan4(v, lane, k.fl);
            const int lk = k.valid ? cur.c - kb : -1;
            if (k.last && lk >= 0 && lk < L) {
                cnt[lk][0] += __popcll(bd0 & seg);
                cnt[lk][1] += __popcll(bd1 & seg);
                cnt[lk][2] += __popcll(bd2 & seg);
#pragma unroll
                for (int w = 0; w < 4; ++w) S[lk][w] += v[w];
            }
            wave_sync();
        }
    }

    // pass 2: stable ranks and the scatter of this wave's own particles,
    // kBatch chunks per memory round trip
    for (int bb = olo; bb < ohi; bb += 64 * kBatch) {
        int cc[kBatch], cd[kBatch], tc[kBatch], tid[kBatch];
        R tf[kBatch], tx[kBatch], ty[kBatch], tz[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int j = bb + 64 * q + lane;
            const bool v = j < ohi;
            cc[q] = v ? cell[j] : 0x7fffffff;
            cd[q] = v ? (int)b.tcode[j] : 5;
            tc[q] = v ? b.tcell[j] : 0;
            tid[q] = v ? b.id[par][j] : 0;
            tf[q] = v ? b.tfrac[j] : R(0);
            tx[q] = v ? b.tux[j] : R(0);
            ty[q] = v ? b.tuy[j] : R(0);
            tz[q] = v ? b.tuz[j] : R(0);
        }
#pragma unroll
        for (int q = 0; q < kBatch; ++q) {
            const int base = bb + 64 * q;
            if (base >= ohi) break;
            const int c_cur = cc[q], code = cd[q];
            const int i = base + lane;
            const Chunk k = chunk_info(c_cur, i, ohi, lane);
            const int d = (code & 3) - 1;
            const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
            const unsigned long long bd0 = __ballot(k.valid && d == -1);
            const unsigned long long bd1 = __ballot(k.valid && d == 0);
            const unsigned long long bd2 = __ballot(k.valid && d == 1);
            const unsigned long long mine = d < 0 ? bd0 : (d == 0 ? bd1 : bd2);
            const int lk = k.valid ? c_cur - kb : -1;
            const bool in = k.valid && c_cur >= 0 && c_cur <= N && lk >= 1 && lk + 1 < L;
            int r = 0;
            if (in) r = rk[lk][d + 1] + __popcll(mine & lanes_lt(lane) & ~lanes_lt(k.fl));
            wave_sync();
            if (in && k.last) {
                rk[lk][0] += __popcll(bd0 & seg);
                rk[lk][1] += __popcll(bd1 & seg);
                rk[lk][2] += __popcll(bd2 & seg);
            }
            if (in) {
                int ni;
                if (d < 0) ni = st[c_cur] - cnt[lk - 1][2] + r;
                else if (d == 0) ni = st[c_cur] + cnt[lk][0] + r;
                else ni = st[c_cur + 1] - cnt[lk][2] + cnt[lk + 1][0] + r;
                if (ni >= 0 && ni < np) {
                    b.cell[po][ni] = tc[q];
                    b.frac[po][ni] = tf[q];
                    b.ux[po][ni] = tx[q];
                    b.uy[po][ni] = ty[q];
                    b.uz[po][ni] = tz[q];
                    b.id[po][ni] = tid[q];
                } else {
                    b.err[0] = 1;
                }
            }
            wave_sync();
        }
    }

    // new cell starts and J' of this wave's nodes
    for (int c = c0 + lane; c < cend; c += 64) {
        const int lk = c - kb;
        b.start[po][c] = st[c] - cnt[lk - 1][2] + cnt[lk][0];
        if (c <= N) {
            const double j = S[lk - 2][3] + S[lk - 1][2] + S[lk][1] + S[lk + 1][0];
            const_cast<R*>(sd.J)[c] = (R)j;
        }
    }
}

// With profiling events the kernels go through hipExtLaunchKernelGGL, which
// stamps the events from their own dispatch packets.
template <typename R>
hipError_t launch_pic(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* push_blocks, int npush,
                      const int2* items, int nitems, long long launch_idx, hipStream_t st, hipEvent_t e0,
                      hipEvent_t e1, hipEvent_t e2, hipEvent_t e3) {
    hipError_t e = hipSuccess;
    if (npush > 0) {
        if (e0 || e1)
            hipExtLaunchKernelGGL((pic_push<R>), dim3(npush), dim3(256), 0, st, e0, e1, 0, sims, beams,
                                  push_blocks, launch_idx);
        else
            hipLaunchKernelGGL((pic_push<R>), dim3(npush), dim3(256), 0, st, sims, beams, push_blocks, launch_idx);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nitems > 0) {
        if (e2 || e3)
            hipExtLaunchKernelGGL((pic_sort_deposit<R>), dim3(nitems), dim3(64), 0, st, e2, e3, 0, sims, beams,
                                  items, launch_idx);
        else
            hipLaunchKernelGGL((pic_sort_deposit<R>), dim3(nitems), dim3(64), 0, st, sims, beams, items,
                               launch_idx);
        e = hipGetLastError();
    }
    return e;
}

template hipError_t launch_pic<double>(const SimDev<double>*, const BeamDev<double>*, const int2*, int,
                                       const int2*, int, long long, hipStream_t, hipEvent_t, hipEvent_t,
                                       hipEvent_t, hipEvent_t);
template hipError_t launch_pic<float>(const SimDev<float>*, const BeamDev<float>*, const int2*, int,
                                      const int2*, int, long long, hipStream_t, hipEvent_t, hipEvent_t,
                                      hipEvent_t, hipEvent_t);

}  // namespace fdtd
