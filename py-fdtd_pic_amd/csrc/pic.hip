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

// Segment bookkeeping of one 64-particle chunk of a cell-sorted range.
struct Chunk {
    int c;                          // old cell (INT_MAX for lanes past the end)
    bool valid, last;               // last: final lane of its cell in this chunk
    int fl;                         // first lane of this lane's cell in the chunk
};

__device__ __forceinline__ Chunk chunk_info(const int32_t* __restrict__ cell, int i, int hi, int lane) {
    Chunk k;
    k.valid = i < hi;
    k.c = k.valid ? cell[i] : 0x7fffffff;
    const int cp = __shfl_up(k.c, 1);
    const int cn = __shfl_down(k.c, 1);
    const unsigned long long bnd = __ballot(k.valid && (lane == 0 || cp != k.c));
    const unsigned long long m = bnd & lanes_le(lane);
    k.fl = m ? 63 - __clzll((long long)m) : 0;
    k.last = k.valid && (lane == 63 || i + 1 >= hi || cn != k.c);
    return k;
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
    __syncthreads();
    const int32_t* st = b.start[par];
    const int32_t* cell = b.cell[par];
    auto clampi = [&](int v) { return min(max(v, 0), np); };
    const int lo = clampi(st[max(c0 - 2, 0)]);
    const int hi = max(lo, clampi(st[min(c0 + G + 1, N + 1)]));

    // pass 1: direction counts and segmented J' sums over cells c0-2 .. c0+G
    for (int base = lo; base < hi; base += 64) {
        const int i = base + lane;
        const Chunk k = chunk_info(cell, i, hi, lane);
        const int code = k.valid ? (int)b.tcode[i] : 5;
        const int d = (code & 3) - 1;
        const int off = (code >> 2) & 3;
        const double a0 = k.valid ? (double)b.ta0[i] : 0.0;
        const double a1 = k.valid ? (double)b.ta1[i] : 0.0;
        const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
        const unsigned long long bd0 = __ballot(k.valid && d == -1);
        const unsigned long long bd1 = __ballot(k.valid && d == 0);
        const unsigned long long bd2 = __ballot(k.valid && d == 1);
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = (q == off ? a0 : 0.0) + (q == off + 1 ? a1 : 0.0);
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double t = __shfl_up(v[q], o);
                if (lane - o >= k.fl) v[q] += t;
            }
        }
        const int lk = k.valid ? k.c - kb : -1;
        if (k.last && lk >= 0 && lk < L) {
            cnt[lk][0] += __popcll(bd0 & seg);
            cnt[lk][1] += __popcll(bd1 & seg);
            cnt[lk][2] += __popcll(bd2 & seg);
#pragma unroll
            for (int q = 0; q < 4; ++q) S[lk][q] += v[q];
        }
        __syncthreads();
    }

    // pass 2: stable ranks and the scatter of this wave's own particles
    const int olo = clampi(st[c0]);
    const int ohi = max(olo, clampi(st[cend]));
    for (int base = olo; base < ohi; base += 64) {
        const int i = base + lane;
        const Chunk k = chunk_info(cell, i, ohi, lane);
        const int code = k.valid ? (int)b.tcode[i] : 5;
        const int d = (code & 3) - 1;
        const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
        const unsigned long long bd0 = __ballot(k.valid && d == -1);
        const unsigned long long bd1 = __ballot(k.valid && d == 0);
        const unsigned long long bd2 = __ballot(k.valid && d == 1);
        const unsigned long long mine = d < 0 ? bd0 : (d == 0 ? bd1 : bd2);
        const int lk = k.valid ? k.c - kb : -1;
        const bool in = k.valid && k.c >= 0 && k.c <= N && lk >= 1 && lk + 1 < L;
        int r = 0;
        if (in) r = rk[lk][d + 1] + __popcll(mine & lanes_lt(lane) & ~lanes_lt(k.fl));
        __syncthreads();
        if (in && k.last) {
            rk[lk][0] += __popcll(bd0 & seg);
            rk[lk][1] += __popcll(bd1 & seg);
            rk[lk][2] += __popcll(bd2 & seg);
        }
        if (in) {
            const int c = k.c;
            int ni;
            if (d < 0) ni = st[c] - cnt[lk - 1][2] + r;
            else if (d == 0) ni = st[c] + cnt[lk][0] + r;
            else ni = st[c + 1] - cnt[lk][2] + cnt[lk + 1][0] + r;
            if (ni >= 0 && ni < np) {
                b.cell[po][ni] = b.tcell[i];
                b.frac[po][ni] = b.tfrac[i];
                b.ux[po][ni] = b.tux[i];
                b.uy[po][ni] = b.tuy[i];
                b.uz[po][ni] = b.tuz[i];
                b.id[po][ni] = b.id[par][i];
            } else {
                b.err[0] = 1;
            }
        }
        __syncthreads();
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
