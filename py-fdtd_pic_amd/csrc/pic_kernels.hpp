// Particle-in-cell kernels for gfx950: relativistic Boris push with linear
// gather, deterministic stable re-sort by cell and CIC current deposit with
// FP64 segmented sums (no floating-point atomics).
//
// The reference has no particle code; the current enters the grid through
// V.Jx in ADE_ExUpdate (BaseFDTD11.py:663-669) as Ex += (dHy - Jx)*UpExMat.
// Units are the normalised ones of the field kernel: E, H' = Z0*Hy, and the
// deposit writes J' = Z0*Jx (J_x*dz in the reference's units).
//
// Step n: push/count, then the field launch of step n (which forms
// J^{n+1/2} from the CIC sums) and the scatter's node blocks, which need not
// wait for it:
//   pic_push_cell   dense beams (one-cell groups): workgroups of four waves
//                   take blocks of 8 cells, two cells per wave; per particle,
//                   gather E^n at x^n and B^n = -(H^{n-1/2} + H^{n+1/2})/(2c)
//                   at x^n - 1/2, Boris, x^{n+1} = x^n + Sc u_z/gamma; per
//                   cell, the particles moving down / up (|d| <= 1 since
//                   Sc*beta < 1), each particle's stable rank r among
//                   (cell, d), and four FP64 sums of the CIC weights of J' at
//                   x^{n+1/2} onto nodes c-1 .. c+2.  The new index is
//                      d=-1: start[c] - up(c-1) + r
//                      d= 0: start[c] + down(c) + r
//                      d=+1: start[c+1] - up(c) + down(c+1) + r
//                   and with the block's counts exchanged through LDS, every
//                   particle whose new cell lies in the block is written to
//                   the other buffer; the movers out of the block are staged.
//   pic_push_count  sparser beams (groups of G cells, one wave each): the
//                   same; a group of up to 128 particles is pushed once and
//                   held in registers across its counts, a larger one takes
//                   a counting pass and then a pass that pushes again
//                   (bitwise the same) and writes; movers out of the group
//                   are staged.
//   pic_scatter     node blocks: J'[m] = S+2(m-2) + S+1(m-1) + S0(m) +
//                   S-1(m+1), the new start[] and the next live range (a
//                   single writer each: the first and the last non-empty
//                   cell); then the staged movers to their new index, and the
//                   particles parked at the previous step copied to the
//                   other buffer.
// Absorbed particles stay in the array (cells 1 and N-2), so the particle
// count is constant and no global scan is needed.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

__device__ __forceinline__ unsigned long long lanes_le(int lane) {
    return lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
}
__device__ __forceinline__ unsigned long long lanes_lt(int lane) { return (1ull << lane) - 1ull; }

// 64-bit DPP move (two 32-bit halves); lanes the control does not feed read 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROWMASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 32-bit DPP move of a float (same lane rules as dpp_f64).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xf, true));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_val(double v) { return dpp_f64<CTRL, ROWMASK>(v); }
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_val(float v) { return dpp_f32<CTRL, ROWMASK>(v); }

// Inclusive scan of four values per lane (FP64, or FP32 in the FP32 mode),
// segmented at fl (the first lane of the lane's segment): row_shr 1,2,4,8
// inside 16-lane rows, then row_bcast:15 (rows 1 and 3 take the last lane of
// the row below) and row_bcast:31 (rows 2 and 3 take lane 31).  A step adds
// its source only when the source lane lies in the lane's own segment, so
// every partial stays a contiguous run of the segment.
template <typename T>
__device__ __forceinline__ void seg_scan4(T (&v)[4], int lane, int fl) {
#define FDTD_SEG_STEP(CTRL, RM, COND)                                      \
    {                                                                      \
        const bool ok = (COND);                                            \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) {                    \
            const T t = dpp_val<CTRL, RM>(v[q]);                           \
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

// Inclusive scan of four FP64 values over the whole wave (the segmented scan
// with every lane in one segment, without its selects: out-of-row DPP
// sources read 0 and masked rows receive 0); lane 63 holds the totals.
__device__ __forceinline__ void wave_scan4(double (&v)[4]) {
#define FDTD_SCAN_STEP(CTRL, RM)                                                      \
    {                                                                                 \
        _Pragma("unroll") for (int q = 0; q < 4; ++q) v[q] += dpp_f64<CTRL, RM>(v[q]); \
    }
    FDTD_SCAN_STEP(0x111, 0xf)
    FDTD_SCAN_STEP(0x112, 0xf)
    FDTD_SCAN_STEP(0x114, 0xf)
    FDTD_SCAN_STEP(0x118, 0xf)
    FDTD_SCAN_STEP(0x142, 0xa)
    FDTD_SCAN_STEP(0x143, 0xc)
#undef FDTD_SCAN_STEP
}

// Each push/count wave owns its own LDS tables (the workgroup's waves never
// share them), and a wave's LDS accesses execute in program order, so a
// wavefront-scope fence (a compiler barrier, no wait on outstanding global
// loads) orders the cross-lane LDS traffic; a workgroup barrier would also
// drain the prefetched loads of the next chunk.
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



// Beam and field arrays are global memory, but their pointers are read from
// descriptors, so the compiler would address them as generic (flat) memory;
// flat accesses share the LDS wait counter with the rank table, which would
// make every LDS access wait for the particle stores.  These casts keep them
// global.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
// read-only descriptors in the constant address space
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
using cref = const __attribute__((address_space(4))) T&;
template <typename T>
__device__ __forceinline__ gptr<T> gp(T* p) { return (gptr<T>)p; }
template <typename T>
__device__ __forceinline__ gptr<const T> gp(const T* p) { return (gptr<const T>)p; }

// Loads and stores at an unsigned 32-bit element index: the address is the
// (scalar) base plus a zero-extended 32-bit byte offset, which the global
// instructions take directly (no 64-bit address per lane).
template <typename T>
__device__ __forceinline__ T ldo(gptr<const T> p, unsigned i) {
    return *(gptr<const T>)((gptr<const char>)p + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void sto(gptr<T> p, unsigned i, T v) {
    *(gptr<T>)((gptr<char>)p + i * (unsigned)sizeof(T)) = v;
}

// 1/sqrt(x): the hardware reciprocal square root in FP32 (~1 ulp; the FP32
// mode's tolerance is 1e-5), IEEE operations in FP64.
__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_r(double x) { return 1.0 / sqrt(x); }
// 2t / (1 + t^2) of the Boris rotation: a reciprocal in FP32, a division in FP64
__device__ __forceinline__ float boris_s(float t) { return 2.0f * t * __builtin_amdgcn_rcpf(1.0f + t * t); }
__device__ __forceinline__ double boris_s(double t) { return 2.0 * t / (1.0 + t * t); }

// Boris push of one live particle; returns d and the CIC weights (node offset off = m0 + 1).
// What the push reads of the descriptors, fetched once per wave before the
// kernel's first store (scalar loads are only legal ahead of it).
template <typename R>
struct PushArgs {
    gptr<const R> E, Hc, Hp;     // E^n, H^{n+1/2}, H^{n-1/2}
    R a, jscale, Sc;
    int N;
};

// The multi-cell push/count runs it twice on the same particles (a counting
// pass and a writing pass), which must agree bit for bit, so no FMA
// contraction here: a contraction decision can depend on how the results
// are used at each call.
// The push after the gather: Eg, Hg are the fields at the particle.
template <typename R>
__device__ __forceinline__ void push_at(const PushArgs<R>& pa, int& c, R& f, R& ux, R& uy, R& uz, R Eg, R Hg,
                                        int& d, int& off, R& a0, R& a1) {
#pragma clang fp contract(off)
    const int N = pa.N;
    // Boris: E = (Eg, 0, 0), B = (0, -Hg/c, 0)
    const R aE = pa.a * Eg;
    const R umx = ux + aE, umy = uy, umz = uz;
    const R igm = rsqrt_r(R(1) + umx * umx + umy * umy + umz * umz);     // 1 / gamma^-
    const R ty = -pa.a * Hg * igm;
    const R sy = boris_s(ty);
    const R upx = umx - umz * ty;
    const R upz = umz + umx * ty;
    ux = umx - upz * sy + aE;
    uz = umz + upx * sy;
    uy = umy;
    const R rg = rsqrt_r(R(1) + ux * ux + uy * uy + uz * uz);
    const R bx = ux * rg, bz = uz * rg;
    const R dx = pa.Sc * bz;
    // CIC of J' at x^{n+1/2}, relative to cell c: xh in (-1/4, 5/4)
    const R xh = f + R(0.5) * dx;
    const int m0 = xh < R(0) ? -1 : (xh >= R(1) ? 1 : 0);
    const R w1 = xh - (R)m0;
    const R q = pa.jscale * bx;
    a0 = q * (R(1) - w1);
    a1 = q * w1;
    off = m0 + 1;
    // new position
    R nf = f + dx;
    d = nf < R(0) ? -1 : (nf >= R(1) ? 1 : 0);
    nf -= (R)d;
    int nc = c + d;
    if (nc < 2) { nc = 1; nf = R(0.5); }
    if (nc > N - 3) { nc = N - 2; nf = R(0.5); }
    c = nc;
    f = nf;
}

// The fields a particle of cell c can gather, wave-uniform when c is: E at
// c and c+1, and the time-averaged H, (H^{n-1/2} + H^{n+1/2})/2, at c-1, c
// and c+1.  Loaded once per one-cell group, ahead of the particles.
template <typename R>
struct CellFields {
    R E0, E1, hm, h0, hp;
};
template <typename R>
__device__ __forceinline__ CellFields<R> cell_fields(const PushArgs<R>& pa, int c) {
#pragma clang fp contract(off)
    const R e0 = ldo(pa.E, c), e1 = ldo(pa.E, c + 1);
    const R pm = ldo(pa.Hp, c - 1), qm = ldo(pa.Hc, c - 1);
    const R p0 = ldo(pa.Hp, c), q0 = ldo(pa.Hc, c);
    const R pp = ldo(pa.Hp, c + 1), qp = ldo(pa.Hc, c + 1);
    return CellFields<R>{e0, e1, R(0.5) * (pm + qm), R(0.5) * (p0 + q0), R(0.5) * (pp + qp)};
}

// The push of a particle of the cell cf was loaded for: the linear gather of
// E at x and of the time-averaged H at x - 1/2 (nodes c-1, c below the cell
// centre, c, c+1 above it), then push_at.
template <typename R>
__device__ __forceinline__ void push_cell(const PushArgs<R>& pa, const CellFields<R>& cf, int& c, R& f, R& ux,
                                          R& uy, R& uz, int& d, int& off, R& a0, R& a1) {
#pragma clang fp contract(off)
    const R Eg = (R(1) - f) * cf.E0 + f * cf.E1;
    const bool hi = f >= R(0.5);
    const R g = hi ? f - R(0.5) : f + R(0.5);
    const R h0 = hi ? cf.h0 : cf.hm;
    const R h1 = hi ? cf.hp : cf.h0;
    const R Hg = (R(1) - g) * h0 + g * h1;
    push_at<R>(pa, c, f, ux, uy, uz, Eg, Hg, d, off, a0, a1);
}

// A particle of a one-cell group through the push: its state (old, then
// new) and dr = rank << 2 | (d + 1), d = new cell - old cell, rank among its
// cell's class d.  The cell itself is not kept: in a one-cell group every
// particle of a cell's index range is in that cell, and the new cell is the
// old one plus d (the graveyard clamps give c - 1 = 1 and c + 1 = N - 2 too).
template <typename R>
struct Pushed {
    int id, dr;
    R f, ux, uy, uz;
};

// Particle state arrays of one buffer (one-cell beams keep no cell array:
// a particle's cell follows from its index and start[]).
template <typename R>
struct StateOut {
    gptr<int32_t> id;
    gptr<R> f, ux, uy, uz;
};

template <typename R>
struct StateIn {
    gptr<const int32_t> id;
    gptr<const R> f, ux, uy, uz;
};

// One particle's state from index s of one buffer to index d of another.
template <typename R>
__device__ __forceinline__ void copy_state(const StateIn<R>& a, unsigned s, const StateOut<R>& b, unsigned d) {
    sto(b.id, d, ldo(a.id, s));
    sto(b.f, d, ldo(a.f, s));
    sto(b.ux, d, ldo(a.ux, s));
    sto(b.uy, d, ldo(a.uy, s));
    sto(b.uz, d, ldo(a.uz, s));
}

// Loads of one 64-particle chunk from index base: unconditional at an
// in-range index (np >= 1); the callers select on use.
template <typename R>
__device__ __forceinline__ void g1_load_one(gptr<const int32_t> pid, gptr<const R> pf, gptr<const R> px,
                                            gptr<const R> py, gptr<const R> pz, int base, int np, int lane,
                                            Pushed<R>& s) {
    const unsigned i = (unsigned)min(base + lane, np - 1);
    s.id = ldo(pid, i);
    s.f = ldo(pf, i);
    s.ux = ldo(px, i);
    s.uy = ldo(py, i);
    s.uz = ldo(pz, i);
}

// The same for kBatch chunks from index base, 64 apart.
template <typename R>
__device__ __forceinline__ void g1_load(gptr<const int32_t> pid, gptr<const R> pf, gptr<const R> px,
                                        gptr<const R> py, gptr<const R> pz, int base, int np, int lane,
                                        Pushed<R> (&s)[kBatch]) {
#pragma unroll
    for (int q = 0; q < kBatch; ++q) g1_load_one<R>(pid, pf, px, py, pz, base + 64 * q, np, lane, s[q]);
}

// One wave's LDS row for a pushed 64-particle chunk (structure of arrays:
// lane l's fields at index l, conflict-free).
template <typename R>
struct OvfRow {
    int id[64], dr[64];
    R f[64], ux[64], uy[64], uz[64];
    __device__ __forceinline__ void put(int l, const Pushed<R>& s) {
        id[l] = s.id; dr[l] = s.dr; f[l] = s.f; ux[l] = s.ux; uy[l] = s.uy; uz[l] = s.uz;
    }
    __device__ __forceinline__ Pushed<R> get(int l) const {
        Pushed<R> s;
        s.id = id[l]; s.dr = dr[l]; s.f = f[l]; s.ux = ux[l]; s.uy = uy[l]; s.uz = uz[l];
        return s;
    }
};

// One 64-particle chunk of a one-cell group (valid: the lane holds one of the
// cell's particles): the push, its move class d, its rank among the cell's
// class d so far (scalar carries n0, n1, n2 plus the earlier lanes of the
// class) and, when SUMS, its CIC weights added to the lane's FP64 sums onto
// nodes c-1 .. c+2.
template <typename R, bool SUMS>
__device__ __forceinline__ void g1_push(const PushArgs<R>& pa, const CellFields<R>& cf, int c, bool valid,
                                        Pushed<R>& s, int& n0, int& n1, int& n2, double (&acc)[4]) {
    int d = 0, off = 1;
    R a0 = R(0), a1 = R(0);
    if (valid) push_cell<R>(pa, cf, c, s.f, s.ux, s.uy, s.uz, d, off, a0, a1);
    const unsigned long long bd0 = __ballot(valid && d == -1);
    const unsigned long long bd1 = __ballot(valid && d == 0);
    const unsigned long long bd2 = __ballot(valid && d == 1);
    // the earlier lanes of each class (the masks are wave-uniform), then the
    // lane's own class with its carry
    const int m0 = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bd0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bd0, 0u));
    const int m1 = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bd1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bd1, 0u));
    const int m2 = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bd2 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bd2, 0u));
    const int r = d < 0 ? n0 + m0 : (d == 0 ? n1 + m1 : n2 + m2);
    s.dr = (r << 2) | (d + 1);
    if (SUMS) {
        // the weights onto nodes c-1+off and c+off, each node's share chosen
        // in R and then converted (exact, so the same sums as choosing among
        // converted weights, with 32-bit selects in FP32)
        acc[0] += (double)(off == 0 ? a0 : R(0));
        acc[1] += (double)(off == 0 ? a1 : (off == 1 ? a0 : R(0)));
        acc[2] += (double)(off == 1 ? a1 : (off == 2 ? a0 : R(0)));
        acc[3] += (double)(off == 2 ? a1 : R(0));
    }
    n0 += __popcll(bd0);
    n1 += __popcll(bd1);
    n2 += __popcll(bd2);
}

// The new state of one particle at index i of o (an index outside [0, np)
// sets *err instead; never expected).
template <typename R>
__device__ __forceinline__ void g1_put(const StateOut<R>& o, int i, int np, const Pushed<R>& s,
                                       gptr<int32_t> err) {
    if ((unsigned)i >= (unsigned)np) {
        err[0] = 1;
        return;
    }
    sto(o.id, (unsigned)i, s.id);
    sto(o.f, (unsigned)i, s.f);
    sto(o.ux, (unsigned)i, s.ux);
    sto(o.uy, (unsigned)i, s.uy);
    sto(o.uz, (unsigned)i, s.uz);
}

// The new state of one pushed chunk of a cell (its particles [.., phi), the
// chunk from index base): the particle of class k = d + 1 and rank r goes to
// index x_k + r of nxt, or of stg when k == staged (-1: none).
template <typename R>
__device__ __forceinline__ void g1_put_one(const StateOut<R>& nxt, const StateOut<R>& stg, const Pushed<R>& s,
                                           int base, int phi, int lane, int x0, int x1, int x2, int staged,
                                           int np, gptr<int32_t> err) {
    if (base + lane < phi) {
        const int k = s.dr & 3, i = (k == 0 ? x0 : (k == 1 ? x1 : x2)) + (s.dr >> 2);
        if (k == staged) g1_put<R>(stg, i, np, s, err);
        else g1_put<R>(nxt, i, np, s, err);
    }
}

// The same for kBatch chunks from index base, 64 apart.
template <typename R>
__device__ __forceinline__ void g1_put_cell(const StateOut<R>& nxt, const StateOut<R>& stg,
                                            const Pushed<R> (&s)[kBatch], int base, int phi, int lane, int x0,
                                            int x1, int x2, int staged, int np, gptr<int32_t> err) {
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
        g1_put_one<R>(nxt, stg, s[q], base + 64 * q, phi, lane, x0, x1, x2, staged, np, err);
}

// The fields of a multi-cell group, staged in the wave's LDS once per group:
// E and the time-averaged H (formed as cell_fields forms it) at the nodes of
// cells c0 - 1 .. c0 + n (entry k holds node c0 - 1 + k, clamped to the grid).
template <typename R>
__device__ __forceinline__ void stage_fields(const PushArgs<R>& pa, R* Et, R* Ht, int c0, int n, int lane) {
#pragma clang fp contract(off)
    for (int k = lane; k < n + 2; k += 64) {
        const unsigned cc = (unsigned)min(max(c0 - 1 + k, 0), pa.N);
        Et[k] = ldo(pa.E, cc);
        Ht[k] = R(0.5) * (ldo(pa.Hp, cc) + ldo(pa.Hc, cc));
    }
}

// The push of a particle of cell c in [c0, c0 + n) with the staged fields:
// the gather of push_cell, then push_at.
template <typename R>
__device__ __forceinline__ void push_tab(const PushArgs<R>& pa, const R* Et, const R* Ht, int c0, int& c, R& f,
                                         R& ux, R& uy, R& uz, int& d, int& off, R& a0, R& a1) {
#pragma clang fp contract(off)
    const int k = c - c0 + 1;
    const R Eg = (R(1) - f) * Et[k] + f * Et[k + 1];
    int j;
    R g;
    if (f >= R(0.5)) { j = k; g = f - R(0.5); } else { j = k - 1; g = f + R(0.5); }
    const R Hg = (R(1) - g) * Ht[j] + g * Ht[j + 1];
    push_at<R>(pa, c, f, ux, uy, uz, Eg, Hg, d, off, a0, a1);
}

// The first wave of a beam's push/count resets the next step's live range
// (the scatter's node threads fill it), carries the union range over and
// records the previous step's graveyard bounds (start[po] still holds that
// step's start[]; this step's scatter overwrites it): a particle parked then
// was copied to the other buffer then and keeps its index.
template <typename R>
__device__ __forceinline__ void begin_step(const BeamDev<R>& b, int par, int N) {
    const int po = par ^ 1;
    gp(b.range[po])[0] = N;
    gp(b.range[po])[1] = -1;
    gp(b.urange[po])[0] = gp((const int32_t*)b.urange[par])[0];
    gp(b.urange[po])[1] = gp((const int32_t*)b.urange[par])[1];
    gp(b.park)[0] = gp((const int32_t*)b.start[po])[2];
    gp(b.park)[1] = gp((const int32_t*)b.start[po])[N - 2];
}

// Groups of a multi-cell beam over the count range [lo, hi] (at least three
// cells): G cells each, the last taking a remainder of one cell (G + 1
// cells), so every group has a first and a last cell of its own.
__device__ __forceinline__ int mc_groups(int lo, int hi, int G) {
    const int n = hi - lo + 1;
    const int g = (n + G - 1) / G;
    return (n % G == 1 && g > 1) ? g - 1 : g;
}
__device__ __forceinline__ int2 mc_group(int lo, int hi, int G, int ngroups, int j) {
    const int g0 = lo + j * G;
    return make_int2(g0, j == ngroups - 1 ? hi + 1 : g0 + G);
}

// kBatch chunks of a multi-cell group in registers (cell, state; id only
// for the writing pass).
template <typename R>
struct McBatch {
    int c[kBatch], id[kBatch];
    R f[kBatch], ux[kBatch], uy[kBatch], uz[kBatch];
};
template <typename R, bool ID>
__device__ __forceinline__ void mc_fetch(gptr<const int32_t> cell, gptr<const int32_t> pid, gptr<const R> pf,
                                         gptr<const R> px, gptr<const R> py, gptr<const R> pz, int base, int np,
                                         int lane, McBatch<R>& m) {
#pragma unroll
    for (int q = 0; q < kBatch; ++q) {
        const unsigned i = (unsigned)min(base + 64 * q + lane, np - 1);
        m.c[q] = ldo(cell, i);
        m.id[q] = ID ? ldo(pid, i) : 0;
        m.f[q] = ldo(pf, i);
        m.ux[q] = ldo(px, i);
        m.uy[q] = ldo(py, i);
        m.uz[q] = ldo(pz, i);
    }
}

// The new state of a multi-cell group's particle (cell c0 + k of the group's
// n cells, move d, rank r; stl: start[] of the group's cells, tot: their
// totals).  New index:
//   d=-1: start[c] - up(c-1) + r      d=0: start[c] + down(c) + r
//   d=+1: start[c+1] - up(c) + down(c+1) + r
// The down-movers of the first cell and the up-movers of the last need a
// neighbouring group's counts: they are staged at start[c] + r of this
// buffer (inside the cell's own range, at or below the mover's own index,
// which this wave has already read) and moved by the node blocks.
template <typename R>
__device__ __forceinline__ void mc_put(const StateOut<R>& nxt, const StateOut<R>& stg, gptr<int32_t> cello,
                                       gptr<int32_t> cellp, const int* stl, const int (*tot)[3], int n, int k,
                                       int d, int r, int c, const Pushed<R>& s, int np, gptr<int32_t> err) {
    int ni;
    bool staged = false;
    if (d == 0) {
        ni = stl[k] + tot[k][0] + r;
    } else if (d < 0) {
        staged = k == 0;
        ni = staged ? stl[k] + r : stl[k] - tot[k - 1][2] + r;
    } else {
        staged = k == n - 1;
        ni = staged ? stl[k] + r : stl[k + 1] - tot[k][2] + tot[k + 1][0] + r;
    }
    if ((unsigned)ni >= (unsigned)np) {
        err[0] = 1;
    } else if (staged) {
        g1_put<R>(stg, ni, np, s, err);
        sto(cellp, (unsigned)ni, c);
    } else {
        g1_put<R>(nxt, ni, np, s, err);
        sto(cello, (unsigned)ni, c);
    }
}

// Multi-cell groups (G > 1): kPushWaves independent waves per workgroup,
// each with its own LDS tables; item kPushWaves * block + wave, items with
// x < 0 are padding.

template <typename R>
__global__ __launch_bounds__(64 * kPushWaves) void pic_push_count(const SimDev<R>* __restrict__ sims,
                                                                  const BeamDev<R>* __restrict__ beams,
                                                                  const int2* __restrict__ items,
                                                                  long long launch_idx) {
    __shared__ int cnt_all[kPushWaves][kPicGMax + 1][3];   // down, stay, up so far per group cell (rank carry)
    __shared__ int tot_all[kPushWaves][kPicGMax + 1][3];   // the group's totals (second pass)
    __shared__ int stl_all[kPushWaves][kPicGMax + 2];      // start[] of the group's cells g0 .. g1
    __shared__ double S_all[kPushWaves][kPicGMax + 1][4];  // J' sums onto nodes c-1 .. c+2
    __shared__ R fld_all[kPushWaves][2][kPicGMax + 3];     // the group's E and averaged H (stage_fields)
    // wave-uniform by construction; readfirstlane lets the compiler see it, so
    // the item and the descriptors come in through scalar loads
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int (&cnt)[kPicGMax + 1][3] = cnt_all[wv];
    int (&tot)[kPicGMax + 1][3] = tot_all[wv];
    int* const stl = stl_all[wv];
    double (&Sl)[kPicGMax + 1][4] = S_all[wv];
    R* const Et = fld_all[wv][0];
    R* const Ht = fld_all[wv][1];
    const int2 it = items[blockIdx.x * kPushWaves + wv];
    if (it.x < 0) return;
    const cref<BeamDev<R>> b = ((cptr<BeamDev<R>>)beams)[it.x];
    const cref<SimDev<R>> sd = ((cptr<SimDev<R>>)sims)[b.sim];
    const int par = (int)((launch_idx - b.launch_begin) & 1);
    const int po = par ^ 1;
    const int lane = threadIdx.x & 63;
    const int N = b.N, np = b.np, nwaves = b.nwaves;
    const int G = max(2, min(b.G, kPicGMax));
    const PushArgs<R> pa{gp((const R*)sd.E[par]), gp((const R*)sd.H[par]), gp((const R*)sd.H[par ^ 1]),
                         b.a, b.jscale, b.Sc, N};
    const gptr<const int32_t> rng_p = gp((const int32_t*)b.range[par]);
    const gptr<const int32_t> st = gp((const int32_t*)b.start[par]);
    const gptr<const int32_t> cell = gp((const int32_t*)b.cell[par]), pid = gp((const int32_t*)b.id[par]);
    const gptr<const R> pf = gp((const R*)b.frac[par]), px = gp((const R*)b.ux[par]);
    const gptr<const R> py = gp((const R*)b.uy[par]), pz = gp((const R*)b.uz[par]);
    const int cmin = rng_p[0], cmax = rng_p[1];
    if (it.y == 0 && lane == 0) begin_step<R>(beams[it.x], par, N);
    if (cmin > cmax) return;
    const int lo_cell = max(cmin - 1, 0), hi_cell = min(cmax + 1, N);   // counts cover the range +-1
    const int ngroups = mc_groups(lo_cell, hi_cell, G);
    for (int gi = it.y; gi < ngroups; gi += nwaves) {
        const int2 gb = mc_group(lo_cell, hi_cell, G, ngroups, gi);
        const int g0 = gb.x, g1 = gb.y, n = g1 - g0;
        // live cells only (the graveyard cells 1 and N-2 are skipped); both
        // bounds come in with one load (lane 0 and lane 1), read back with readlane
        const int sb = ldo(st, (unsigned)(lane == 0 ? max(g0, 2) : min(g1, N - 2)));
        for (int k = lane; k < n; k += 64) {
            cnt[k][0] = cnt[k][1] = cnt[k][2] = 0;
            Sl[k][0] = Sl[k][1] = Sl[k][2] = Sl[k][3] = 0.0;
        }
        for (int k = lane; k <= n; k += 64) stl[k] = ldo(st, (unsigned)min(g0 + k, N + 1));
        stage_fields<R>(pa, Et, Ht, g0, n, lane);
        wave_sync();
        const int plo = min(max(__builtin_amdgcn_readlane(sb, 0), 0), np);
        const int phi = max(plo, min(max(__builtin_amdgcn_readlane(sb, 1), 0), np));
        // A group of up to 64 * kBatch live particles is done in one pass:
        // its ranks come out of the counting pass and its pushed state waits
        // in registers for the totals (a lone sparse beam's groups of ~100
        // particles are latency-bound, so a second pass would double their
        // time).  Larger groups take two passes (below).
        const bool one = phi - plo <= 64 * kBatch;
        Pushed<R> hs[kBatch];
        int hc[kBatch], hk[kBatch], hd[kBatch], hr[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; ++q) hk[q] = -1;      // -1: no particle held
        // Pass 1: the moves, counts and CIC sums.  Particle loads one batch
        // ahead: the next batch is in flight while this one is processed
        // (the wave fences below do not wait for it); unconditional loads at
        // an in-range index (np >= 1), selects on use.
        McBatch<R> nb;
        if (one) mc_fetch<R, true>(cell, pid, pf, px, py, pz, plo, np, lane, nb);
        else mc_fetch<R, false>(cell, pid, pf, px, py, pz, plo, np, lane, nb);
        for (int bb = plo; bb < phi; bb += 64 * kBatch) {
            McBatch<R> cb = nb;
            if (bb + 64 * kBatch < phi) mc_fetch<R, false>(cell, pid, pf, px, py, pz, bb + 64 * kBatch, np, lane, nb);
#pragma unroll
            for (int q = 0; q < kBatch; ++q) {
                const int base = bb + 64 * q;
                if (base >= phi) break;
                const int i = base + lane;
                const int c_old = i < phi ? cb.c[q] : 0x7fffffff;
                const Chunk k = chunk_info(c_old, i, phi, lane);
                int c = c_old, d = 0, off = 1;
                R f = cb.f[q], ux = cb.ux[q], uy = cb.uy[q], uz = cb.uz[q];
                R a0 = R(0), a1 = R(0);
                if (k.valid) push_tab<R>(pa, Et, Ht, g0, c, f, ux, uy, uz, d, off, a0, a1);
                const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
                const unsigned long long bd0 = __ballot(k.valid && d == -1);
                const unsigned long long bd1 = __ballot(k.valid && d == 0);
                const unsigned long long bd2 = __ballot(k.valid && d == 1);
                if (one && k.valid) {
                    const unsigned long long mine = d < 0 ? bd0 : (d == 0 ? bd1 : bd2);
                    hk[q] = c_old - g0;
                    hd[q] = d;
                    hr[q] = cnt[hk[q]][d + 1] + __popcll(mine & lanes_lt(lane) & ~lanes_lt(k.fl));
                    hc[q] = c;
                    hs[q].id = cb.id[q];
                    hs[q].f = f; hs[q].ux = ux; hs[q].uy = uy; hs[q].uz = uz;
                }
                // the chunk's segment sums of the weights onto nodes c-1+off,
                // c+off, in R (FP32: at most 64 weights, then added into the
                // FP64 cell sums in a fixed order)
                R v[4];
                v[0] = off == 0 ? a0 : R(0);
                v[1] = off == 0 ? a1 : (off == 1 ? a0 : R(0));
                v[2] = off == 1 ? a1 : (off == 2 ? a0 : R(0));
                v[3] = off == 2 ? a1 : R(0);
                seg_scan4<R>(v, lane, k.fl);
                const int lk = k.valid ? c_old - g0 : -1;
                if (k.valid && k.last) {
                    cnt[lk][0] += __popcll(bd0 & seg);
                    cnt[lk][1] += __popcll(bd1 & seg);
                    cnt[lk][2] += __popcll(bd2 & seg);
#pragma unroll
                    for (int w = 0; w < 4; ++w) Sl[lk][w] += (double)v[w];
                }
                wave_sync();
            }
        }
        // the group's totals: out to the node blocks, and kept for pass 2
        const gptr<int32_t> cup = gp(b.cup), cdown = gp(b.cdown);
        const gptr<double> Sg = gp(b.S);
        for (int k = lane; k < n; k += 64) {
            const int c = g0 + k;
            cup[c] = cnt[k][2];
            cdown[c] = cnt[k][0];
#pragma unroll
            for (int w = 0; w < 4; ++w) Sg[4 * (size_t)c + w] = Sl[k][w];
            tot[k][0] = cnt[k][0];
            tot[k][1] = cnt[k][1];
            tot[k][2] = cnt[k][2];
            cnt[k][0] = cnt[k][1] = cnt[k][2] = 0;
        }
        wave_sync();
        // The new state (mc_put): from registers for a one-pass group; else
        // pass 2, the same pushes again (bitwise the same moves; the group's
        // ~1,000 particles come back from L2/MALL) with the ranks.
        const StateOut<R> nxt{gp(b.id[po]), gp(b.frac[po]), gp(b.ux[po]), gp(b.uy[po]), gp(b.uz[po])};
        const StateOut<R> stg{gp(b.id[par]), gp(b.frac[par]), gp(b.ux[par]), gp(b.uy[par]), gp(b.uz[par])};
        const gptr<int32_t> cello = gp(b.cell[po]), cellp = gp(b.cell[par]);
        const gptr<int32_t> err = gp(b.err);
        if (one) {
#pragma unroll
            for (int q = 0; q < kBatch; ++q)
                if (hk[q] >= 0)
                    mc_put<R>(nxt, stg, cello, cellp, stl, tot, n, hk[q], hd[q], hr[q], hc[q], hs[q], np, err);
            continue;
        }
        if (plo < phi) mc_fetch<R, true>(cell, pid, pf, px, py, pz, plo, np, lane, nb);
        for (int bb = plo; bb < phi; bb += 64 * kBatch) {
            McBatch<R> cb = nb;
            if (bb + 64 * kBatch < phi) mc_fetch<R, true>(cell, pid, pf, px, py, pz, bb + 64 * kBatch, np, lane, nb);
#pragma unroll
            for (int q = 0; q < kBatch; ++q) {
                const int base = bb + 64 * q;
                if (base >= phi) break;
                const int i = base + lane;
                const int c_old = i < phi ? cb.c[q] : 0x7fffffff;
                const Chunk k = chunk_info(c_old, i, phi, lane);
                Pushed<R> s;
                s.id = cb.id[q];
                s.f = cb.f[q]; s.ux = cb.ux[q]; s.uy = cb.uy[q]; s.uz = cb.uz[q];
                int c = c_old, d = 0, off = 1;
                R a0 = R(0), a1 = R(0);
                if (k.valid) push_tab<R>(pa, Et, Ht, g0, c, s.f, s.ux, s.uy, s.uz, d, off, a0, a1);
                const unsigned long long seg = lanes_le(lane) & ~lanes_lt(k.fl);
                const unsigned long long bd0 = __ballot(k.valid && d == -1);
                const unsigned long long bd1 = __ballot(k.valid && d == 0);
                const unsigned long long bd2 = __ballot(k.valid && d == 1);
                const unsigned long long mine = d < 0 ? bd0 : (d == 0 ? bd1 : bd2);
                const int lk = k.valid ? c_old - g0 : 0;
                int r = 0;
                if (k.valid) r = cnt[lk][d + 1] + __popcll(mine & lanes_lt(lane) & ~lanes_lt(k.fl));
                wave_sync();
                if (k.valid && k.last) {
                    cnt[lk][0] += __popcll(bd0 & seg);
                    cnt[lk][1] += __popcll(bd1 & seg);
                    cnt[lk][2] += __popcll(bd2 & seg);
                }
                if (k.valid) mc_put<R>(nxt, stg, cello, cellp, stl, tot, n, lk, d, r, c, s, np, err);
                wave_sync();
            }
        }
    }
}

// The pair's eight FP64 CIC sums (a: cell A onto nodes c-1 .. c+2, b: cell
// B) over the wave's 64 lanes, in a fixed order, through the wave's own LDS
// rows red[0..8]: lane l adds the eight values of sum l/8 held by lanes
// 8(l%8) .. 8(l%8)+7, then lane k < 8 adds the eight partials of sum k
// (about 20 VALU operations against ~150 for two DPP wave scans).  The wave
// orders its own LDS accesses, so wavefront fences suffice.
__device__ __forceinline__ double pair_sums(double (*red)[64], const double (&a)[4], const double (&b)[4],
                                            int lane) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        red[w][lane] = a[w];
        red[4 + w][lane] = b[w];
    }
    wave_sync();
    const int k = lane >> 3, p = lane & 7;
    double t = red[k][8 * p];
#pragma unroll
    for (int j = 1; j < 8; ++j) t += red[k][8 * p + j];
    red[8][lane] = t;
    wave_sync();
    double u = 0.0;
    if (lane < 8) {
        u = red[8][8 * lane];
#pragma unroll
        for (int j = 1; j < 8; ++j) u += red[8][8 * lane + j];
    }
    wave_sync();
    return u;
}

// Waves per SIMD that pic_push_cell asks the register allocator for: FP32
// fits five (96 VGPRs, so a config-3 beam's ~5,000 waves run in one round;
// left alone the allocator takes 99 and four), FP64 needs ~160 (three).
template <typename R>
constexpr int kPushCellOcc = sizeof(R) == 4 ? 5 : 3;

static_assert(sizeof(OvfRow<float>) == kOvfRowBytes<float> && sizeof(OvfRow<double>) == kOvfRowBytes<double>,
              "OvfRow layout");
template <typename R>
constexpr size_t push_cell_lds_bytes(int xr) {
    return (size_t)kPushWaves * 2 * (size_t)xr * sizeof(OvfRow<R>);
}

// One-cell groups (G = 1): one item (beam, workgroup index) per workgroup;
// xr LDS rows per cell in the dynamic shared memory (push_cell_lds_bytes).
// single >= 0: the launch holds that beam alone and workgroup b's item is
// (single, b), so the item list is not read (one dependent load fewer).
template <typename R>
__global__ __launch_bounds__(64 * kPushWaves, kPushCellOcc<R>) void pic_push_cell(const SimDev<R>* __restrict__ sims,
                                                                 const BeamDev<R>* __restrict__ beams,
                                                                 const int2* __restrict__ items,
                                                                 long long launch_idx, int xr, int single) {
    __shared__ int xcnt[2][kPushWaves][2];      // per pair, down(A) and up(B) (double buffered)
    __shared__ double red_all[kPushWaves][9][64];   // per wave: the CIC sum reduction (pair_sums)
    extern __shared__ __align__(16) unsigned char pc_rows[];
    // wave-uniform by construction; readfirstlane lets the compiler see it, so
    // the item and the descriptors come in through scalar loads
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // this wave's rows: cell A's [0, xr), cell B's [xr, 2 xr)
    OvfRow<R>* const rows = reinterpret_cast<OvfRow<R>*>(pc_rows) + (size_t)wv * 2 * xr;
    const int2 it = single >= 0 ? make_int2(single, (int)blockIdx.x) : items[blockIdx.x];
    // the descriptors through the constant address space (the kernels never
    // write them), so their fields can be loaded where they are used, after
    // stores too, instead of all being held from the start
    const cref<BeamDev<R>> b = ((cptr<BeamDev<R>>)beams)[it.x];
    const cref<SimDev<R>> sd = ((cptr<SimDev<R>>)sims)[b.sim];
    const int par = (int)((launch_idx - b.launch_begin) & 1);
    const int po = par ^ 1;
    const int lane = threadIdx.x & 63;
    const int N = b.N, np = b.np, nwaves = b.nwaves;
    const PushArgs<R> pa{gp((const R*)sd.E[par]), gp((const R*)sd.H[par]), gp((const R*)sd.H[par ^ 1]),
                         b.a, b.jscale, b.Sc, N};
    const gptr<const int32_t> rng_p = gp((const int32_t*)b.range[par]);
    const gptr<const int32_t> st = gp((const int32_t*)b.start[par]);
    const gptr<const int32_t> pid = gp((const int32_t*)b.id[par]);
    const gptr<const R> pf = gp((const R*)b.frac[par]), px = gp((const R*)b.ux[par]);
    const gptr<const R> py = gp((const R*)b.uy[par]), pz = gp((const R*)b.uz[par]);
    const int cmin = rng_p[0], cmax = rng_p[1];
    if (it.y == 0 && wv == 0 && lane == 0) begin_step<R>(beams[it.x], par, N);
    if (cmin > cmax) return;
    const int lo_cell = max(cmin - 1, 0), hi_cell = min(cmax + 1, N);   // counts cover the range +-1
    // One-cell groups.  A workgroup takes blocks of kPushWaves pairs of
    // adjacent cells (2 * kPushWaves cells), one pair (A, B) per wave with
    // both cells' loads in flight together.  Every particle of a cell's
    // range is in that cell, so the per-class carries are wave-uniform
    // scalars and a particle's rank is its class carry plus the earlier
    // lanes of its class (mbcnt); no LDS table is needed.  Each cell keeps
    // its own carries, lane sums and reduction, in chunk order.
    //
    // The new state is written here.  A particle's new index needs its
    // own cell's counts and one neighbour's: up(c-1) for d = -1, down(c+1)
    // for d = +1.  The waves of the workgroup exchange their pairs'
    // totals through LDS, so every particle whose new cell lies in the
    // block goes straight to the other buffer; the down-movers of the
    // block's first cell and the up-movers of its last cell are staged at
    // start[c] + rank (inside the cell's own old range) and moved by the
    // scatter's node threads.  The kPushWaves waves of a workgroup share
    // its item (beam, workgroup index), so all of them run the same block
    // loop and reach every barrier.
    const int npairs = (hi_cell - lo_cell + 2) >> 1;
    const int nblk = (npairs + kPushWaves - 1) / kPushWaves;
    const int nwg = max(nwaves / kPushWaves, 1);
    int dbuf = 0;
    for (int bk = it.y; bk < nblk; bk += nwg, dbuf ^= 1) {
        const int gi = bk * kPushWaves + wv;
        const bool pv = gi < npairs;
        const int gA = lo_cell + 2 * gi, gB = gA + 1;
        const bool hasB = pv && gB <= hi_cell;
        // particles of live cells only: the graveyard cells 1 and N-2 can
        // hold most of a beam once it has drifted out, and their particles
        // neither move nor deposit (their cells' counts stay 0 below).
        // Three bounds in one load: start of A (lane 0), start of B
        // (lane 1), end of B (lane 2; an empty B when the pair has one cell).
        const int bi = lane == 0 ? gA : (lane == 1 || !hasB ? gB : gB + 1);
        const int sb = ldo(st, (unsigned)min(max(bi, 2), N - 2));
        // the cells' fields, loaded alongside the bounds (clamped cells
        // when a cell holds no live particle)
        const CellFields<R> cfA = cell_fields<R>(pa, min(max(gA, 2), N - 3));
        const CellFields<R> cfB = cell_fields<R>(pa, min(max(gB, 2), N - 3));
        const int pA = min(max(__builtin_amdgcn_readlane(sb, 0), 0), np);
        const int pB = pv ? max(pA, min(max(__builtin_amdgcn_readlane(sb, 1), 0), np)) : pA;
        const int pE = pv ? max(pB, min(max(__builtin_amdgcn_readlane(sb, 2), 0), np)) : pA;
        double accA[4] = {0.0, 0.0, 0.0, 0.0}, accB[4] = {0.0, 0.0, 0.0, 0.0};
        const int len = max(pB - pA, pE - pB);
        // cells of up to 64 * kBatch live particles (the common case): one
        // pass, the batch held in registers across the exchange; up to
        // 64 * (kBatch + xr): the same, each cell's further chunks pushed
        // after the batch and their new state parked in the wave's LDS rows
        // (a config-3 beam of ~100 particles per cell has a few such cells,
        // and a second pass there would set the launch's tail; a lone dense
        // beam gets more rows); larger ones: a counting pass, then a pass
        // that pushes again (bitwise the same) and writes.  Either way the
        // chunks are pushed, ranked and summed in the same order.
        const bool one = len <= 64 * (kBatch + xr);
        const int nx = one ? max(0, (len - 64 * kBatch + 63) / 64) : 0;   // chunks in LDS rows, <= xr
        int tA0 = 0, tA2 = 0, tB0 = 0, tB2 = 0;        // the pair's totals: A down, A up, B down, B up
        Pushed<R> qa[kBatch], qb[kBatch];
        {
            int nA0 = 0, nA1 = 0, nA2 = 0, nB0 = 0, nB1 = 0, nB2 = 0;
            const int oend = one ? min(len, 64 * kBatch) : len;
            for (int o = 0; o < oend; o += 64 * kBatch) {
                g1_load<R>(pid, pf, px, py, pz, pA + o, np, lane, qa);
                g1_load<R>(pid, pf, px, py, pz, pB + o, np, lane, qb);
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    g1_push<R, true>(pa, cfA, gA, pA + o + 64 * q + lane < pB, qa[q], nA0, nA1, nA2, accA);
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    g1_push<R, true>(pa, cfB, gB, pB + o + 64 * q + lane < pE, qb[q], nB0, nB1, nB2, accB);
            }
            for (int e = 0; e < nx; ++e) {
                const int ox = 64 * (kBatch + e);
                Pushed<R> x;
                g1_load_one<R>(pid, pf, px, py, pz, pA + ox, np, lane, x);
                g1_push<R, true>(pa, cfA, gA, pA + ox + lane < pB, x, nA0, nA1, nA2, accA);
                rows[e].put(lane, x);
                g1_load_one<R>(pid, pf, px, py, pz, pB + ox, np, lane, x);
                g1_push<R, true>(pa, cfB, gB, pB + ox + lane < pE, x, nB0, nB1, nB2, accB);
                rows[xr + e].put(lane, x);
            }
            tA0 = nA0; tA2 = nA2; tB0 = nB0; tB2 = nB2;
        }
        // the pair's totals to the block, the neighbours' back (double
        // buffered: a wave writes this buffer again two barriers later,
        // after every wave has read it)
        if (lane == 0) {
            xcnt[dbuf][wv][0] = tA0;
            xcnt[dbuf][wv][1] = tB2;
        }
        __syncthreads();
        // the outputs: the other buffer, and this buffer for the staged
        // movers (slot start[c] + r of a cell's r-th staged mover lies in the
        // cell's own range, at or below the mover's own index: only this wave
        // reads that range, and the slot's particle is already in registers)
        const StateOut<R> nxt{gp(b.id[po]), gp(b.frac[po]), gp(b.ux[po]), gp(b.uy[po]), gp(b.uz[po])};
        const StateOut<R> stg{gp(b.id[par]), gp(b.frac[par]), gp(b.ux[par]), gp(b.uy[par]), gp(b.uz[par])};
        const gptr<int32_t> err = gp(b.err);
        const bool first = wv == 0, last = wv == kPushWaves - 1;
        const int upPrev = first ? 0 : xcnt[dbuf][wv - 1][1];      // up(A-1): B of the pair below
        const int downNext = last ? 0 : xcnt[dbuf][wv + 1][0];     // down(B+1): A of the pair above
        // new indices (start[A] = pA, start[B] = pB, start[B+1] = pE for
        // a pair with live particles):
        //   A: down -> pA - up(A-1) + r, staged at pA + r in a block's
        //      first pair; stay -> pA + down(A) + r; up -> pB - up(A) + down(B) + r
        //   B: down -> pB - up(A) + r; stay -> pB + down(B) + r;
        //      up -> pE - up(B) + down(B+1) + r, staged at pB + r in a
        //      block's last pair
        // per class k = d + 1, the new index of the class's rank 0
        const int xa0 = first ? pA : pA - upPrev, xa1 = pA + tA0, xa2 = pB - tA2 + tB0;
        const int xb0 = pB - tA2, xb1 = pB + tB0, xb2 = last ? pB : pE - tB2 + downNext;
        if (one) {
            if (len > 0) {
                g1_put_cell<R>(nxt, stg, qa, pA, pB, lane, xa0, xa1, xa2, first ? 0 : -1, np, err);
                g1_put_cell<R>(nxt, stg, qb, pB, pE, lane, xb0, xb1, xb2, last ? 2 : -1, np, err);
            }
            for (int e = 0; e < nx; ++e) {
                const int ox = 64 * (kBatch + e);
                g1_put_one<R>(nxt, stg, rows[e].get(lane), pA + ox, pB, lane, xa0, xa1, xa2, first ? 0 : -1, np,
                              err);
                g1_put_one<R>(nxt, stg, rows[xr + e].get(lane), pB + ox, pE, lane, xb0, xb1, xb2, last ? 2 : -1,
                              np, err);
            }
        } else {
            int nA0 = 0, nA1 = 0, nA2 = 0, nB0 = 0, nB1 = 0, nB2 = 0;   // rank carries again
            for (int o = 0; o < len; o += 64 * kBatch) {
                g1_load<R>(pid, pf, px, py, pz, pA + o, np, lane, qa);
                g1_load<R>(pid, pf, px, py, pz, pB + o, np, lane, qb);
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    g1_push<R, false>(pa, cfA, gA, pA + o + 64 * q + lane < pB, qa[q], nA0, nA1, nA2, accA);
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    g1_push<R, false>(pa, cfB, gB, pB + o + 64 * q + lane < pE, qb[q], nB0, nB1, nB2, accB);
                g1_put_cell<R>(nxt, stg, qa, pA + o, pB, lane, xa0, xa1, xa2, first ? 0 : -1, np, err);
                g1_put_cell<R>(nxt, stg, qb, pB + o, pE, lane, xb0, xb1, xb2, last ? 2 : -1, np, err);
            }
        }
        if (!pv) continue;
        const gptr<int32_t> cup = gp(b.cup), cdown = gp(b.cdown);
        const gptr<double> Sg = gp(b.S);
        // lane k < 8 holds S[4 gA + k] (k >= 4: cell B's sums)
        const double sk = pair_sums(red_all[wv], accA, accB, lane);
        if (lane < (hasB ? 8 : 4)) Sg[4 * (size_t)gA + lane] = sk;
        if (lane == 0) {
            cup[gA] = tA2;
            cdown[gA] = tA0;
            if (hasB) {
                cup[gB] = tB2;
                cdown[gB] = tB0;
            }
        }
    }
}

// A per-cell count of push/count (cup or cdown), which exists on the live
// range [cmin, cmax] +-1 this step; zero elsewhere.
__device__ __forceinline__ int count_in(gptr<const int32_t> a, int c, int cmin, int cmax) {
    return (cmin <= cmax && c >= cmin - 1 && c <= cmax + 1) ? a[c] : 0;
}

// One particle's state from index s of one buffer to index d of the other,
// with its cell when the beam keeps a cell array (mc).
template <typename R>
__device__ __forceinline__ void move_one(const StateIn<R>& a, const StateOut<R>& b, gptr<const int32_t> cells,
                                         gptr<int32_t> celld, bool mc, unsigned s, unsigned d) {
    copy_state<R>(a, s, b, d);
    if (mc) sto(celld, d, ldo(cells, s));
}

// The second part of a beam's node blocks (block k of nmove): the movers
// push/count staged at start[c] + rank of this buffer, i.e. the down-movers
// of the first cell of each push/count block (one-cell beams: blocks of
// 2 * kPushWaves cells; multi-cell beams: the groups of mc_group) and the
// up-movers of its last cell (kXMove threads per block boundary, each taking
// every kXMove-th mover of both cells), and the particles parked at the
// previous step, [park[0], start[2]) and [start[N-2], park[1]), copied to the
// other buffer at the same index (from then on they are in both buffers and
// never copied again).  Multi-cell beams carry their cell array along.
template <typename R>
__device__ __forceinline__ void move_block(const BeamDev<R>& b, int par, int k, int t, int cmin, int cmax) {
    const int po = par ^ 1;
    const int N = b.N, np = b.np, nmove = b.nmove;
    const bool mc = b.G != 1;
    const int G = max(2, min(b.G, kPicGMax));
    const gptr<const int32_t> st = gp((const int32_t*)b.start[par]);
    const gptr<const int32_t> cup = gp((const int32_t*)b.cup), cdown = gp((const int32_t*)b.cdown);
    const gptr<const int32_t> park = gp((const int32_t*)b.park);
    const gptr<int32_t> err = gp(b.err);
    const StateOut<R> dstb{gp(b.id[po]), gp(b.frac[po]), gp(b.ux[po]), gp(b.uy[po]), gp(b.uz[po])};
    const StateIn<R> srcb{gp((const int32_t*)b.id[par]), gp((const R*)b.frac[par]), gp((const R*)b.ux[par]),
                          gp((const R*)b.uy[par]), gp((const R*)b.uz[par])};
    const gptr<const int32_t> cells = gp((const int32_t*)b.cell[par]);
    const gptr<int32_t> celld = gp(b.cell[po]);
    constexpr int kBlkCells = 2 * kPushWaves;
    if (cmin <= cmax) {
        const int lo_cell = max(cmin - 1, 0), hi_cell = min(cmax + 1, N);
        const int nblk = mc ? mc_groups(lo_cell, hi_cell, G) : (hi_cell - lo_cell + kBlkCells) / kBlkCells;
        for (int j = k * 256 + t; j < nblk * kXMove; j += nmove * 256) {
            const int g = j / kXMove, r0 = j % kXMove;
            int md, mu;
            if (mc) {
                const int2 gb = mc_group(lo_cell, hi_cell, G, nblk, g);
                md = gb.x;
                mu = gb.y - 1;
            } else {
                md = lo_cell + g * kBlkCells;
                mu = md + kBlkCells - 1;
            }
            if (md >= 2 && md <= N - 3) {
                const int n = cdown[md];
                const int src0 = st[md], dst0 = st[md] - count_in(cup, md - 1, cmin, cmax);
                if (src0 < 0 || src0 + n > np || dst0 < 0 || dst0 + n > np) err[0] = 1;
                else
                    for (int r = r0; r < n; r += kXMove)
                        move_one<R>(srcb, dstb, cells, celld, mc, (unsigned)(src0 + r), (unsigned)(dst0 + r));
            }
            if (mu >= 2 && mu <= min(hi_cell, N - 3)) {
                const int n = cup[mu];
                const int src0 = st[mu];
                const int dst0 = st[mu + 1] - count_in(cup, mu, cmin, cmax) + count_in(cdown, mu + 1, cmin, cmax);
                if (src0 < 0 || src0 + n > np || dst0 < 0 || dst0 + n > np) err[0] = 1;
                else
                    for (int r = r0; r < n; r += kXMove)
                        move_one<R>(srcb, dstb, cells, celld, mc, (unsigned)(src0 + r), (unsigned)(dst0 + r));
            }
        }
    }
    const int a0 = park[0], a1 = max(a0, st[2]);
    const int b1 = park[1], b0 = min(st[N - 2], b1);
    const int na = a1 - a0, nb = b1 - b0;
    for (int j = k * 256 + t; j < na + nb; j += nmove * 256) {
        const unsigned i = (unsigned)(j < na ? a0 + j : b0 + (j - na));
        move_one<R>(srcb, dstb, cells, celld, mc, i, i);
    }
}

// One 256-thread scatter block (node work; push/count has written every
// particle that stays within its push/count block): entry bl = (beam,
// -(k + 1)); k < nnode: J', the new start[] and the next live range over the
// union range; nnode <= k < nnode + nmove: move_block k - nnode.  t is the
// thread's index in
// the block.  Shared by pic_scatter and the fused field + scatter launch.
template <typename R>
__device__ __forceinline__ void scatter_block(const SimDev<R>* __restrict__ sims,
                                              const BeamDev<R>* __restrict__ beams, const int2 bl,
                                              long long launch_idx, const int t) {
    const BeamDev<R>& b = beams[bl.x];
    const SimDev<R>& sd = sims[b.sim];
    const int par = (int)((launch_idx - b.launch_begin) & 1);
    const int po = par ^ 1;
    // every descriptor field is read here, ahead of the kernel's first store
    // (scalar loads)
    const int N = b.N, nnode = b.nnode;
    const gptr<const int32_t> st = gp((const int32_t*)b.start[par]);
    const gptr<const int32_t> cup = gp((const int32_t*)b.cup), cdown = gp((const int32_t*)b.cdown);
    const gptr<int32_t> st_o = gp(b.start[po]);
    const gptr<const double> Sg = gp((const double*)b.S);
    const gptr<const int32_t> rng_p = gp((const int32_t*)b.range[par]), urng_p = gp((const int32_t*)b.urange[par]);
    const gptr<int32_t> rng_o = gp(b.range[po]), urng_o = gp(b.urange[po]);
    const gptr<R> J = gp(const_cast<R*>(sd.J));
    const int cmin = rng_p[0], cmax = rng_p[1];
    // counts and sums exist on [cmin-1, cmax+1] this step; zero elsewhere
    const int kk = -bl.y - 1;
    if (kk >= nnode) {
        move_block<R>(b, par, kk - nnode, t, cmin, cmax);
        return;
    }
    const int k = kk;
    const int u0 = urng_p[0], u1 = urng_p[1];
    const int e0 = max(u0 - 3, 0), e1 = min(u1 + 4, N);
    const int nlo = st[2] + count_in(cdown, 2, cmin, cmax);      // new start of cell 2 (cell 1 sends nothing up)
    const int nhi = st[N - 2] - count_in(cup, N - 3, cmin, cmax);    // new start of cell N-2
    for (int m = e0 + k * 256 + t; m <= e1; m += nnode * 256) {
        J[m] = (R)j_from_sums(Sg, cmin, cmax, m);     // the value the field step forms
        const int ns = st[m] - count_in(cup, m - 1, cmin, cmax) + count_in(cdown, m, cmin, cmax);
        st_o[m] = ns;
        if (m >= 2 && m <= N - 3) {
            const int ns1 = st[m + 1] - count_in(cup, m, cmin, cmax) + count_in(cdown, m + 1, cmin, cmax);
            if (ns1 > ns) {
                if (ns == nlo) {
                    rng_o[0] = m;
                    urng_o[0] = min(u0, m);
                }
                if (ns1 == nhi) {
                    rng_o[1] = m;
                    urng_o[1] = max(u1, m);
                }
            }
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void pic_scatter(const SimDev<R>* __restrict__ sims,
                                                   const BeamDev<R>* __restrict__ beams,
                                                   const int2* __restrict__ blocks, long long launch_idx) {
    scatter_block<R>(sims, beams, blocks[blockIdx.x], launch_idx, (int)threadIdx.x);
}

}  // namespace fdtd
