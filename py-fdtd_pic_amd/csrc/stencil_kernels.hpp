#pragma once
// Temporally blocked 1-D Yee/CPML/ADE/Kerr stencil for gfx950.
//
// One wavefront (64 lanes) owns a tile of 64*K consecutive cells of one
// simulation; lane l holds cells base + l*K .. base + l*K + K-1 in registers.
// The tile advances S steps per launch without touching memory: the only
// neighbour traffic per step is one DPP wave shift of H (for the E half) and
// one of E (for the H half).  Each step shrinks the valid region by one cell
// on each side, so a tile loaded over [out0-S, out0+U+S) with U = 64K-2S
// writes back exactly its interior [out0, out0+U).  State is ping-ponged
// between two buffers so neighbouring tiles never read what another tile of
// the same launch wrote.
//
// Per cell and step, in the order of IntegratorLinLor1D / IntegratorNL1D
// (Solver_Engine.py:297-316):
//   ADE polarisation (ADE_PolarisationCurrent_Ex, BaseFDTD11.py:609-633)
//   E update + CPML psi_E (ADE_ExUpdate :663-669, CPML_Psi_e_Update :364-376)
//   soft source / TFSF (Solver_Engine.py:307-310)
//   slab: D update + E creation (ADE_DxUpdate :750-760, ADE_ExCreate :712-725,
//         AcubicFinder/NonLinExUpdate :803-877)
//   H update + CPML psi_H (ADE_HyUpdate :640-656, CPML_Psi_m_Update :381-393)
//   probes / snapshots (probeSim :16-54, vidMake :57-68)
#include <hip/hip_ext.h>

#include <type_traits>

#include "device_common.hpp"

namespace fdtd {

template <typename R, int K>
__device__ __forceinline__ R pick(const R (&a)[K], int j) {
    R v = a[0];
#pragma unroll
    for (int i = 1; i < K; ++i) v = (i == j) ? a[i] : v;
    return v;
}

// Wave-uniform lane broadcast / insert (v_readlane / v_writelane).
__device__ __forceinline__ float lane_read(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_read(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Lane of the calling thread in its wave (a tile runs as one wave of a
// multi-wave workgroup).
__device__ __forceinline__ int wave_lane() { return (int)(threadIdx.x & 63u); }

constexpr int kChunks = 4;    // S <= 255: a launch spans at most 4 x 64 steps

// Kerr small-z series y(z) = sum_k (-1)^k C(3k,k)/(2k+1) z^k; FP64 keeps 14
// terms, FP32 8 (truncation < 5e-10 relative for z < 0.01, below FP32 rounding).
template <typename R>
__device__ __forceinline__ double kerr_series(double z) {
    double y;
    if constexpr (sizeof(R) == 8) {
        y = 300830572.0;
        y = fma(y, -z, 50067108.0);
        y = fma(y, -z, 8414640.0);
        y = fma(y, -z, 1430715.0);
        y = fma(y, -z, 246675.0);
        y = fma(y, -z, 43263.0);
        y = fma(y, -z, 7752.0);
        y = fma(y, -z, 1428.0);
    } else {
        y = 1428.0;
    }
    y = fma(y, -z, 273.0);
    y = fma(y, -z, 55.0);
    y = fma(y, -z, 12.0);
    y = fma(y, -z, 3.0);
    y = fma(y, -z, 1.0);
    y = fma(y, -z, 1.0);
    return y;
}

constexpr double kKerrSeriesCut = 0.01;

// FP32 mode: the root is x (1 + delta) with x = u/eps in FP64 and the small
// correction delta = y - 1 = -z (1 - 3z + 12z^2 - 55z^3 + 273z^4) in FP32
// (z < 0.01: |delta| < 0.01; the first term left out, 1428 z^6, and the rest
// of the alternating tail stay below 1.5e-9 relative, 1/40 of an FP32 ulp;
// one FMA fewer per slab cell and step than with the z^5 term: config 2
// 12.73 against 12.80 us per launch).  The series/closed-form choice stays
// on z.
__device__ __forceinline__ float kerr_delta_f32(float z) {
    float t = 273.0f;
    t = fmaf(t, -z, 55.0f);
    t = fmaf(t, -z, 12.0f);
    t = fmaf(t, -z, 3.0f);
    t = fmaf(t, -z, 1.0f);
    return -z * t;
}
// z of the small-z test: FP32 in the FP32 mode, where z only selects the
// branch (the 1e-5 bar leaves no need for an FP64 copy).
template <typename R>
using KerrZ = typename std::conditional<sizeof(R) == 4, float, double>::type;

template <typename R>
__device__ __forceinline__ R kerr_series_e(KerrZ<R> x, double kz, KerrZ<R>& z) {
    if constexpr (sizeof(R) == 4) {
        // E = x (1 + delta) rounded once in FP32: E is stored in FP32 and
        // formed again from D - P every step, so nothing accumulates
        const float zf = (float)kz * x * x;
        z = zf;
        return fmaf(x, kerr_delta_f32(zf), x);
    } else {
        z = kz * x * x;
        return (R)(x * kerr_series<R>(z));
    }
}

// ADE + D update for one slab cell (no branches: ade == 0 is encoded as
// zero coefficients).  FP64 keeps the reference form P' = A P + B Pm + C E
// (ADE_PolarisationCurrent_Ex :620-631); FP32 keeps the velocity form
// V' = -B V - (1-A-B) P + C E, P += V' (sA = -B, sB = 1-A-B).
// HASJ: the step carries a current (general tiles); otherwise Jv is zero and
// the jD * Jv term is left out (jD * 0 is not folded under IEEE rules).
template <typename R, bool KAHAN, bool HASJ>
__device__ __forceinline__ KerrZ<R> slab_u(R kD, R jD, R dH, R Jv, R E, R& D, R& cD, R& P, R& Pm, R& cP,
                                           R sA, R sB, R sC) {
    if constexpr (sizeof(R) == 4) {
        const R Vn = sA * Pm - sB * P + sC * E;
        Pm = Vn;
        if (KAHAN) kahan_add(P, cP, Vn); else P = P + Vn;
    } else {
        const R Pn = fmr(sA, P, fmr(sB, Pm, sC * E));
        Pm = P;
        P = Pn;
    }
    const R g = HASJ ? fmr(-jD, Jv, dH) : dH;
    if (KAHAN) kahan_fma(D, cD, g, kD); else D = fmr(g, kD, D);
    // FP32: u = (D - cD) - (P - cP) from the compensated values, in FP32.
    // D - P is exact when D and P are within a factor 2 (the cancelling case)
    // and otherwise rounded relative to a u of their size, so the rounded sum
    // is within ~1e-7 of u itself.  FP64: cD = cP = 0.
    if constexpr (sizeof(R) == 4)
        return KAHAN ? (D - P) + (cP - cD) : (D - P);
    else
        return D - P;
}

// E of a slab cell from u = D - P (per KIND); ie = 1/e.
template <typename R, int KIND>
__device__ __forceinline__ R slab_e_simple(const SimDev<R>& sd, double u, R D, R ie) {
    if constexpr (KIND == SLAB_KERR_REF) {
        const double dd = (double)D / sd.kerr_e0;
        const double A = cubic_root0_1728(sd.kerr_a, sd.kerr_b, sd.kerr_c, -(fabs(dd) * fabs(dd)));
        return (R)((double)D / (sd.kerr_e0 * sd.kerr_lin + sd.kerr_e0 * sd.kerr_chi3 * A));
    } else {
        return (R)(u * (double)ie);
    }
}

// Cells whose z reached the series cut take the closed form; one wave-level
// test per step keeps the common path free of per-cell branches.
template <typename R, int K>
__device__ __forceinline__ void kerr_fixup(R (&E)[K], const KerrZ<R> (&x)[K], const KerrZ<R> (&z)[K],
                                           const double (&ied)[K], double chi3) {
    using Z = KerrZ<R>;
    constexpr Z cut = (Z)kKerrSeriesCut;
    Z zmax = z[0];
#pragma unroll
    for (int j = 1; j < K; ++j) zmax = zmax > z[j] ? zmax : z[j];
    if (__any(zmax >= cut)) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (z[j] >= cut) E[j] = (R)kerr_root_e((double)x[j] / ied[j], 1.0 / ied[j], chi3);
    }
}

// Register-resident state of one lane's K cells.  Kept as separate arrays
// passed by reference (one struct of ~25 arrays is not scalar-replaced and
// would live in scratch memory).
#define TILE_PARAMS                                                                                      \
    R(&E)[K], R(&H)[K], R(&cE)[K], R(&cH)[K], R(&ke)[K], R(&kh)[K], R(&D)[K], R(&P)[K], R(&Pm)[K],  \
        R(&cD)[K], R(&cP)[K], R(&sA)[K], R(&sB)[K], R(&sC)[K], R(&ie)[K], double(&ied)[K],            \
        double(&kz)[K], R(&psE)[K], R(&psH)[K], R(&bE)[K], R(&cEc)[K], R(&CbE)[K], R(&bH)[K],           \
        R(&cHc)[K], R(&C2H)[K], const unsigned slab_mask, const unsigned interior_mask
#define TILE_ARGS                                                                                       \
    E, H, cE, cH, ke, kh, D, P, Pm, cD, cP, sA, sB, sC, ie, ied, kz, psE, psH, bE, cEc, CbE, bH, cHc, C2H, \
        slab_mask, interior_mask

// Uniform per-launch values of one tile.
template <typename R>
struct TileCtx {
    int lane, t0, base, N, s_lo, s_hi, t_first, T;
    long long gstep0, t_begin;
    int src_lane, src_j, tf_lane, tf_j;
    int npr, pr_rel[2], pr_t0[2], pr_t1[2], pr_id[2];
    unsigned probe_bits;
    int nprobe, snap_interval, snap_rows;
    R kD, jD;
    double chi3;
    const R *srcE, *srcH, *Jc;
    double* Hprev;          // energy: H' before the launch's last H update, or null
    const double* pS;       // beam CIC sums (J' formed here), or null: J' read from Jc
    int pcmin, pcmax;
    bool jpre;              // J' of the lane's cells formed at set-up (K <= 2): Jpre
    double Jpre[2];
    double* probe_out;      // FP64 samples (FP32 fields: hi - c of the Kahan pair)
    R* snap;
    R (*prl)[2][64];        // the wave's LDS rows of probe samples: [probe][E, cE][step in the chunk]
    // workgroup halo exchange (M_GROUP): this wave's place in its workgroup,
    // the exchange period in steps (= halo cells on an inner side) and the
    // emulated vacuum cells (E-form Kerr slab-boundary tiles)
    int wv, hx;
    unsigned emu_mask;
};

// What a tile's step loop has to handle (compile-time, one loop per class).
enum : int {
    M_PURE_SLAB = 1,   // every cell is a slab cell
    M_MIXED = 2,       // slab and interior cells: selects per cell
    M_STRIP = 4,       // CPML psi (coefficients are zero off the strips)
    M_SRC = 8,         // soft source / TFSF cell in the tile
    M_PROBE = 16,      // up to two probes in the tile's interior
    M_GENERAL = 32,    // current J, snapshots, more than two probes
    M_GROUP = 64,      // the workgroup's four windows exchange inner halos every hx steps
};

// Inner halos of a workgroup's four adjacent windows (M_GROUP).  Window w
// starts 64K - 2hx cells after window w - 1, so its first hx cells are the
// owned cells hx .. 2hx - 1 from the end of window w - 1, and its last hx
// cells are the owned cells hx .. 2hx - 1 of window w + 1.  Between
// exchanges each window recomputes its halos itself; the invalid front that
// enters at a window edge moves one cell per step, so after hx steps it has
// just reached the owned cells and the halos are replaced by the owners'
// values.  Every state array of a cell goes across (E, H and their Kahan
// compensations, D, P, Pm, cD, cP, psi_E, psi_H); an emulated vacuum cell
// (D-form Kerr tiles) sends the compensation of its E, E - (D - cD), as the
// write-back stores it, and the receiver emulates its halo cells again.
// Double buffered: a buffer is written again only after the next barrier,
// which every wave reaches after reading it.  All four waves run the same
// steps (one simulation per workgroup), so they pass the same barriers.
constexpr int kTileWaves = 4;        // tiles (waves) per workgroup
constexpr int kHaloMaxLanes = 8;     // hx / K lanes per inner side
template <typename R, int K, bool KAHAN>
__device__ __forceinline__ void halo_exchange(TILE_PARAMS, const TileCtx<R>& c, int buf) {
    constexpr int NV = 11;
    __shared__ R xb[2][4][2][NV][kHaloMaxLanes][K];
    constexpr int st = kHaloMaxLanes * K;    // between the arrays of one (buffer, wave, side)
    const int hl = c.hx / K;
    const int lane = c.lane;
    const int w = c.wv;
    // side 0: cells hx .. 2hx-1 (the left neighbour's right halo); side 1:
    // cells 64K-2hx .. 64K-hx-1 (the right neighbour's left halo)
    int side = -1, q = 0;
    if (w > 0 && lane >= hl && lane < 2 * hl) { side = 0; q = lane - hl; }
    if (w < 3 && lane >= 64 - 2 * hl && lane < 64 - hl) { side = 1; q = lane - (64 - 2 * hl); }
    if (side >= 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            R ce = cE[j];
            if (KAHAN && ((c.emu_mask >> j) & 1u)) ce = (R)((double)E[j] - ((double)D[j] - (double)cD[j]));
            R* o = &xb[buf][w][side][0][q][j];
            o[0 * st] = E[j]; o[1 * st] = H[j]; o[2 * st] = ce; o[3 * st] = cH[j];
            o[4 * st] = D[j]; o[5 * st] = P[j]; o[6 * st] = Pm[j]; o[7 * st] = cD[j];
            o[8 * st] = cP[j]; o[9 * st] = psE[j]; o[10 * st] = psH[j];
        }
    }
    __syncthreads();
    int from = -1, fside = 0, fq = 0;
    if (w > 0 && lane < hl) { from = w - 1; fside = 1; fq = lane; }
    if (w < 3 && lane >= 64 - hl) { from = w + 1; fside = 0; fq = lane - (64 - hl); }
    if (from >= 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const R* o = &xb[buf][from][fside][0][fq][j];
            E[j] = o[0 * st]; H[j] = o[1 * st]; cE[j] = o[2 * st]; cH[j] = o[3 * st];
            D[j] = o[4 * st]; P[j] = o[5 * st]; Pm[j] = o[6 * st]; cD[j] = o[7 * st];
            cP[j] = o[8 * st]; psE[j] = o[9 * st]; psH[j] = o[10 * st];
            if ((c.emu_mask >> j) & 1u) {
                D[j] = E[j];
                cD[j] = cE[j];
                P[j] = Pm[j] = cP[j] = R(0);
            }
        }
    }
}

// Slab E half for all K cells of a pure-slab lane (straight-line: the K FP64
// chains interleave).
template <typename R, int K, bool KAHAN, int KIND, bool HASJ>
__device__ __forceinline__ void e_pure_slab(TILE_PARAMS, const TileCtx<R>& c, const SimDev<R>& sd,
                                            const R (&dH)[K], const R (&Jv)[K]) {
    KerrZ<R> x[K], z[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const KerrZ<R> u = slab_u<R, KAHAN, HASJ>(c.kD, c.jD, dH[j], Jv[j], E[j], D[j], cD[j], P[j], Pm[j],
                                            cP[j], sA[j], sB[j], sC[j]);
        if constexpr (KIND == SLAB_KERR_E) {
            x[j] = u * ie[j];          // ie = 1/eps_inf in R (FP64: the same value as ied)
            E[j] = kerr_series_e<R>(x[j], kz[j], z[j]);
        } else {
            E[j] = slab_e_simple<R, KIND>(sd, (double)u, D[j], ie[j]);
        }
    }
    if constexpr (KIND == SLAB_KERR_E) kerr_fixup<R, K>(E, x, z, ied, c.chi3);
}

// Mixed tile: every cell takes the interior update (+ CPML psi), slab cells
// then take E from D - P through selects.
template <typename R, int K, bool KAHAN, int KIND, bool STRIP, bool HASJ>
__device__ __forceinline__ void e_mixed(TILE_PARAMS, const TileCtx<R>& c, const SimDev<R>& sd,
                                        const R (&dH)[K], const R (&Jv)[K]) {
    R Ev[K], cEv[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        Ev[j] = E[j]; cEv[j] = cE[j];
        const R a = dH[j] - Jv[j];
        if (KAHAN) kahan_fma(Ev[j], cEv[j], a, ke[j]); else Ev[j] = fmr(a, ke[j], Ev[j]);
        if (STRIP) {
            psE[j] = fmr(bE[j], psE[j], cEc[j] * dH[j]);
            Ev[j] = fmr(-CbE[j], psE[j], Ev[j]);
        }
    }
    KerrZ<R> x[K], z[K];
    R Es[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const bool sl = (slab_mask >> j) & 1u;
        R Dn = D[j], cDn = cD[j], Pn = P[j], Pmn = Pm[j], cPn = cP[j];
        const KerrZ<R> u = slab_u<R, KAHAN, HASJ>(c.kD, c.jD, dH[j], Jv[j], E[j], Dn, cDn, Pn, Pmn, cPn, sA[j],
                                            sB[j], sC[j]);
        if constexpr (KIND == SLAB_KERR_E) {
            x[j] = u * ie[j];
            Es[j] = kerr_series_e<R>(x[j], sl ? kz[j] : 0.0, z[j]);
        } else {
            x[j] = 0;
            z[j] = 0;
            Es[j] = slab_e_simple<R, KIND>(sd, (double)u, Dn, ie[j]);
        }
        D[j] = sl ? Dn : D[j]; cD[j] = sl ? cDn : cD[j];
        P[j] = sl ? Pn : P[j]; Pm[j] = sl ? Pmn : Pm[j]; cP[j] = sl ? cPn : cP[j];
    }
    if constexpr (KIND == SLAB_KERR_E) kerr_fixup<R, K>(Es, x, z, ied, c.chi3);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const bool sl = (slab_mask >> j) & 1u;
        E[j] = sl ? Es[j] : Ev[j];
        cE[j] = sl ? cE[j] : cEv[j];
    }
}

template <typename R, int K, bool KAHAN, int KIND, int MODE>
__device__ __forceinline__ void run_steps(TILE_PARAMS, const TileCtx<R>& c, const SimDev<R>& sd) {
    constexpr bool SLAB_CODE = (MODE & (M_PURE_SLAB | M_MIXED)) != 0 && KIND >= SLAB_LINEAR;
    constexpr bool STRIP = (MODE & M_STRIP) != 0;
    constexpr bool SRC = (MODE & M_SRC) != 0;
    constexpr bool PROBE = (MODE & M_PROBE) != 0;
    constexpr bool GEN = (MODE & M_GENERAL) != 0;
    constexpr bool GROUP = (MODE & M_GROUP) != 0;
    int xleft = c.hx, xbuf = 0;      // steps to the next halo exchange, its buffer
    // source values of this launch, one step per lane (read back with
    // v_readlane): no global memory access on the per-step path
    // (chunk q of 64 steps sits in srcEc[0] while it runs; the arrays rotate
    // at chunk ends so every index is a compile-time constant)
    // The source and TFSF cells enter through 0/1 masks per lane and cell:
    // E + 1*v is E + v exactly and E + 0*v is E, so the injection needs no
    // branch on the lane or the cell.
    // Probe samples of the current 64 steps: the lane holding the probe cell
    // writes the field and (FP32) its Kahan compensation into the wave's LDS
    // rows c.prl, step j of the chunk at index j (one masked LDS store per
    // value, against a v_readlane, its wait states and a select per value
    // when the samples were gathered lane by lane in registers: config 2
    // launches 12.78 against 13.27 us); at the chunk's end lane j reads step
    // j back, combines the pair in FP64 and stores it
    R srcEc[kChunks], srcHc[kChunks], ms[K], mt[K];
    if constexpr (SRC) {
#pragma unroll
        for (int j = 0; j < K; ++j) {
            ms[j] = (c.src_lane >= 0 && c.lane == c.src_lane && j == c.src_j) ? R(1) : R(0);
            mt[j] = (c.tf_lane >= 0 && c.lane == c.tf_lane && j == c.tf_j) ? R(1) : R(0);
        }
#pragma unroll
        for (int q = 0; q < kChunks; ++q) {
            const int ix = q * 64 + c.lane;
            const bool in = ix < (c.s_hi - c.s_lo);
            srcEc[q] = (in && c.src_lane >= 0) ? c.srcE[c.t_first + ix] : R(0);
            srcHc[q] = (in && c.tf_lane >= 0) ? c.srcH[c.t_first + ix] : R(0);
        }
    }
    // the step body, called twice per loop trip: with two copies the
    // loop-carried registers alternate instead of being moved back each step
    auto step = [&](const int s) {
        const int ix = s - c.s_lo;
        R Jv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) Jv[j] = R(0);
        if constexpr (GEN) {
            if (c.pS && c.jpre) {
#pragma unroll
                for (int j = 0; j < K; ++j) Jv[j] = (R)c.Jpre[j < 2 ? j : 0];
            } else if (c.pS) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const int n = c.base + j;
                    if (n >= 0 && n <= c.N) Jv[j] = (R)j_from_sums(c.pS, c.pcmin, c.pcmax, n);
                }
            } else if (c.Jc) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const int n = c.base + j;
                    if (n >= 0 && n <= c.N) Jv[j] = c.Jc[n];
                }
            }
        }
        // ---------------- E half ----------------
        const R Hl = shr1(H[K - 1]);
        R dH[K];
#pragma unroll
        for (int j = 0; j < K; ++j) dH[j] = H[j] - ((j == 0) ? Hl : H[j - 1]);
        if constexpr (SLAB_CODE && (MODE & M_PURE_SLAB) && !STRIP) {
            e_pure_slab<R, K, KAHAN, KIND, GEN>(TILE_ARGS, c, sd, dH, Jv);
        } else if constexpr (SLAB_CODE) {
            e_mixed<R, K, KAHAN, KIND, STRIP, GEN>(TILE_ARGS, c, sd, dH, Jv);
        } else {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const R a = dH[j] - Jv[j];
                if (KAHAN) kahan_fma(E[j], cE[j], a, ke[j]); else E[j] = fmr(a, ke[j], E[j]);
                if (STRIP) {
                    psE[j] = fmr(bE[j], psE[j], cEc[j] * dH[j]);
                    E[j] = fmr(-CbE[j], psE[j], E[j]);
                }
            }
        }
        if constexpr (SRC) {
            const R sv = lane_read(srcEc[0], ix & 63);
#pragma unroll
            for (int j = 0; j < K; ++j) E[j] = fmr(ms[j], sv, E[j]);
            // ---------------- H half (TFSF correction first) ----------------
            const R hv = lane_read(srcHc[0], ix & 63);
#pragma unroll
            for (int j = 0; j < K; ++j) H[j] = fmr(-mt[j], hv, H[j]);
        }
        if constexpr (GEN) {
            // energy: the interior's H' before the launch's last H update
            if (c.Hprev && s == c.s_hi - 1) {
#pragma unroll
                for (int j = 0; j < K; ++j)
                    if ((interior_mask >> j) & 1u)
                        c.Hprev[c.base + j] = (double)H[j] - (KAHAN ? (double)cH[j] : 0.0);
            }
        }
        const R Er = shl1(E[0]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const R dE = ((j == K - 1) ? Er : E[j + 1]) - E[j];
            if (KAHAN) kahan_fma(H[j], cH[j], dE, kh[j]); else H[j] = fmr(dE, kh[j], H[j]);
            if (STRIP) {
                psH[j] = fmr(bH[j], psH[j], cHc[j] * dE);
                H[j] = fmr(C2H[j], psH[j], H[j]);
            }
        }
        // ---------------- probes / snapshots ----------------
        if constexpr (PROBE && !GEN) {
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (k < c.npr && c.lane == c.pr_rel[k] / K) {
                    c.prl[k][0][ix & 63] = pick<R, K>(E, c.pr_rel[k] % K);
                    if (KAHAN) c.prl[k][1][ix & 63] = pick<R, K>(cE, c.pr_rel[k] % K);
                }
        }
        // end of a 64-step chunk: flush probe samples, rotate source chunks
        if constexpr (PROBE || SRC) {
            if (((ix & 63) == 63) || s == c.s_hi - 1) {
                if constexpr (PROBE && !GEN) {
                    // the owning lanes' LDS stores come before the other lanes' reads
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    const int tt = c.t_first + (ix & ~63) + c.lane;
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        if (k < c.npr && c.lane <= (ix & 63) && tt >= c.pr_t0[k] && tt <= c.pr_t1[k])
                            // FP64 in both precisions: an FP32 field as hi - c of
                            // its Kahan pair (the value fdtd_read_field returns)
                            c.probe_out[(long long)c.pr_id[k] * c.T + tt] =
                                KAHAN ? (double)c.prl[k][0][c.lane] - (double)c.prl[k][1][c.lane]
                                      : (double)c.prl[k][0][c.lane];
                    // and the reads before the next chunk's stores
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
                if constexpr (SRC) {
#pragma unroll
                    for (int q = 0; q + 1 < kChunks; ++q) { srcEc[q] = srcEc[q + 1]; srcHc[q] = srcHc[q + 1]; }
                }
            }
        }
        if constexpr (GEN) {
            const int t = c.t_first + ix;
            if (c.probe_bits) {
                for (int p = 0; p < c.nprobe; ++p) {
                    if (!((c.probe_bits >> p) & 1u)) continue;
                    if (t < sd.probe_t0[p] || t > sd.probe_t1[p]) continue;
                    const int rel = sd.probe_cell[p] - c.t0;
                    if (c.lane == rel / K)
                        c.probe_out[(long long)p * c.T + t] =
                            KAHAN ? (double)pick<R, K>(E, rel % K) - (double)pick<R, K>(cE, rel % K)
                                  : (double)pick<R, K>(E, rel % K);
                }
            }
            if (c.snap_interval > 0 && t > 0 && (t % c.snap_interval) == 0) {
                const int row = t / c.snap_interval;
                if (row < c.snap_rows) {
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        if ((interior_mask >> j) & 1u)
                            c.snap[(long long)row * (c.N + 1) + c.base + j] = E[j];
                }
            }
        }
        if constexpr (GROUP) {
            if (--xleft == 0) {
                xleft = c.hx;
                if (s + 1 < c.s_hi) {
                    halo_exchange<R, K, KAHAN>(TILE_ARGS, c, xbuf);
                    xbuf ^= 1;
                }
            }
        }
    };
    int s = c.s_lo;
    for (; s + 1 < c.s_hi; s += 2) {
        step(s);
        step(s + 1);
    }
    if (s < c.s_hi) step(s);
}

// Wave-uniform copy of a 64-bit value (both halves through readfirstlane).
__device__ __forceinline__ unsigned long long wave_uniform(unsigned long long v) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffull));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Device-coherent loads and stores of the persistent halo exchange: relaxed
// atomics at agent scope are issued with the device-scope cache bits, so each
// access is coherent across the XCDs' L2 caches by itself, and no fence has to
// write back or invalidate a whole L2 (an agent-scope fence per block and
// tile made a 42-step block take 55 us against 12.8 us for a one-block launch).
template <typename T>
__device__ __forceinline__ T ld_dev(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_dev(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits (wave-uniform) until *flag >= need; false when the real-time counter
// passes deadline first.  The halo loads that follow are device-scope loads
// issued after the flag was seen; the workgroup-scope acquire orders them
// after it (it waits for the flag load and drops this CU's L1 lines).
__device__ __forceinline__ bool wait_flag(unsigned long long* flag, unsigned long long need, long long deadline) {
    for (;;) {
        const unsigned long long v = wave_uniform(ld_dev(flag));
        if (v >= need) break;
        if ((long long)__builtin_amdgcn_s_memrealtime() > deadline) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// One tile: 64 lanes x K cells, S steps (the whole per-tile program); with
// PERSIST, pa.nblocks blocks of S steps with the neighbour exchange between
// them (a separate instance: the one-block kernel keeps its own register
// allocation).
template <typename R, int K, bool KAHAN, int KIND, bool PERSIST, bool GROUP = false>
__device__ __forceinline__ void tile_body(const SimDev<R>& sd, int out0_, long long launch_idx, long long gstep0,
                                          int nsteps, int S, int tid, const PersistArgs& pa) {
    constexpr bool DFORM = KIND >= SLAB_LINEAR;
    const int lane = wave_lane();
    // the window [t0, t0 + 64K) and the cells it owns (writes back, probes):
    // a tile entry holds the first owned cell, the window reaching S cells
    // to either side; with GROUP it holds the window start, and wave w of the
    // workgroup owns all but its outer halo of S cells (first and last wave)
    // and its inner halos of hx cells
    const int wv = GROUP ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
    const int hx = GROUP ? pa.halo : 0;
    const int t0 = GROUP ? out0_ : out0_ - S;
    const int out0 = GROUP ? t0 + (wv == 0 ? S : hx) : out0_;
    const int U = GROUP ? 64 * K - (wv == 0 ? S : hx) - (wv == kTileWaves - 1 ? S : hx) : 64 * K - 2 * S;
    const int N = sd.N;
    const int par = (int)((launch_idx - sd.launch_begin) & 1);
    const int F = sd.F, Rr = sd.Rr, Ls = sd.Ls, Rs = sd.Rs;

    // Everything the step loop reads from the descriptor is copied into
    // registers: the loop's global stores (probes, snapshots) could alias the
    // descriptor, so the compiler would otherwise reload it every step.
    TileCtx<R> c;
    // probe sample rows, one set per wave of the 256-thread workgroup
    __shared__ R pr_rows[kTileWaves][2][2][64];
    c.prl = pr_rows[__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))];
    c.lane = lane;
    c.t0 = t0;
    c.wv = wv;
    c.hx = hx;
    c.emu_mask = 0;
    c.base = c.t0 + lane * K;
    c.N = N;
    c.T = sd.T;
    c.gstep0 = gstep0;
    c.t_begin = sd.t_begin;
    c.kD = sd.kD; c.jD = sd.jD; c.chi3 = sd.kerr_chi3;
    c.nprobe = sd.nprobe; c.snap_interval = sd.snap_interval; c.snap_rows = sd.snap_rows;
    c.srcE = sd.srcE; c.srcH = sd.srcH; c.Jc = sd.J; c.probe_out = sd.probe_out; c.snap = sd.snap;
    c.Hprev = sd.Hprev;
    c.pS = sd.pS;
    c.pcmin = 1;
    c.pcmax = 0;
    c.jpre = false;
    c.Jpre[0] = c.Jpre[1] = 0.0;
    if (c.pS) {
        // J' of a launch is the same at each of its steps (the sums of this
        // launch's push).  With K <= 2 it is formed here: the sums come in
        // from clamped cells without waiting for the live range, which only
        // masks them afterwards (the sum of j_from_sums, in its order).
        double raw[K <= 2 ? K : 1][4];
        if constexpr (K <= 2 && !PERSIST) {
#pragma unroll
            for (int j = 0; j < K; ++j)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int cc = min(max(c.base + j - 2 + w, 0), N + 1);
                    raw[j][w] = c.pS[4 * (size_t)cc + (3 - w)];
                }
        }
        c.pcmin = sd.prange[par][0];
        c.pcmax = sd.prange[par][1];
        if constexpr (K <= 2 && !PERSIST) {
            c.jpre = true;
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int n = c.base + j;
                double t = 0.0;
                if (n >= 0 && n <= N) {
#pragma unroll
                    for (int w = 0; w < 4; ++w) {
                        const int cc = n - 2 + w;
                        const double v = (c.pcmin <= c.pcmax && cc >= c.pcmin - 1 && cc <= c.pcmax + 1) ? raw[j][w]
                                                                                                       : 0.0;
                        t = w == 0 ? v : t + v;
                    }
                }
                c.Jpre[j] = t;
            }
        }
    }

    R E[K], H[K], cE[K], cH[K], ke[K], kh[K];
    R D[K], P[K], Pm[K], cD[K], cP[K], sA[K], sB[K], sC[K], ie[K];
    double ied[K], kz[K];
    R psE[K], psH[K], bE[K], cEc[K], CbE[K], bH[K], cHc[K], C2H[K];
    unsigned slab_mask = 0, strip_mask = 0, interior_mask = 0;
    const R* Ein = sd.E[par];
    const R* Hin = sd.H[par];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int n = c.base + j;
        const bool valid = (n >= 0) && (n <= N);
        E[j] = valid ? Ein[n] : R(0);
        H[j] = valid ? Hin[n] : R(0);
        cE[j] = cH[j] = R(0);
        if (KAHAN && valid) { cE[j] = sd.cE[par][n]; cH[j] = sd.cH[par][n]; }
        ke[j] = valid ? sd.ke0 : R(0);
        kh[j] = valid ? sd.kh0 : R(0);
        D[j] = P[j] = Pm[j] = cD[j] = cP[j] = sA[j] = sB[j] = sC[j] = R(0);
        ie[j] = R(1);
        psE[j] = psH[j] = bE[j] = cEc[j] = CbE[j] = bH[j] = cHc[j] = C2H[j] = R(0);
        const bool inslab = valid && n >= F && n < Rr;
        if (KIND == SLAB_DIEL && inslab) { ke[j] = sd.keS; kh[j] = sd.khS; }
        if (DFORM && inslab) {
            const int i = n - F;
            slab_mask |= 1u << j;
            D[j] = sd.D[par][i]; P[j] = sd.P[par][i]; Pm[j] = sd.Pm[par][i];
            if (KAHAN) { cD[j] = sd.cD[par][i]; cP[j] = sd.cP[par][i]; }
            sA[j] = sd.sl_A[i]; sB[j] = sd.sl_B[i]; sC[j] = sd.sl_C[i]; ie[j] = sd.sl_e[i];
        }
        if (valid && (n < Ls || n >= Rs)) {
            const int i = n < Ls ? n : Ls + (n - Rs);
            strip_mask |= 1u << j;
            ke[j] = sd.s_ke[i]; kh[j] = sd.s_kh[i];
            bE[j] = sd.s_bE[i]; cEc[j] = sd.s_cE[i]; CbE[j] = sd.s_CbE[i];
            bH[j] = sd.s_bH[i]; cHc[j] = sd.s_cH[i]; C2H[j] = sd.s_C2H[i];
            psE[j] = sd.psiE[par][i]; psH[j] = sd.psiH[par][i];
        }
        if (valid && n >= out0 && n < out0 + U) interior_mask |= 1u << j;
        ied[j] = (double)ie[j];
        kz[j] = c.chi3 * ied[j];
    }
    constexpr unsigned full = (K >= 32) ? 0xffffffffu : ((1u << K) - 1u);
    // tile classes (wave-uniform)
    const bool tile_strip = __any(strip_mask != 0);
    const bool pure_slab = DFORM && __all(slab_mask == full);
    const bool any_slab = DFORM && __any(slab_mask != 0);
    const int tile_end = c.t0 + 64 * K;
    const int src = sd.src_cell;
    const bool tile_src = src >= c.t0 && src < tile_end;
    c.src_lane = tile_src ? (src - c.t0) / K : -1;
    c.src_j = tile_src ? (src - c.t0) % K : 0;
    const bool tile_tf = sd.tfsf && (src - 1) >= c.t0 && (src - 1) < tile_end;
    c.tf_lane = tile_tf ? (src - 1 - c.t0) / K : -1;
    c.tf_j = tile_tf ? (src - 1 - c.t0) % K : 0;
    c.probe_bits = 0;
    for (int p = 0; p < c.nprobe; ++p) {
        const int cc = sd.probe_cell[p];
        if (cc >= out0 && cc < out0 + U && cc <= N) c.probe_bits |= 1u << p;
    }
    c.npr = __popc(c.probe_bits);
    // first two probes of the tile (constant indices keep TileCtx in registers)
    c.pr_rel[0] = c.pr_rel[1] = c.pr_id[0] = c.pr_id[1] = c.pr_t0[0] = c.pr_t0[1] = 0;
    c.pr_t1[0] = c.pr_t1[1] = -1;
    {
        bool have0 = false;
        for (int p = 0; p < c.nprobe; ++p) {
            if (!((c.probe_bits >> p) & 1u)) continue;
            const int rel = sd.probe_cell[p] - c.t0, a = sd.probe_t0[p], z = sd.probe_t1[p];
            if (!have0) { c.pr_rel[0] = rel; c.pr_t0[0] = a; c.pr_t1[0] = z; c.pr_id[0] = p; have0 = true; }
            else { c.pr_rel[1] = rel; c.pr_t0[1] = a; c.pr_t1[1] = z; c.pr_id[1] = p; break; }
        }
    }

    // one lean loop per tile class.  With a beam, J' is formed from the CIC
    // sums of this step's live range [cmin, cmax] and is zero off the nodes
    // cmin-2 .. cmax+3; a tile whose window misses them runs without J.
    bool needJ = c.Jc != nullptr;
    if (c.pS) needJ = c.pcmin <= c.pcmax && c.t0 + 64 * K - 1 >= c.pcmin - 2 && c.t0 <= c.pcmax + 3;
    const bool gen = needJ || c.snap_interval > 0 || c.npr > 2 || c.Hprev != nullptr;
    const bool extras = tile_strip || tile_src || tile_tf || c.npr > 0;
    // A slab-boundary tile of the E-form Kerr slab runs the pure-slab loop:
    // a vacuum cell is a slab cell with D = E, P = 0, eps = 1, chi3 = 0 when
    // the D and E updates share the coefficient (kD == ke0, jD == 1); FP64
    // then gives the vacuum update bit for bit.
    unsigned emu_mask = 0;
    bool emu_tile = false;
    auto emulate = [&](int j) {
        D[j] = E[j];
        cD[j] = cE[j];
        P[j] = Pm[j] = cP[j] = R(0);
    };
    if constexpr (KIND == SLAB_KERR_E) {
        const bool all_valid = __all(c.base >= 0 && c.base + K - 1 <= N);
        emu_tile = any_slab && !pure_slab && !extras && !gen && all_valid && sd.ke0 == sd.kD && sd.jD == R(1);
        if (emu_tile) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                if ((slab_mask >> j) & 1u) continue;
                emu_mask |= 1u << j;
                emulate(j);
                sA[j] = sB[j] = sC[j] = R(0);
                ie[j] = R(1);
                ied[j] = 1.0;
                kz[j] = 0.0;
            }
        }
    }

    const int nblocks = PERSIST ? pa.nblocks : 1;
    for (int blk = 0; blk < nblocks; ++blk) {
        const long long li = launch_idx + blk;             // this block's global launch index
        const int pb = (int)((li - sd.launch_begin) & 1);  // the buffer it reads
        if (PERSIST && blk > 0) {
            // both neighbours have written block li - 1: reload the halo cells
            // (outside the interior), which lie in their interiors
            const int2 nb = pa.nbr[tid];
            const long long deadline = (long long)__builtin_amdgcn_s_memrealtime() + pa.spin_ticks;
            bool ok = true;
            if (nb.x >= 0) ok = wait_flag(pa.flags + nb.x, (unsigned long long)li, deadline);
            if (ok && nb.y >= 0) ok = wait_flag(pa.flags + nb.y, (unsigned long long)li, deadline);
            if (!ok) {
                if (lane == 0) __hip_atomic_store(pa.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int n = c.base + j;
                if (((interior_mask >> j) & 1u) || n < 0 || n > N) continue;
                E[j] = ld_dev(sd.E[pb] + n);
                H[j] = ld_dev(sd.H[pb] + n);
                if (KAHAN) { cE[j] = ld_dev(sd.cE[pb] + n); cH[j] = ld_dev(sd.cH[pb] + n); }
                if ((slab_mask >> j) & 1u) {
                    const int i = n - F;
                    D[j] = ld_dev(sd.D[pb] + i); P[j] = ld_dev(sd.P[pb] + i); Pm[j] = ld_dev(sd.Pm[pb] + i);
                    if (KAHAN) { cD[j] = ld_dev(sd.cD[pb] + i); cP[j] = ld_dev(sd.cP[pb] + i); }
                }
                if ((strip_mask >> j) & 1u) {
                    const int i = n < Ls ? n : Ls + (n - Rs);
                    psE[j] = ld_dev(sd.psiE[pb] + i); psH[j] = ld_dev(sd.psiH[pb] + i);
                }
                if ((emu_mask >> j) & 1u) emulate(j);
            }
        }
        // valid local steps of this block: t = gstep + s - t_begin in [0, T)
        const long long gs = gstep0 + (long long)blk * S;
        c.gstep0 = gs;
        c.s_lo = (int)max(0LL, c.t_begin - gs);
        c.s_hi = (int)min((long long)nsteps, (long long)c.T + c.t_begin - gs);
        c.t_first = (int)(gs + c.s_lo - c.t_begin);
        constexpr int GM = GROUP ? M_GROUP : 0;
        c.emu_mask = emu_mask;
        if (gen)
            run_steps<R, K, KAHAN, KIND, M_MIXED | M_STRIP | M_SRC | M_GENERAL | GM>(TILE_ARGS, c, sd);
        else if ((pure_slab || emu_tile) && !extras)
            run_steps<R, K, KAHAN, KIND, M_PURE_SLAB | GM>(TILE_ARGS, c, sd);
        else if (!any_slab && !extras)
            run_steps<R, K, KAHAN, KIND, GM>(TILE_ARGS, c, sd);
        else if (!any_slab)
            run_steps<R, K, KAHAN, KIND, M_STRIP | M_SRC | M_PROBE | GM>(TILE_ARGS, c, sd);
        else
            run_steps<R, K, KAHAN, KIND, M_MIXED | M_STRIP | M_SRC | M_PROBE | GM>(TILE_ARGS, c, sd);

        // ---------------- write back the interior ----------------
        // (persistent: device-scope stores, which the neighbours' halo loads see)
        const int po = pb ^ 1;
        auto put = [&](R* p, R v) {
            if constexpr (PERSIST) st_dev(p, v); else *p = v;
        };
#pragma unroll
        for (int j = 0; j < K; ++j) {
            if (!((interior_mask >> j) & 1u)) continue;
            const int n = c.base + j;
            put(sd.E[po] + n, E[j]);
            put(sd.H[po] + n, H[j]);
            if (KAHAN) {
                // an emulated vacuum cell carries its compensation in cD: E - cE = D - cD
                const R ce = ((emu_mask >> j) & 1u) ? (R)((double)E[j] - ((double)D[j] - (double)cD[j])) : cE[j];
                put(sd.cE[po] + n, ce);
                put(sd.cH[po] + n, cH[j]);
            }
            if ((slab_mask >> j) & 1u) {
                const int i = n - F;
                put(sd.D[po] + i, D[j]); put(sd.P[po] + i, P[j]); put(sd.Pm[po] + i, Pm[j]);
                if (KAHAN) { put(sd.cD[po] + i, cD[j]); put(sd.cP[po] + i, cP[j]); }
            }
            if ((strip_mask >> j) & 1u) {
                const int i = n < Ls ? n : Ls + (n - Rs);
                put(sd.psiE[po] + i, psE[j]); put(sd.psiH[po] + i, psH[j]);
            }
        }
        if constexpr (PERSIST) {
            // publish: the workgroup-scope release waits until every lane's
            // device-scope stores above have completed, then the flag goes out
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            if (lane == 0) st_dev(pa.flags + tid, (unsigned long long)(li + 1));
        }
    }
}

// Tile t runs as wave t % kTileWaves of workgroup t / kTileWaves, and
// consecutive workgroups go round-robin to the 8 XCDs (workgroup b runs on
// XCD b % 8), so the costly slab tiles, which are contiguous along the grid,
// spread over every XCD.  (A contiguous run of tiles per XCD keeps
// neighbours in one L2 but put all of config 2's slab on three XCDs: 25.4 us
// launches against 18.0 us here.)  Four tiles per workgroup rather than one
// cut the workgroups to dispatch by four: config 2 235.5 -> 244.0 Gcell/s
// (FP32), 239.5 -> 243.4 (FP64); two or eight per workgroup measured ~198
// (FP32) and ~181 (FP64).  A tile entry is (sim | narrow << 24, first
// interior cell): narrow tiles hold KW/2 cells per lane, the others KW, so
// tiles with costly slab cells can be made narrower than plain ones within
// one launch.
template <typename R, int KW, bool KAHAN, int KIND, bool PERSIST = false, bool GROUP = false>
__device__ __forceinline__ void field_tile(const SimDev<R>* __restrict__ sims, const int2* __restrict__ tiles,
                                           int tid, long long launch_idx, long long gstep0, int nsteps, int S,
                                           const PersistArgs& pa) {
    const int2 tl = tiles[tid];
    const SimDev<R>& sd = sims[tl.x & 0xffffff];
    if constexpr (GROUP) {
        tile_body<R, KW, KAHAN, KIND, false, true>(sd, tl.y, launch_idx, gstep0, nsteps, S, tid, pa);
        return;
    }
    if constexpr (KW > 1 && !PERSIST) {
        if (tl.x >> 24) {
            tile_body<R, KW / 2, KAHAN, KIND, false>(sd, tl.y, launch_idx, gstep0, nsteps, S, tid, pa);
            return;
        }
    }
    tile_body<R, KW, KAHAN, KIND, PERSIST>(sd, tl.y, launch_idx, gstep0, nsteps, S, tid, pa);
}

// GROUP: the halo-exchange instance; its tile list holds whole workgroups
// (ntiles a multiple of kTileWaves, every wave of a workgroup from one
// simulation), so every wave reaches the exchange barriers.
template <typename R, int KW, bool KAHAN, int KIND, bool PERSIST, bool GROUP = false>
__global__ __launch_bounds__(64 * kTileWaves) void stencil_tiles(const SimDev<R>* __restrict__ sims,
                                                                 const int2* __restrict__ tiles, int ntiles,
                                                                 long long launch_idx, long long gstep0,
                                                                 int nsteps, int S, PersistArgs pa) {
    // wave-uniform tile index, so the descriptors come in through scalar loads
    const int tid = kTileWaves * (int)blockIdx.x + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (tid >= ntiles) return;
    field_tile<R, KW, KAHAN, KIND, PERSIST, GROUP>(sims, tiles, tid, launch_idx, gstep0, nsteps, S, pa);
}

// Host-side launcher (called from host.hip).
// With profiling events the launch goes through hipExtLaunchKernelGGL, which
// stamps ev0/ev1 from the kernel's own dispatch packet (kernel execution time
// only, the figure a kernel trace reports).
template <typename R, int K, bool KAHAN, int KIND, bool PERSIST, bool GROUP = false>
static hipError_t launch_kp(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                            long long gstep0, int nsteps, int S, const PersistArgs& pa, hipStream_t st,
                            hipEvent_t ev0, hipEvent_t ev1) {
    if (GROUP && ntiles % kTileWaves) return hipErrorInvalidValue;
    const int grid = (ntiles + kTileWaves - 1) / kTileWaves;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL((stencil_tiles<R, K, KAHAN, KIND, PERSIST, GROUP>), dim3(grid), dim3(64 * kTileWaves), 0,
                              st, ev0, ev1, 0, sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa);
    else
        hipLaunchKernelGGL((stencil_tiles<R, K, KAHAN, KIND, PERSIST, GROUP>), dim3(grid), dim3(64 * kTileWaves), 0, st,
                           sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa);
    return hipGetLastError();
}

// Persistent instances exist for K <= kPersistMaxK (the host declines wider tiles).
constexpr int kPersistMaxK = 4;

template <typename R, int K, bool KAHAN, int KIND>
static hipError_t launch_k(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                           long long gstep0, int nsteps, int S, const PersistArgs& pa, hipStream_t st,
                           hipEvent_t ev0, hipEvent_t ev1) {
    if (pa.halo > 0) {
        // halo-exchange instances: K = 2 and 4 (hx / K <= kHaloMaxLanes)
        if constexpr (K == 2 || K == 4) {
            if (pa.halo % K || pa.halo / K > kHaloMaxLanes || pa.nblocks > 1) return hipErrorInvalidValue;
            return launch_kp<R, K, KAHAN, KIND, false, true>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa,
                                                            st, ev0, ev1);
        }
        return hipErrorInvalidValue;
    }
    if (pa.nblocks > 1) {
        if constexpr (K <= kPersistMaxK)
            return launch_kp<R, K, KAHAN, KIND, true>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st,
                                                     ev0, ev1);
        return hipErrorInvalidValue;
    }
    return launch_kp<R, K, KAHAN, KIND, false>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st, ev0, ev1);
}

template <typename R, bool KAHAN, int KIND>
hipError_t launch_kind(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                       long long gstep0, int nsteps, int K, int S, const PersistArgs& pa, hipStream_t st,
                       hipEvent_t ev0, hipEvent_t ev1) {
    switch (K) {
        case 1: return launch_k<R, 1, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st, ev0, ev1);
        case 2: return launch_k<R, 2, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st, ev0, ev1);
        case 4: return launch_k<R, 4, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st, ev0, ev1);
        case 8: return launch_k<R, 8, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, pa, st, ev0, ev1);
        default: return hipErrorInvalidValue;
    }
}

// Workgroups of the persistent field kernel for K cells per lane that one
// CU holds at once (register and LDS limits): a persistent launch needs
// every tile resident.  0 where no persistent instance exists.
template <typename R, bool KAHAN, int KIND>
int field_blocks_per_cu(int K) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    switch (K) {
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, stencil_tiles<R, 1, KAHAN, KIND, true>, 64 * kTileWaves, 0); break;
        case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, stencil_tiles<R, 2, KAHAN, KIND, true>, 64 * kTileWaves, 0); break;
        case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, stencil_tiles<R, 4, KAHAN, KIND, true>, 64 * kTileWaves, 0); break;
        default: break;
    }
    return e == hipSuccess ? nb : 0;
}

}  // namespace fdtd
