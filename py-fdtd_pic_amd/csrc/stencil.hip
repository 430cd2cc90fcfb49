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
#include "device_common.hpp"

namespace fdtd {

template <typename R, int K>
__device__ __forceinline__ R pick(const R (&a)[K], int j) {
    R v = a[0];
#pragma unroll
    for (int i = 1; i < K; ++i) v = (i == j) ? a[i] : v;
    return v;
}
template <typename R, int K>
__device__ __forceinline__ void add_at(R (&a)[K], int j, R x) {
#pragma unroll
    for (int i = 0; i < K; ++i)
        if (i == j) a[i] += x;
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
// dst with lane `lane` replaced by the uniform value v (a lane-id select)
template <typename R>
__device__ __forceinline__ R lane_write(R v, int lane, R dst) {
    return ((int)threadIdx.x == lane) ? v : dst;
}
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

// ADE + D update for one slab cell (no branches: ade == 0 is encoded as
// zero coefficients).  FP64 keeps the reference form P' = A P + B Pm + C E
// (ADE_PolarisationCurrent_Ex :620-631); FP32 keeps the velocity form
// V' = -B V - (1-A-B) P + C E, P += V' (sA = -B, sB = 1-A-B).
template <typename R, bool KAHAN>
__device__ __forceinline__ double slab_u(R kD, R jD, R dH, R Jv, R E, R& D, R& cD, R& P, R& Pm, R& cP,
                                         R sA, R sB, R sC) {
    if constexpr (sizeof(R) == 4) {
        const R Vn = sA * Pm - sB * P + sC * E;
        Pm = Vn;
        if (KAHAN) kahan_add(P, cP, Vn); else P = P + Vn;
    } else {
        const R Pn = sA * P + sB * Pm + sC * E;
        Pm = P;
        P = Pn;
    }
    const R inc = (dH - jD * Jv) * kD;
    if (KAHAN) kahan_add(D, cD, inc); else D = D + inc;
    // FP32: u = (D - cD) - (P - cP) from the compensated values, the hi part
    // formed in FP32 (exact when D and P are within a factor 2, else rounded
    // like the stored E), the compensation added in FP64.  FP64: cD = cP = 0.
    if constexpr (sizeof(R) == 4)
        return KAHAN ? (double)(D - P) + (double)(cP - cD) : (double)(D - P);
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
__device__ __forceinline__ void kerr_fixup(R (&E)[K], const double (&x)[K], const double (&z)[K],
                                           const double (&ied)[K], double chi3) {
    double zmax = z[0];
#pragma unroll
    for (int j = 1; j < K; ++j) zmax = fmax(zmax, z[j]);
    if (__any(zmax >= kKerrSeriesCut)) {
#pragma unroll
        for (int j = 0; j < K; ++j)
            if (z[j] >= kKerrSeriesCut) E[j] = (R)kerr_root_e(x[j] / ied[j], 1.0 / ied[j], chi3);
    }
}

template <typename R, int K, bool KAHAN, int KIND>
__global__ __launch_bounds__(64) void stencil_tiles(const SimDev<R>* __restrict__ sims,
                                                    const int2* __restrict__ tiles, int ntiles,
                                                    int tiles_per_xcd, long long launch_idx,
                                                    long long gstep0, int nsteps, int S) {
    constexpr bool DFORM = KIND >= SLAB_LINEAR;
    // XCD-aware order: block b runs on XCD b%8; give each XCD a contiguous run of tiles.
    const int b = blockIdx.x;
    const int tid = (b & 7) * tiles_per_xcd + (b >> 3);
    if (tid >= ntiles) return;
    const int2 tl = tiles[tid];
    const SimDev<R>& sd = sims[tl.x];
    const int lane = threadIdx.x;
    const int U = 64 * K - 2 * S;
    const int out0 = tl.y;
    const int t0 = out0 - S;                 // first cell of the tile
    const int base = t0 + lane * K;          // my first cell
    const int N = sd.N;
    const int par = (int)((launch_idx - sd.launch_begin) & 1);
    const int F = sd.F, Rr = sd.Rr, Ls = sd.Ls, Rs = sd.Rs;
    // Everything the step loop reads from the descriptor is copied into
    // registers here: the loop's global stores (probes, snapshots) could
    // alias the descriptor, so the compiler would otherwise reload it per step.
    const double chi3 = sd.kerr_chi3;
    const R kD = sd.kD, jD = sd.jD;
    const long long t_begin = sd.t_begin;
    const int T = sd.T;
    const int nprobe = sd.nprobe, snap_interval = sd.snap_interval, snap_rows = sd.snap_rows;
    const R* __restrict__ srcE = sd.srcE;
    const R* __restrict__ srcH = sd.srcH;
    const R* __restrict__ Jc = sd.J;
    R* __restrict__ probe_out = sd.probe_out;
    R* __restrict__ snap = sd.snap;

    R E[K], H[K], cE[K], cH[K], ke[K], kh[K];
    R D[K], P[K], Pm[K], cD[K], cP[K], sA[K], sB[K], sC[K], ie[K];
    R psE[K], psH[K], bE[K], cEc[K], CbE[K], bH[K], cHc[K], C2H[K];
    unsigned slab_mask = 0, strip_mask = 0, interior_mask = 0;

    const R* Ein = sd.E[par];
    const R* Hin = sd.H[par];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int n = base + j;
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
    }
    double ied[K], kz[K];      // 1/e and chi3/e per cell, FP64
#pragma unroll
    for (int j = 0; j < K; ++j) { ied[j] = (double)ie[j]; kz[j] = chi3 * ied[j]; }
    constexpr unsigned full = (K >= 32) ? 0xffffffffu : ((1u << K) - 1u);
    // tile classes (wave-uniform): pure slab / pure interior / mixed
    const bool tile_strip = __any(strip_mask != 0);
    const bool pure_slab = DFORM && __all(slab_mask == full);
    const bool any_slab = DFORM && __any(slab_mask != 0);
    const int tile_end = t0 + 64 * K;
    const int src = sd.src_cell;
    const bool tile_src = src >= t0 && src < tile_end;
    const int src_lane = (src - t0) / K, src_j = (src - t0) % K;
    const bool tile_tf = sd.tfsf && (src - 1) >= t0 && (src - 1) < tile_end;
    const int tf_lane = (src - 1 - t0) / K, tf_j = (src - 1 - t0) % K;
    // probes in this tile's interior: at most two kept in registers (cell
    // offset, window); more than two fall back to the descriptor loop
    unsigned probe_bits = 0;
    for (int p = 0; p < nprobe; ++p) {
        const int c = sd.probe_cell[p];
        if (c >= out0 && c < out0 + U && c <= N) probe_bits |= 1u << p;
    }
    int pr_rel[2] = {0, 0}, pr_t0[2] = {0, 0}, pr_t1[2] = {-1, -1}, pr_id[2] = {0, 0};
    const int npr_local = __popc(probe_bits);
    {
        int k = 0;
        for (int p = 0; p < nprobe && k < 2; ++p)
            if ((probe_bits >> p) & 1u) {
                pr_rel[k] = sd.probe_cell[p] - t0; pr_t0[k] = sd.probe_t0[p]; pr_t1[k] = sd.probe_t1[p];
                pr_id[k] = p; ++k;
            }
    }

    // valid local steps of this launch: t = gstep0 + s - t_begin in [0, T)
    const int s_lo = (int)max(0LL, t_begin - gstep0);
    const int s_hi = (int)min((long long)nsteps, (long long)T + t_begin - gstep0);
    // fast path: a plain tile (all slab or all interior, no CPML strip, source,
    // probe, current or snapshot) runs a loop with no per-step uniform checks
    const bool plain = (pure_slab || !any_slab) && !tile_strip && !tile_src && !tile_tf && npr_local == 0 &&
                       !Jc && snap_interval == 0;
    if (plain && pure_slab) {
        for (int s = s_lo; s < s_hi; ++s) {
            const R Hl = shr1(H[K - 1]);
            double x[K], z[K];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const R dH = H[j] - ((j == 0) ? Hl : H[j - 1]);
                const double u = slab_u<R, KAHAN>(kD, jD, dH, R(0), E[j], D[j], cD[j], P[j], Pm[j], cP[j],
                                                  sA[j], sB[j], sC[j]);
                if constexpr (KIND == SLAB_KERR_E) {
                    x[j] = u * ied[j];
                    z[j] = kz[j] * x[j] * x[j];
                    E[j] = (R)(x[j] * kerr_series<R>(z[j]));
                } else {
                    E[j] = slab_e_simple<R, KIND>(sd, u, D[j], ie[j]);
                }
            }
            if constexpr (KIND == SLAB_KERR_E) kerr_fixup<R, K>(E, x, z, ied, chi3);
            const R Er = shl1(E[0]);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const R inc = (((j == K - 1) ? Er : E[j + 1]) - E[j]) * kh[j];
                if (KAHAN) kahan_add(H[j], cH[j], inc); else H[j] = H[j] + inc;
            }
        }
    } else if (plain) {
        for (int s = s_lo; s < s_hi; ++s) {
            const R Hl = shr1(H[K - 1]);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const R inc = (H[j] - ((j == 0) ? Hl : H[j - 1])) * ke[j];
                if (KAHAN) kahan_add(E[j], cE[j], inc); else E[j] = E[j] + inc;
            }
            const R Er = shl1(E[0]);
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const R inc = (((j == K - 1) ? Er : E[j + 1]) - E[j]) * kh[j];
                if (KAHAN) kahan_add(H[j], cH[j], inc); else H[j] = H[j] + inc;
            }
        }
    } else {
        // source values of this launch, one step per lane (read back with
        // v_readlane), and probe samples collected with v_writelane and
        // stored once after the loop: no memory access on the per-step path
        const int t_first = (int)(gstep0 + s_lo - t_begin);
        R srcEc[kChunks], srcHc[kChunks], pacc[2][kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int tt = t_first + c * 64 + lane;
            const bool in = (c * 64 + lane) < (s_hi - s_lo);
            srcEc[c] = (tile_src && in) ? srcE[tt] : R(0);
            srcHc[c] = (tile_tf && in) ? srcH[tt] : R(0);
            pacc[0][c] = pacc[1][c] = R(0);
        }
        for (int s = s_lo; s < s_hi; ++s) {
            const int t = (int)(gstep0 + s - t_begin);
            const int ix = s - s_lo;
            R Jv[K];
#pragma unroll
            for (int j = 0; j < K; ++j) Jv[j] = R(0);
            if (Jc) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const int n = base + j;
                    if (n >= 0 && n <= N) Jv[j] = Jc[n];
                }
            }
            // ---------------- E half ----------------
            const R Hl = shr1(H[K - 1]);
            R dHs[K];
#pragma unroll
            for (int j = 0; j < K; ++j) dHs[j] = H[j] - ((j == 0) ? Hl : H[j - 1]);

            if (pure_slab && !tile_strip) {
                // straight-line over the K cells: the FP64 chains interleave
                if constexpr (KIND == SLAB_KERR_E) {
                    double x[K], z[K];
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const double u = slab_u<R, KAHAN>(kD, jD, dHs[j], Jv[j], E[j], D[j], cD[j], P[j], Pm[j], cP[j],
                                                          sA[j], sB[j], sC[j]);
                        x[j] = u * ied[j];
                        z[j] = kz[j] * x[j] * x[j];
                        E[j] = (R)(x[j] * kerr_series<R>(z[j]));
                    }
                    kerr_fixup<R, K>(E, x, z, ied, chi3);
                } else {
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const double u = slab_u<R, KAHAN>(kD, jD, dHs[j], Jv[j], E[j], D[j], cD[j], P[j], Pm[j], cP[j],
                                                          sA[j], sB[j], sC[j]);
                        E[j] = slab_e_simple<R, KIND>(sd, u, D[j], ie[j]);
                    }
                }
            } else if (!any_slab && !tile_strip) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const R inc = (dHs[j] - Jv[j]) * ke[j];
                    if (KAHAN) kahan_add(E[j], cE[j], inc); else E[j] = E[j] + inc;
                }
            } else {
                // mixed tile: every cell takes the interior update (CPML psi
                // coefficients are zero off the strips, so no per-cell test);
                // slab cells then take E from D - P through selects, and the
                // Kerr closed form hides behind one wave-level test.
                R Ev[K], cEv[K];
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    Ev[j] = E[j]; cEv[j] = cE[j];
                    const R inc = (dHs[j] - Jv[j]) * ke[j];
                    if (KAHAN) kahan_add(Ev[j], cEv[j], inc); else Ev[j] = Ev[j] + inc;
                }
                if (tile_strip) {
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        psE[j] = bE[j] * psE[j] + cEc[j] * dHs[j];
                        Ev[j] = Ev[j] - CbE[j] * psE[j];
                    }
                }
                if (DFORM && any_slab) {
                    double x[K], z[K];
                    R Es[K];
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const bool sl = (slab_mask >> j) & 1u;
                        R Dn = D[j], cDn = cD[j], Pn = P[j], Pmn = Pm[j], cPn = cP[j];
                        const double u = slab_u<R, KAHAN>(kD, jD, dHs[j], Jv[j], E[j], Dn, cDn, Pn, Pmn, cPn,
                                                          sA[j], sB[j], sC[j]);
                        if constexpr (KIND == SLAB_KERR_E) {
                            x[j] = u * ied[j];
                            z[j] = sl ? kz[j] * x[j] * x[j] : 0.0;
                            Es[j] = (R)(x[j] * kerr_series<R>(z[j]));
                        } else {
                            x[j] = z[j] = 0.0;
                            Es[j] = slab_e_simple<R, KIND>(sd, u, Dn, ie[j]);
                        }
                        D[j] = sl ? Dn : D[j]; cD[j] = sl ? cDn : cD[j];
                        P[j] = sl ? Pn : P[j]; Pm[j] = sl ? Pmn : Pm[j]; cP[j] = sl ? cPn : cP[j];
                    }
                    if constexpr (KIND == SLAB_KERR_E) kerr_fixup<R, K>(Es, x, z, ied, chi3);
#pragma unroll
                    for (int j = 0; j < K; ++j) {
                        const bool sl = (slab_mask >> j) & 1u;
                        E[j] = sl ? Es[j] : Ev[j];
                        cE[j] = sl ? cE[j] : cEv[j];
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < K; ++j) { E[j] = Ev[j]; cE[j] = cEv[j]; }
                }
            }
            if (tile_src) {
                const R sv = lane_read(pick<R, kChunks>(srcEc, ix >> 6), ix & 63);
                if (lane == src_lane) add_at<R, K>(E, src_j, sv);
            }
            // ---------------- H half ----------------
            if (tile_tf) {
                const R hv = lane_read(pick<R, kChunks>(srcHc, ix >> 6), ix & 63);
                if (lane == tf_lane) add_at<R, K>(H, tf_j, -hv);
            }
            const R Er = shl1(E[0]);
            if (!tile_strip) {
#pragma unroll
                for (int j = 0; j < K; ++j) {
                    const R dE = ((j == K - 1) ? Er : E[j + 1]) - E[j];
                    const R inc = dE * kh[j];
                    if (KAHAN) kahan_add(H[j], cH[j], inc); else H[j] = H[j] + inc;
                }
            } else {
#pragma unroll
                for (int j = 0; j < K; ++j) {   // psi coefficients are zero off the strips
                    const R dE = ((j == K - 1) ? Er : E[j + 1]) - E[j];
                    const R inc = dE * kh[j];
                    if (KAHAN) kahan_add(H[j], cH[j], inc); else H[j] = H[j] + inc;
                    psH[j] = bH[j] * psH[j] + cHc[j] * dE;
                    H[j] = H[j] + C2H[j] * psH[j];
                }
            }
            // ---------------- probes / snapshots ----------------
            if (npr_local) {
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    if (k < npr_local) {
                        const R v = lane_read(pick<R, K>(E, pr_rel[k] % K), pr_rel[k] / K);
#pragma unroll
                        for (int c = 0; c < kChunks; ++c)
                            if (c == (ix >> 6)) pacc[k][c] = lane_write(v, ix & 63, pacc[k][c]);
                    }
                if (npr_local > 2) {
                    for (int p = 0; p < nprobe; ++p) {
                        if (!((probe_bits >> p) & 1u) || p == pr_id[0] || p == pr_id[1]) continue;
                        if (t < sd.probe_t0[p] || t > sd.probe_t1[p]) continue;
                        const int rel = sd.probe_cell[p] - t0;
                        if (lane == rel / K) probe_out[(long long)p * T + t] = pick<R, K>(E, rel % K);
                    }
                }
            }
            if (snap_interval > 0 && t > 0 && (t % snap_interval) == 0) {
                const int row = t / snap_interval;
                if (row < snap_rows) {
#pragma unroll
                    for (int j = 0; j < K; ++j)
                        if ((interior_mask >> j) & 1u) snap[(long long)row * (N + 1) + base + j] = E[j];
                }
            }
        }
        if (npr_local) {
#pragma unroll
            for (int k = 0; k < 2; ++k)
                if (k < npr_local) {
#pragma unroll
                    for (int c = 0; c < kChunks; ++c) {
                        const int ix = c * 64 + lane;
                        const int tt = t_first + ix;
                        if (ix < s_hi - s_lo && tt >= pr_t0[k] && tt <= pr_t1[k])
                            probe_out[(long long)pr_id[k] * T + tt] = pacc[k][c];
                    }
                }
        }
    }

    // ---------------- write back the interior ----------------
    const int po = par ^ 1;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (!((interior_mask >> j) & 1u)) continue;
        const int n = base + j;
        sd.E[po][n] = E[j];
        sd.H[po][n] = H[j];
        if (KAHAN) { sd.cE[po][n] = cE[j]; sd.cH[po][n] = cH[j]; }
        if ((slab_mask >> j) & 1u) {
            const int i = n - F;
            sd.D[po][i] = D[j]; sd.P[po][i] = P[j]; sd.Pm[po][i] = Pm[j];
            if (KAHAN) { sd.cD[po][i] = cD[j]; sd.cP[po][i] = cP[j]; }
        }
        if ((strip_mask >> j) & 1u) {
            const int i = n < Ls ? n : Ls + (n - Rs);
            sd.psiE[po][i] = psE[j]; sd.psiH[po][i] = psH[j];
        }
    }
}

// Host-side launcher (called from host.hip).
template <typename R, int K, bool KAHAN, int KIND>
static hipError_t launch_k(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                           long long gstep0, int nsteps, int S, hipStream_t st) {
    const int per_xcd = (ntiles + 7) / 8;
    const int grid = per_xcd * 8;
    hipLaunchKernelGGL((stencil_tiles<R, K, KAHAN, KIND>), dim3(grid), dim3(64), 0, st, sims, tiles, ntiles,
                       per_xcd, launch_idx, gstep0, nsteps, S);
    return hipGetLastError();
}

template <typename R, bool KAHAN, int KIND>
static hipError_t launch_kind(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                              long long gstep0, int nsteps, int K, int S, hipStream_t st) {
    switch (K) {
        case 1: return launch_k<R, 1, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 2: return launch_k<R, 2, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 4: return launch_k<R, 4, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 8: return launch_k<R, 8, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        default: return hipErrorInvalidValue;
    }
}

// kind: 0 (no D-form slab: NONE or DIEL share the DIEL kernel), SLAB_LINEAR,
// SLAB_KERR_REF (FP64 only), SLAB_KERR_E.
template <typename R, bool KAHAN>
hipError_t launch_stencil(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                          long long gstep0, int nsteps, int K, int S, int kind, hipStream_t st) {
    switch (kind) {
        case SLAB_NONE:
        case SLAB_DIEL: return launch_kind<R, KAHAN, SLAB_DIEL>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st);
        case SLAB_LINEAR: return launch_kind<R, KAHAN, SLAB_LINEAR>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st);
        case SLAB_KERR_E: return launch_kind<R, KAHAN, SLAB_KERR_E>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st);
        case SLAB_KERR_REF:
            if constexpr (sizeof(R) == 8)
                return launch_kind<R, KAHAN, SLAB_KERR_REF>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_stencil<double, false>(const SimDev<double>*, const int2*, int, long long,
                                                  long long, int, int, int, int, hipStream_t);
template hipError_t launch_stencil<float, true>(const SimDev<float>*, const int2*, int, long long,
                                                long long, int, int, int, int, hipStream_t);

}  // namespace fdtd
