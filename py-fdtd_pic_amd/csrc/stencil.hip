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

template <typename R>
__device__ __forceinline__ R slab_create(const SimDev<R>& sd, R D, R cD, R P, R cP, R e) {
    // FP32: u formed in FP64 from the compensated (hi - lo) values.
    const double u = ((double)D - (double)cD) - ((double)P - (double)cP);
    if (sd.slab_kind == SLAB_KERR_E) return (R)kerr_root_e(u, (double)e, sd.kerr_chi3);
    if (sd.slab_kind == SLAB_KERR_REF) {
        const double dd = (double)D / sd.kerr_e0;
        const double A = cubic_root0_1728(sd.kerr_a, sd.kerr_b, sd.kerr_c, -(fabs(dd) * fabs(dd)));
        return (R)((double)D / (sd.kerr_e0 * sd.kerr_lin + sd.kerr_e0 * sd.kerr_chi3 * A));
    }
    return (R)(u / (double)e);
}

template <typename R, int K, bool KAHAN>
__global__ __launch_bounds__(64) void stencil_tiles(const SimDev<R>* __restrict__ sims,
                                                    const int2* __restrict__ tiles, int ntiles,
                                                    int tiles_per_xcd, long long launch_idx,
                                                    long long gstep0, int nsteps, int S) {
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
    const bool dform = sd.slab_kind >= SLAB_LINEAR;

    R E[K], H[K], cE[K], cH[K], ke[K], kh[K];
    R D[K], P[K], Pm[K], cD[K], cP[K], sA[K], sB[K], sC[K], se[K];
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
        se[j] = R(1);
        psE[j] = psH[j] = bE[j] = cEc[j] = CbE[j] = bH[j] = cHc[j] = C2H[j] = R(0);
        const bool inslab = valid && n >= F && n < Rr;
        if (inslab && sd.slab_kind == SLAB_DIEL) { ke[j] = sd.keS; kh[j] = sd.khS; }
        if (inslab && dform) {
            const int i = n - F;
            slab_mask |= 1u << j;
            D[j] = sd.D[par][i]; P[j] = sd.P[par][i]; Pm[j] = sd.Pm[par][i];
            if (KAHAN) { cD[j] = sd.cD[par][i]; cP[j] = sd.cP[par][i]; }
            sA[j] = sd.sl_A[i]; sB[j] = sd.sl_B[i]; sC[j] = sd.sl_C[i]; se[j] = sd.sl_e[i];
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
    const bool tile_slab = __any(slab_mask != 0);
    const bool tile_strip = __any(strip_mask != 0);
    const int tile_end = t0 + 64 * K;
    const int src = sd.src_cell;
    const bool tile_src = src >= t0 && src < tile_end;
    const int src_lane = (src - t0) / K, src_j = (src - t0) % K;
    const bool tile_tf = sd.tfsf && (src - 1) >= t0 && (src - 1) < tile_end;
    const int tf_lane = (src - 1 - t0) / K, tf_j = (src - 1 - t0) % K;
    unsigned probe_bits = 0;
    for (int p = 0; p < sd.nprobe; ++p) {
        const int c = sd.probe_cell[p];
        if (c >= out0 && c < out0 + U && c <= N) probe_bits |= 1u << p;
    }

    for (int s = 0; s < nsteps; ++s) {
        const long long tl64 = gstep0 + s - sd.t_begin;
        if (tl64 < 0 || tl64 >= sd.T) continue;
        const int t = (int)tl64;
        R Jv[K];
#pragma unroll
        for (int j = 0; j < K; ++j) Jv[j] = R(0);
        if (sd.J) {
#pragma unroll
            for (int j = 0; j < K; ++j) {
                const int n = base + j;
                if (n >= 0 && n <= N) Jv[j] = sd.J[n];
            }
        }
        // ---------------- E half ----------------
        const R Hl = shr1(H[K - 1]);
        R dHs[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const R Hm = (j == 0) ? Hl : H[j - 1];
            dHs[j] = H[j] - Hm;
        }
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const R dH = dHs[j];
            const bool is_slab = tile_slab && ((slab_mask >> j) & 1u);
            if (is_slab) {
                if (sd.ade) {
                    if constexpr (sizeof(R) == 4) {   // velocity form: V' = -B V - (1-A-B) P + C E ; P += V'
                        const R Vn = sA[j] * Pm[j] - sB[j] * P[j] + sC[j] * E[j];
                        Pm[j] = Vn;
                        if (KAHAN) kahan_add(P[j], cP[j], Vn); else P[j] = P[j] + Vn;
                    } else {
                        const R Pn = sA[j] * P[j] + sB[j] * Pm[j] + sC[j] * E[j];
                        Pm[j] = P[j];
                        P[j] = Pn;
                    }
                }
                const R inc = (dH - sd.jD * Jv[j]) * sd.kD;
                if (KAHAN) kahan_add(D[j], cD[j], inc); else D[j] = D[j] + inc;
                E[j] = slab_create<R>(sd, D[j], cD[j], P[j], cP[j], se[j]);
                if (tile_strip && ((strip_mask >> j) & 1u)) psE[j] = bE[j] * psE[j] + cEc[j] * dH;
            } else {
                const R inc = (dH - Jv[j]) * ke[j];
                if (KAHAN) kahan_add(E[j], cE[j], inc); else E[j] = E[j] + inc;
                if (tile_strip && ((strip_mask >> j) & 1u)) {
                    psE[j] = bE[j] * psE[j] + cEc[j] * dH;
                    E[j] = E[j] - CbE[j] * psE[j];
                }
            }
        }
        if (tile_src) {
            const R sv = sd.srcE[t];
            if (lane == src_lane) add_at<R, K>(E, src_j, sv);
        }
        // ---------------- H half ----------------
        if (tile_tf) {
            const R hv = sd.srcH[t];
            if (lane == tf_lane) add_at<R, K>(H, tf_j, -hv);
        }
        const R Er = shl1(E[0]);
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const R Ep = (j == K - 1) ? Er : E[j + 1];
            const R dE = Ep - E[j];
            const R inc = dE * kh[j];
            if (KAHAN) kahan_add(H[j], cH[j], inc); else H[j] = H[j] + inc;
            if (tile_strip && ((strip_mask >> j) & 1u)) {
                psH[j] = bH[j] * psH[j] + cHc[j] * dE;
                H[j] = H[j] + C2H[j] * psH[j];
            }
        }
        // ---------------- probes / snapshots ----------------
        if (probe_bits) {
            for (int p = 0; p < sd.nprobe; ++p) {
                if (!((probe_bits >> p) & 1u)) continue;
                if (t < sd.probe_t0[p] || t > sd.probe_t1[p]) continue;
                const int rel = sd.probe_cell[p] - t0;
                if (lane == rel / K) sd.probe_out[(long long)p * sd.T + t] = pick<R, K>(E, rel % K);
            }
        }
        if (sd.snap_interval > 0 && t > 0 && (t % sd.snap_interval) == 0) {
            const int row = t / sd.snap_interval;
            if (row < sd.snap_rows) {
#pragma unroll
                for (int j = 0; j < K; ++j)
                    if ((interior_mask >> j) & 1u) sd.snap[(long long)row * (N + 1) + base + j] = E[j];
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
template <typename R, int K, bool KAHAN>
static hipError_t launch_k(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                           long long gstep0, int nsteps, int S, hipStream_t st) {
    const int per_xcd = (ntiles + 7) / 8;
    const int grid = per_xcd * 8;
    hipLaunchKernelGGL((stencil_tiles<R, K, KAHAN>), dim3(grid), dim3(64), 0, st, sims, tiles, ntiles,
                       per_xcd, launch_idx, gstep0, nsteps, S);
    return hipGetLastError();
}

template <typename R, bool KAHAN>
hipError_t launch_stencil(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                          long long gstep0, int nsteps, int K, int S, hipStream_t st) {
    switch (K) {
        case 1: return launch_k<R, 1, KAHAN>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 2: return launch_k<R, 2, KAHAN>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 4: return launch_k<R, 4, KAHAN>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        case 8: return launch_k<R, 8, KAHAN>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, st);
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_stencil<double, false>(const SimDev<double>*, const int2*, int, long long,
                                                  long long, int, int, int, hipStream_t);
template hipError_t launch_stencil<float, true>(const SimDev<float>*, const int2*, int, long long,
                                                long long, int, int, int, hipStream_t);
template hipError_t launch_stencil<float, false>(const SimDev<float>*, const int2*, int, long long,
                                                 long long, int, int, int, hipStream_t);

}  // namespace fdtd
