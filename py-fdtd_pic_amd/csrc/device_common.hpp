// Device-side types and helpers shared by the stencil and PIC kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdtd {

constexpr int kMaxProbes = 16;

// Per-simulation descriptor as the kernels see it.  One per simulation,
// read by the tile kernel with wave-uniform (scalar) loads.  State arrays
// are ping-ponged: a launch reads [par] and writes [par^1].
template <typename R>
struct SimDev {
    int32_t N, T, Ls, Rs, F, Rr;
    int32_t slab_kind, ade, src_cell, tfsf;
    int32_t nprobe, snap_interval, snap_rows, has_current;
    int64_t launch_begin;     // global launch index of this simulation's first launch
    int64_t t_begin;          // global step at which the simulation's step 0 runs
    int32_t probe_cell[kMaxProbes], probe_t0[kMaxProbes], probe_t1[kMaxProbes];
    R ke0, kh0, keS, khS, kD, jD;
    double kerr_a, kerr_b, kerr_c, kerr_e0, kerr_lin, kerr_chi3;
    R *E[2], *H[2], *cE[2], *cH[2];             // N+1 cells
    R *D[2], *P[2], *Pm[2], *cD[2], *cP[2];     // R-F slab cells (Pm holds V in FP32)
    R *psiE[2], *psiH[2];                       // strip cells
    const R *s_ke, *s_kh, *s_bE, *s_cE, *s_CbE, *s_bH, *s_cH, *s_C2H;
    const R *sl_A, *sl_B, *sl_C, *sl_e;         // FP64: A,B,C ; FP32: -B, 1-A-B, C
    const R *srcE, *srcH;
    double *probe_out;                          // nprobe x T, FP64 (an FP32 field as hi - c of its
                                                // Kahan pair)
    R *snap;                                    // snap_rows x (N+1)
    double* Hprev;                              // N+1 or null: H' before the last H update of a launch
                                                // (compensated, FP64), for the energy sum
    const R *J;                                 // N+1 (PIC current, read every step)
    // with a beam attached the field step forms J' itself from the beam's
    // per-cell CIC sums (j_from_sums), so it does not wait for the scatter
    const double* pS;                           // 4 per cell, or null
    const int32_t* prange[2];                   // the beam's live range per parity
};

// One simulation's energy sum (energy.hip): the launch's final E, H (and
// FP32 compensations), H' before its last H update, and where W goes.
template <typename R>
struct EnergyJob {
    const R *E, *H, *cE, *cH;         // cE, cH null in FP64
    const double* Hprev;
    double* out;
    int N;
};

enum SlabKind { SLAB_NONE = 0, SLAB_DIEL = 1, SLAB_LINEAR = 2, SLAB_KERR_REF = 3, SLAB_KERR_E = 4 };

// A field launch that runs nblocks blocks of S steps instead of one.  Every
// tile stays resident for the whole launch with its window in registers:
// after a block it writes its interior (as a one-block launch does) and
// publishes flags[tile] = the global index of the block it finished; before
// the next block it waits until both neighbour tiles (nbr: left, right; -1
// none) have finished the same block, then reloads only its halo cells, which
// lie in their interiors (the host guarantees S <= U, uniform tiles, and no
// more tiles than the device holds at once).  A wait that passes
// spin_ticks of the 100 MHz real-time counter sets *err and ends the tile
// (a bounded wait: the launch always drains).  nblocks = 1 and flags = null
// is the plain one-block launch.
struct PersistArgs {
    int nblocks;
    unsigned long long* flags;
    const int2* nbr;
    int32_t* err;
    long long spin_ticks;
    int halo;        // > 0: the workgroup halo-exchange instance, exchange period (cells = steps)
};

// Particle beam of one simulation.  Particles are kept sorted by cell
// (stable), structure of arrays, ping-ponged per step like the fields;
// start[c] is the first particle of cell c (start[N+1] = np).  Live
// particles occupy cells [2, N-3]; cells 1 and N-2 park absorbed ones, which
// keep their index from then on.
constexpr int kPicGMax = 64;        // cells per push/count group (LDS tables; the last group of a
                                    // range may take one more)
// push/count waves per workgroup (a CU holds a limited number of workgroups;
// single-wave workgroups would cap the resident waves; 8 measured slower:
// config 3 push/count 13.0 against 11.6 us); the host pads the work list to
// whole workgroups of this many waves
constexpr int kPushWaves = 4;
constexpr int kXMove = 8;           // move_block threads per push/count block boundary
constexpr int kBatch = 2;           // chunks of 64 particles fetched per memory round trip
// pic_push_cell's LDS rows (dynamic shared memory): one pushed 64-particle
// chunk each (id, move and rank, frac, ux, uy, uz), xr rows for each of a
// wave's two cells; at most 8 in FP32, 4 in FP64 (about 96 KiB per workgroup)
template <typename R>
constexpr size_t kOvfRowBytes = 64 * (2 * sizeof(int32_t) + 4 * sizeof(R));
template <typename R>
constexpr int kPushCellMaxRows = sizeof(R) == 4 ? 8 : 4;
template <typename R>
struct BeamDev {
    int64_t launch_begin;  // the simulation's, copied: the parity without a load of the sim descriptor
    int32_t sim, N, np, G;
    int32_t nwaves;      // push/count waves of this beam (grid-stride over cell groups; one-cell
                         // groups: whole workgroups of kPushWaves waves)
    int32_t nnode;       // scatter node blocks of this beam (grid-stride over nodes)
    int32_t nmove;       // node blocks for its staged movers and parked particles (move_block)
    R a;                 // q dt / (2 m c)
    R jscale;            // Z0 q w c: J' of one particle per unit u_x/gamma
    R Sc;                // c dt / dz
    int32_t* cell[2];    // multi-cell beams only (a one-cell beam's cells follow from start[])
    R *frac[2], *ux[2], *uy[2], *uz[2];
    int32_t* id[2];
    int32_t* start[2];   // N+2
    int32_t* range[2];   // [cmin, cmax] of the live particles' cells (empty: cmin > cmax)
    int32_t* urange[2];  // union of the live ranges so far (where J and start[] can change)
    int32_t* park;       // [2]: the previous step's start[2] and start[N-2] (push/count -> scatter):
                         // particles below / at or above them are parked and in both buffers already
    // per-cell results of the push/count pass, valid on the live range +-1
    int32_t *cup, *cdown;  // particles moving up / down out of the cell
    double* S;             // 4 per cell: J' sums onto nodes c-1, c, c+1, c+2
    int32_t* err;        // set when an index check fails (never expected)
};

// J'[m] of a beam from its per-cell CIC sums S (4 per cell: onto nodes
// c-1 .. c+2), which exist on the live range [cmin, cmax] +-1 this step:
// J'[m] = S+2(m-2) + S+1(m-1) + S0(m) + S-1(m+1), summed in this order.
template <typename P>
__device__ __forceinline__ double j_from_sums(P S, int cmin, int cmax, int m) {
    auto Sq = [&](int c, int w) {
        return (cmin <= cmax && c >= cmin - 1 && c <= cmax + 1) ? S[4 * (size_t)c + w] : 0.0;
    };
    return Sq(m - 2, 3) + Sq(m - 1, 2) + Sq(m, 1) + Sq(m + 1, 0);
}

// ---- cross-lane shifts inside one wavefront (DPP, GFX9 wave_shr/wave_shl) ----
// shr1: lane i receives lane i-1 (lane 0 receives 0); shl1: lane i receives lane i+1.
__device__ __forceinline__ int dpp_shr1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_shl1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float shr1(float v) { return __int_as_float(dpp_shr1_i(__float_as_int(v))); }
__device__ __forceinline__ float shl1(float v) { return __int_as_float(dpp_shl1_i(__float_as_int(v))); }
__device__ __forceinline__ double shr1(double v) {
    long long b = __double_as_longlong(v);
    int lo = dpp_shr1_i((int)(b & 0xffffffffll)), hi = dpp_shr1_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double shl1(double v) {
    long long b = __double_as_longlong(v);
    int lo = dpp_shl1_i((int)(b & 0xffffffffll)), hi = dpp_shl1_i((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Kahan-compensated s += inc; (s - c) is the compensated value.
template <typename R>
__device__ __forceinline__ void kahan_add(R& s, R& c, R inc) {
    R y = inc - c;
    R t = s + y;
    c = (t - s) - y;
    s = t;
}

// a * b + c.  FP64: rounded once (an explicit FMA); the FP64 translation
// units are compiled without FMA contraction (-ffp-contract=off), so every
// fused product is one written out here and a cell's update is the same
// sequence of roundings in every tile class and every (K, S) tiling.  FP32:
// the plain expression, fused as the compiler chooses (the FP32 field units
// keep contraction on; their results are held to a tolerance, and the
// compensated sums absorb the difference).
__device__ __forceinline__ double fmr(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmr(float a, float b, float c) { return a * b + c; }

// Kahan-compensated s += a * b (FP64: a * b - c rounded once).
template <typename R>
__device__ __forceinline__ void kahan_fma(R& s, R& c, R a, R b) {
    R y = fmr(a, b, -c);
    R t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---- Kerr, reference path: Re(roots[0]) of a x^3 + b x^2 + c x + d (FP64) ----
// Restates the 1728.org closed form vendored at CubicEquationSolver.py:29-105
// (branch order and tests kept), as consumed by Nonlin_Cubic_Solver
// (BaseFDTD11.py:834-838): only roots[0] is used, |d| <= 1e-8 gives 0.
__device__ inline double cubic_root0_1728(double a, double b, double c, double d) {
    if (!(fabs(d) > 1e-8)) return 0.0;
    if (a == 0.0 && b == 0.0) return (-d * 1.0) / c;
    if (a == 0.0) {
        double disc = c * c - 4.0 * b * d;
        if (disc >= 0.0) return (-c + sqrt(disc)) / (2.0 * b);
        return -c / (2.0 * b);               // real part of the complex pair
    }
    const double f = ((3.0 * c / a) - (pow(b, 2.0) / pow(a, 2.0))) / 3.0;
    const double g = (((2.0 * pow(b, 3.0)) / pow(a, 3.0)) - ((9.0 * b * c) / pow(a, 2.0)) + (27.0 * d / a)) / 27.0;
    const double h = pow(g, 2.0) / 4.0 + pow(f, 3.0) / 27.0;
    if (f == 0.0 && g == 0.0 && h == 0.0) {
        if ((d / a) >= 0.0) return pow(d / (1.0 * a), 1.0 / 3.0) * -1.0;
        return pow(-d / (1.0 * a), 1.0 / 3.0);
    }
    if (h <= 0.0) {
        const double i = sqrt((pow(g, 2.0) / 4.0) - h);
        const double j = pow(i, 1.0 / 3.0);
        const double k = acos(-(g / (2.0 * i)));
        return 2.0 * j * cos(k / 3.0) - (b / (3.0 * a));
    }
    const double sh = sqrt(h);
    const double Rv = -(g / 2.0) + sh;
    const double S = Rv >= 0.0 ? pow(Rv, 1.0 / 3.0) : pow(-Rv, 1.0 / 3.0) * -1.0;
    const double Tv = -(g / 2.0) - sh;
    const double U = Tv >= 0.0 ? pow(Tv, 1.0 / 3.0) : pow(-Tv, 1.0 / 3.0) * -1.0;
    return (S + U) - (b / (3.0 * a));
}

// ---- Kerr, physical path: real root of chi3 E^3 + eps E = u (chi3 >= 0), FP64 ----
// With x = u/eps and z = chi3 x^2/eps (27 z / 4 is the ratio q^2/(4 p^3/27) that
// sets the sign of the depressed cubic's discriminant), E = x*y where
// y = 1 - z y^3.  For |z| < 0.01 y is the alternating Fuss-Catalan series
// sum_k (-1)^k C(3k,k)/(2k+1) z^k (14 terms, error < 1e-16); otherwise the
// hyperbolic closed form E = -2 s sinh(asinh(3q/(2p s))/3), s = sqrt(p/3).
__device__ inline double kerr_root_e(double u, double eps, double chi3) {
    const double x = u / eps;
    if (chi3 == 0.0) return x;
    const double z = chi3 * x * x / eps;
    if (z < 0.01) {
        double y = 300830572.0;
        y = fma(y, -z, 50067108.0);
        y = fma(y, -z, 8414640.0);
        y = fma(y, -z, 1430715.0);
        y = fma(y, -z, 246675.0);
        y = fma(y, -z, 43263.0);
        y = fma(y, -z, 7752.0);
        y = fma(y, -z, 1428.0);
        y = fma(y, -z, 273.0);
        y = fma(y, -z, 55.0);
        y = fma(y, -z, 12.0);
        y = fma(y, -z, 3.0);
        y = fma(y, -z, 1.0);
        y = fma(y, -z, 1.0);
        return x * y;
    }
    const double p = eps / chi3, q = -u / chi3;
    const double s = sqrt(p / 3.0);
    return -2.0 * s * sinh(asinh((1.5 * q / p) / s) / 3.0);
}

}  // namespace fdtd
