/*
 * fdtdpic.h -- C ABI of libfdtdpic.so, the MI355X (gfx950) FDTD(+PIC) step loop.
 *
 * The reference drives its loop from Python: Solver_Engine.Integrator*1D
 * (Solver_Engine.py:142-371) calls ~8 numba kernels of BaseFDTD11.py per
 * step on numpy float64 arrays owned by the Variables / CPML_Variables
 * jitclasses (MasterController.py:145-440).  This library keeps the whole
 * step loop on the GPU: the Python drop-in modules (Solver_Engine.py,
 * BaseFDTD11.py, ...) build the same coefficient arrays the reference
 * builds, hand them over once through fdtd_add_sim, advance with fdtd_run,
 * and read fields / probes back with fdtd_read_*.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative FDTD_E_* code;
 *     fdtd_last_error() returns the thread-local message of the last failure.
 *     The library never calls exit(); the reference's sys.exit() paths
 *     (BaseFDTD11.py:36-66, Environment_Setup.py:59-155) are raised by the
 *     Python layer.
 *   - Host arrays are plain pointers with explicit lengths; values are FP64 on
 *     the host side whatever the device precision.
 *   - A context owns a compute stream; the contexts of one device share one
 *     upload stream.  fdtd_run is asynchronous, fdtd_sync waits.
 *     fdtd_add_sim and pic_attach return once their uploads are done,
 *     without waiting for running launches.
 *     Contexts are not thread-safe; use one per host thread.
 */
#ifndef FDTDPIC_H
#define FDTDPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_OK 0
#define FDTD_E_ARG (-1)      /* invalid argument / shape */
#define FDTD_E_HIP (-2)      /* HIP runtime error */
#define FDTD_E_STATE (-3)    /* call not valid in the current state */
#define FDTD_E_NOMEM (-4)    /* device allocation failed or budget exceeded */

#define FDTD_F64 0
#define FDTD_F32 1

/* slab_kind: how E is formed in the slab [F,R) */
#define FDTD_SLAB_NONE 0     /* no slab */
#define FDTD_SLAB_DIEL 1     /* E += keS*(dH-J), H += khS*dE  (IntegratorFreeSpace1D + UpdateCoef) */
#define FDTD_SLAB_LINEAR 2   /* D += kD*(dH-jD*J); E = (D-P)/e          (ADE_DxUpdate + ADE_ExCreate) */
#define FDTD_SLAB_KERR_REF 3 /* D as above; E = D/(e0*lin + e0*chi3*A), A = Re roots[0] of the x=|E|^2 cubic
                                (AcubicFinder + NonLinExUpdate, BaseFDTD11.py:803-877) */
#define FDTD_SLAB_KERR_E 4   /* D as above; E = real root of chi3 E^3 + e E = D - P (physical model) */

#define FDTD_MAX_PROBES 16

/* Fields addressable by fdtd_read_field / fdtd_write_field */
#define FDTD_FIELD_E 0       /* Ex, length N+1 */
#define FDTD_FIELD_H 1       /* Hy, length N+1 */
#define FDTD_FIELD_D 2       /* Dx over the slab, length R-F */
#define FDTD_FIELD_P 3       /* polarisationCurr over the slab */
#define FDTD_FIELD_PM 4      /* previous polarisation P^{n-1} (tempVarPol) over the slab */
#define FDTD_FIELD_PSIE 5    /* psi_Ex, length N+1 (zero outside the strips) */
#define FDTD_FIELD_PSIH 6    /* psi_Hy, length N+1 */
#define FDTD_FIELD_J 7       /* Jx as used by the E update, length N+1 */

/* kernel kinds for fdtd_kernel_timing */
#define FDTD_KERNEL_STENCIL 0  /* field stencil (E/H/ADE/Kerr/CPML) */
#define FDTD_KERNEL_PUSH 1     /* particle gather + Boris push + per-cell counts, ranks, CIC sums */
#define FDTD_KERNEL_SORT 2     /* particle scatter (stable re-sort) + J', cell starts */
#define FDTD_KERNEL_FIELD_SORT 3 /* field stencil and particle scatter in one launch (PIC fusion on) */

typedef struct fdtd_ctx fdtd_ctx;

/*
 * One simulation.  Cells are 0..N.  E is updated as
 *     E[n] += ke(n)*(H[n]-H[n-1]-J[n]);  psi_E = bE psi_E + cE dH;  E -= CbE psi_E
 * and H as
 *     H[n] += kh(n)*(E[n+1]-E[n]);       psi_H = bH psi_H + cH dE;  H += C2H psi_H
 * (ADE_ExUpdate BaseFDTD11.py:663-669, CPML_Psi_e_Update :364-376,
 *  ADE_HyUpdate :640-656, CPML_Psi_m_Update :381-393).
 * Cells outside the two strips [0,Ls) and [Rs,N] use ke0/kh0 (or keS/khS in a
 * DIEL slab); strip cells carry their own coefficients (ke = UpExMat*den_Exdz
 * etc.); a cell that is never updated has ke or kh = 0 there.
 */
typedef struct {
    int32_t N;                 /* cells 0..N */
    int32_t steps;             /* steps this simulation runs (reference P.timeSteps) */
    double ke0, kh0;           /* interior Yee coefficients (UpExMat, UpHyMat) */
    int32_t Ls, Rs;            /* strips [0,Ls) and [Rs,N] */
    /* strip coefficients, length Ls + (N+1-Rs), left strip then right strip */
    const double *s_ke, *s_kh, *s_bE, *s_cE, *s_CbE, *s_bH, *s_cH, *s_C2H;
    /* slab [F,R) */
    int32_t F, R, slab_kind, ade;
    double keS, khS;           /* DIEL slab coefficients */
    double kD, jD;             /* D += kD*(dH - jD*J) */
    const double *sl_A, *sl_B, *sl_C;  /* ADE P' = A P + B Pm + C E (ADE_PolarisationCurrent_Ex :620-631) */
    const double *sl_e;        /* LINEAR: divisor e (eps0); KERR_E: eps_inf per cell */
    double kerr_a, kerr_b, kerr_c;     /* KERR_REF cubic a x^3 + b x^2 + c x + d, d = -(D/e0)^2 */
    double kerr_e0, kerr_lin, kerr_chi3; /* KERR_REF: e0, sqrt(1.2), chi3Stat; KERR_E: chi3 in kerr_chi3 */
    /* source: E[src_cell] += srcE[t]; if tfsf: H[src_cell-1] -= srcH[t] (Solver_Engine.py:307-310) */
    int32_t src_cell, tfsf;
    const double *srcE, *srcH; /* length steps (srcH may be NULL when tfsf == 0) */
    /* probes: record E[probe_cell] after step t for t in [probe_t0, probe_t1] (probeSim :16-54) */
    int32_t nprobe;
    int32_t probe_cell[FDTD_MAX_PROBES];
    int32_t probe_t0[FDTD_MAX_PROBES], probe_t1[FDTD_MAX_PROBES];
    /* snapshots: row t/interval <- E after step t when t>0, t%interval==0, row<rows (vidMake :57-68) */
    int32_t snap_interval, snap_rows;
    /* PIC: 1 if a current array is attached (J is then re-deposited every step) */
    int32_t has_current;
} fdtd_sim_desc_t;

/* ---- context ---------------------------------------------------------- */
/* precision: FDTD_F64 or FDTD_F32 (FP32 fields, Kahan-compensated D, P and
 * fields, FP64 Kerr root).  Replaces the per-call numba dispatch of
 * Solver_Engine.py:294-316 with one device-resident loop. */
int fdtd_ctx_create(int precision, fdtd_ctx **out);
int fdtd_ctx_destroy(fdtd_ctx *ctx);
/* Temporal tiling: K cells per lane (1,2,4,8), S steps per launch (<= 32K-1;
 * up to 64K-h-1 with the workgroup halo exchange of period h, below; the run
 * fails with FDTD_E_ARG when neither holds).  S = 0 picks the library
 * default for K. */
int fdtd_set_tiling(fdtd_ctx *ctx, int K, int S);
/* Mixed widths in one launch: tiles whose window meets a Lorentz/Kerr slab
 * hold K_slab = K/2 cells per lane, the others K (K_slab = K: uniform).  The
 * costly slab cells then get more lanes while plain cells use fewer waves. */
int fdtd_set_tiling_mixed(fdtd_ctx *ctx, int K, int K_slab, int S);
/* PIC step launches: with on = 1 (the default) the particle scatter runs in
 * the same launch as the field step when that launch is small (K = 1 or 2
 * and at most 2,048 tiles; with several slab kinds in one context, the last
 * kind's launch), so a step is two launches instead of three.  A larger
 * field launch keeps its own, since the fused kernel's register count (the
 * field tile's) would cut the scatter's occupancy.  0 always launches them
 * one after the other.  Results are bitwise the same either way.  No
 * counterpart in the reference (which has no particles); it selects how
 * this library schedules its own kernels. */
int fdtd_set_pic_fusion(fdtd_ctx *ctx, int on);
/* LDS rows per cell of the one-cell particle kernel (dense beams): a cell's
 * particles beyond the 128 held in registers are pushed once and their new
 * state waits in rows of 64 until the block's counts are known; a fuller
 * cell takes a counting pass and a second, bitwise identical, pushing pass.
 * rows = 0 (the default) gives a launch whose workgroups fit one per CU the
 * rows its fullest beam needed at attach (plus a quarter) and a larger
 * launch one row; 1..8 forces a count (capped by the device's shared memory,
 * 4 in FP64).  Results are bitwise the same for every count.  No counterpart
 * in the reference (which has no particles). */
int fdtd_set_pic_lds_rows(fdtd_ctx *ctx, int rows);

/* Persistent field launches (on = 1): when the context's tiles all fit on
 * the device at once (one slab kind, uniform tiles, S <= 64K - 2S, no beam),
 * one launch runs up to 4,096 blocks of S steps.  Each tile keeps its window
 * in registers for the whole launch; between blocks it writes its interior,
 * publishes a per-tile flag, waits for its two neighbour tiles to reach the
 * same block, and reloads only its halo cells.  The waits are bounded (2 s);
 * one that times out makes fdtd_sync and the reads fail with FDTD_E_STATE.
 * Otherwise, and for the steps left over, launches run one block each.
 * Results are bitwise those of one-block launches of the same (K, S).  No
 * counterpart in the reference; it replaces the per-S-steps relaunch of this
 * library's own loop (the `for counts in range(P.timeSteps)` bodies of
 * Solver_Engine.py:167,236,294). */
int fdtd_set_persistent(fdtd_ctx *ctx, int on);
/* *active = 1 when the next fdtd_run uses persistent launches. */
int fdtd_persistent_active(fdtd_ctx *ctx, int *active);
/* Workgroup halo exchange (h > 0; 0 turns it off, the default): the four
 * waves of a field workgroup take adjacent windows of one simulation that
 * overlap by 2h cells and, every h steps, replace each other's inner halos
 * with the owners' values through LDS, so only the workgroup's two outer
 * windows carry S-cell halos (a window otherwise recomputes S cells on each
 * side, 2S of its 64K).  No counterpart in the reference, whose loop updates
 * each cell once per step (Solver_Engine.py:297-316).  Used when the tiling
 * is uniform with K = 2 or 4, h is a multiple of K with h/K <= 8, h <= S and
 * 64K - h - S >= 1, no beam is attached and persistent launches are off;
 * otherwise declined (fdtd_halo_exchange_active reports 0).  Results equal
 * the oracle to the same tolerances as the default tiles; they differ from
 * other tilings at the rounding level, as tilings do among themselves. */
int fdtd_set_halo_exchange(fdtd_ctx *ctx, int h);
int fdtd_halo_exchange_active(fdtd_ctx *ctx, int *active);

/* ---- simulations ------------------------------------------------------ */
/* Uploads coefficients, allocates zeroed state (FieldInit + CPML_FieldInit,
 * BaseFDTD11.py:36-82,195-219) and makes the simulation active from the next
 * fdtd_run.  The descriptor's host arrays may be freed on return. */
int fdtd_add_sim(fdtd_ctx *ctx, const fdtd_sim_desc_t *desc, int *sim_id);
/* Frees a simulation's device memory (it stops being advanced). */
int fdtd_remove_sim(fdtd_ctx *ctx, int sim_id);
/* Zeroes a simulation's state (fields, polarisation, CPML psi, probes) and
 * restarts it at step 0 from the next fdtd_run (a fresh FieldInit without
 * re-uploading coefficients).  Asynchronous on the context stream. */
int fdtd_reset_sim(fdtd_ctx *ctx, int sim_id);
/* Steps a simulation has completed. */
int fdtd_sim_progress(fdtd_ctx *ctx, int sim_id, int64_t *steps_done);

/* ---- checkpoint / resume ---------------------------------------------- */
/* The reference holds a run's state in memory only (Ex_History,
 * Solver_Engine.py:57-68, allocated at MasterController.py:160-166) and has
 * no checkpoint.  A checkpoint is one simulation's device state (both
 * ping-pong buffers of fields, ADE and CPML state, probes and snapshot rows
 * so far, the attached beam's particles and per-cell results, the energy
 * rows) plus its step position, written into a caller buffer of
 * fdtd_state_size bytes.  fdtd_state_save drains the context stream first.
 * fdtd_state_load copies one into a simulation built from the same
 * description in a context of the same precision (with the same beam
 * attached and the same energy setting; FDTD_E_ARG otherwise): it continues
 * from the saved step, and in a context of the same tiling, with launches
 * of whole S-step slices, its results are bitwise those of the run that was
 * saved.  Coefficients are not part of a checkpoint. */
int fdtd_state_size(fdtd_ctx *ctx, int sim_id, int64_t *bytes);
int fdtd_state_save(fdtd_ctx *ctx, int sim_id, void *buf, int64_t len);
int fdtd_state_load(fdtd_ctx *ctx, int sim_id, const void *buf, int64_t len);

/* ---- stepping --------------------------------------------------------- */
/* Advances every active, unfinished simulation by up to nsteps steps
 * (the `for counts in range(P.timeSteps)` bodies of Solver_Engine.py:167,236,294).
 * Asynchronous on the context stream. */
int fdtd_run(fdtd_ctx *ctx, int64_t nsteps);
/* Advances until every active simulation has run all its steps. */
int fdtd_run_all(fdtd_ctx *ctx);
int fdtd_sync(fdtd_ctx *ctx);
/* *busy = 1 while launches queued by fdtd_run are still running, else 0.
 * Lets a batching host admit work (fdtd_add_sim, pic_attach upload on a
 * stream of their own and do not wait for running launches) while the
 * device steps; no counterpart in the reference, whose loop is blocking. */
int fdtd_busy(fdtd_ctx *ctx, int *busy);

/* ---- particle beams (PIC) ---------------------------------------------- */
/* The reference has no particle code: its current sink V.Jx
 * (MasterController.py:159) enters the grid in ADE_ExUpdate
 * (BaseFDTD11.py:663-669) as Ex += (dHy - Jx)*UpExMat, in units of J_x*dz.
 * A beam is advanced every step by a relativistic Boris push with linear
 * gather of E and the time-averaged H, and deposits J' = Z0*Jx by CIC at the
 * half-step position; a context holding a beam runs one step per launch. */
typedef struct {
    int64_t n;          /* macro-particles */
    double q, m;        /* charge (C) and mass (kg) of one physical particle */
    double w;           /* macro-particle weight per unit transverse area (m^-2) */
    double dt, dz;      /* time step (s) and cell size (m) of the simulation */
    const double *x;    /* positions in cell units (E node k at x = k), in [2, N-2) */
    const double *u;    /* gamma*v/c, n x 3 (ux, uy, uz) */
    int32_t group;      /* cells per push/count group, 1..64 (0: picked from the density) */
} fdtd_beam_desc_t;

/* Attaches a beam to a simulation created with has_current = 1 that has not
 * run a step yet.  Particles are stored sorted by (cell, position, u) so the
 * run does not depend on the input order. */
int pic_attach(fdtd_ctx *ctx, int sim_id, const fdtd_beam_desc_t *beam);
/* Particle state in input order: x (n), u (n x 3), alive (n; may be NULL).
 * Absorbed particles report alive = 0 and a position in cell 1 or N-2. */
int pic_read(fdtd_ctx *ctx, int sim_id, double *x, double *u, int8_t *alive, int64_t n);

/* ---- data movement (host arrays FP64) --------------------------------- */
int fdtd_read_field(fdtd_ctx *ctx, int sim_id, int field, double *out, int64_t len);
int fdtd_write_field(fdtd_ctx *ctx, int sim_id, int field, const double *in, int64_t len);
/* probes as nprobe x steps, row k holds probe k (x1ColBe/x1ColAf/Port1/Port2) */
int fdtd_read_probes(fdtd_ctx *ctx, int sim_id, double *out, int64_t len);
/* snapshots as snap_rows x (N+1) (Ex_History, MasterController.py:160) */
int fdtd_read_snapshots(fdtd_ctx *ctx, int sim_id, double *out, int64_t len);

/* ---- energy ----------------------------------------------------------- */
/* FP64 discrete field energy of one simulation, taken on the device after
 * every `every`-th launch in which it steps (every = 0: off), up to `rows`
 * values:  W = sum_n E_n^2 + sum_n H'_n H'prev_n  (E and H' = Z0 Hy after the
 * launch's last step, H'prev before that step's H update; the leapfrog
 * invariant of a closed lossless cavity, conserved to rounding).  FP32
 * fields enter through their Kahan compensations; every term and the sum
 * are FP64, in a fixed order.  The reference has no accumulator: its
 * per-step observables are the probes of Solver_Engine.py:360-370.  A
 * simulation with energy on runs the general tile class and never a
 * persistent launch.  Setting it again clears the rows. */
int fdtd_set_energy(fdtd_ctx *ctx, int sim_id, int every, int rows);
/* Copies up to cap recorded values and the step index after which each was
 * taken; *count = rows recorded so far. */
int fdtd_read_energy(fdtd_ctx *ctx, int sim_id, double *values, int64_t *steps, int64_t cap, int64_t *count);

/* ---- spectra ---------------------------------------------------------- */
/* FP64 DFT X(f) = sum_t x_t exp(-2 pi i f t dt) dt of probe `probe` of each
 * listed simulation, computed on the device from the resident probe series
 * (the per-frequency sums behind TransformHandler.RefTester,
 * TransformHandler.py:34-78, and the two-pass reflection protocol of
 * Solver_Engine.py:275-371).  out: nsims x nf x (re, im).  Synchronous. */
int fdtd_probe_spectra(fdtd_ctx *ctx, int nsims, const int *sim_ids, int probe, const double *freqs, int nf,
                       double dt, double *out);

/* The same spectra with no allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc), this
 * variant stages its tables and results in a static device array of the
 * library (4 MiB + 56 KiB per GPU, held until the process ends) instead of
 * the context's grow-on-demand buffer, and calls no hipMalloc / hipFree.
 * Limits: nsims <= 4096, nf <= 1024, nsims * nf <= 262144 (FDTD_E_ARG
 * beyond them).  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two threads, two contexts) race; the call is synchronous, so
 * consecutive calls are safe.  Results equal fdtd_probe_spectra's bit for
 * bit. */
int fdtd_probe_spectra_noalloc(fdtd_ctx *ctx, int nsims, const int *sim_ids, int probe, const double *freqs,
                               int nf, double dt, double *out);

/* ---- single reference update kernels ---------------------------------- */
/* One BaseFDTD11 per-step function on the device, FP64, in the reference's
 * expression order (no FMA contraction), for callers that drive the
 * per-step functions one by one (Solver_Engine.py:297-316 style).  arrays
 * are host buffers (copied in, outputs copied back); ranges holds two
 * half-open index ranges [lo0,hi0) [lo1,hi1) (empty when hi <= lo).
 *   op  function (BaseFDTD11.py)        arrays                                scalars
 *   0   ADE_ExUpdate :663-669           Ex*, Hy, Jx, UpExMat, den_Exdz        -
 *   1   ADE_HyUpdate :640-656           Hy*, Ex, UpHySelf, UpHyMat, den_Hydz  -
 *   2   CPML_Psi_e_Update :364-376      Ex*, psi_Ex*, Hy, beX, ceX, Cb        -
 *   3   CPML_Psi_m_Update :381-393      Hy*, psi_Hy*, Ex, bmY, cmY, C2        -
 *   4   ADE_DxUpdate :750-760           Dx*, Hy, den_Exdz                     dt/dz
 *   5   ADE_ExCreate :712-725           Ex*, Dx, P                            eps0
 *   6   ADE_PolarisationCurrent_Ex :609-633  P*, P_prev, Ex                  A, B, C
 *   7   AcubicFinder :803-853           Acubic*, Dx                           cub, qua, one, eps0
 *   8   NonLinExUpdate :858-877         Ex*, Dx, Acubic                       eps0, chi3, sqrt(1.2)
 * and the per-step functions of the older BaseFDTD.py that differ in form:
 *   op  function (BaseFDTD.py)          arrays                                scalars
 *   9   CPML_ExUpdate :355-358,         Ex*, Hy, UpExSelf, c1, c2             -
 *       ADE_ExUpdate :401-408           (Ex = Ex*Self + dH*c1*c2)
 *   10  CPML_Psi_e_Update :331-336      Ex*, psi_Ex*, Hy, beX, ceX, Cb        -   (Ex += Cb psi)
 *   11  CPML_Psi_m_Update :339-344      Hy*, psi_Hy*, Ex, bmY, cmY, C2        -   (Ex[n] - Ex[n+1])
 *   12  ADE_PolarisationCurrent_Ex :378-390  P*, Pprev, Pprev2, Ex            A, B, C
 *   13  ADE_DxUpdate :417-422           Dx*, Ex*, Hy, P                       dt, dz, eps0
 * (* = written).  Needs an FP64 context. */
int fdtd_step_op(fdtd_ctx *ctx, int op, int narr, double *const *arrays, const int64_t *lens, int nscal,
                 const double *scalars, const int32_t *ranges);

/* ---- diagnostics ------------------------------------------------------ */
const char *fdtd_last_error(void);
/* Device time (ms) of the sampled field launches since the last reset (see
 * fdtd_set_profile for how they are sampled); also how many launches the
 * samples hold. */
int fdtd_timing(fdtd_ctx *ctx, double *ms, int64_t *launches, int reset);
/* Device time (ms) and launch count of one kernel kind (FDTD_KERNEL_*) since
 * the last reset, under profiling. */
int fdtd_kernel_timing(fdtd_ctx *ctx, int kind, double *ms, int64_t *launches);
/* Profiling (every = 0: off).  Without a beam, runs of `every` consecutive
 * field launches are bracketed by two stream markers (HIP events recorded on
 * the context's stream), so fdtd_timing reports the summed duration of the
 * runs and the number of launches in them: their ratio is the average launch
 * duration used for the roofline, gaps between launches included, and no
 * launch carries an event of its own.  With a beam, every `every`-th step's
 * kernels (1: all) carry start/stop events stamped by their own dispatch
 * packets, and fdtd_kernel_timing reports them per kernel kind.  An event
 * pair costs more host time than a launch, so sampling keeps a long run
 * GPU-bound. */
int fdtd_set_profile(fdtd_ctx *ctx, int every);
/* Bytes of device memory the context currently holds. */
int fdtd_device_bytes(fdtd_ctx *ctx, int64_t *bytes);
/* Library version string. */
const char *fdtd_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FDTDPIC_H */
