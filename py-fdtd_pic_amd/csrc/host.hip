// C-ABI of libfdtdpic.so: contexts, per-simulation device state, tile lists
// and the launch loop around the stencil kernel.  See include/fdtdpic.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fdtdpic.h"
#include "device_common.hpp"
#include "host_internal.hpp"

namespace fdtd {
template <typename R, bool KAHAN>
hipError_t launch_stencil(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                          long long gstep0, int nsteps, int K, int S, int kind, const PersistArgs& pa,
                          hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
template <typename R, bool KAHAN>
int field_capacity_per_cu(int K, int kind);
struct OpArgs {
    double* a[8];
    double s[4];
    int lo[2], hi[2];
};
hipError_t launch_step_op(int op, const OpArgs& g, hipStream_t st);
hipError_t launch_dft(const double* const* series, const int* lens, int nseries, const double* freqs, int nf, double dt,
                      double* out, hipStream_t st);
template <typename R>
hipError_t launch_energy(const EnergyJob<R>* jobs, int n, hipStream_t st);
template <typename R>
hipError_t launch_push(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* cell_items, int ncell,
                       const int2* group_items, int ngroup, long long launch_idx, int xr, int single, hipStream_t st,
                       hipEvent_t e0, hipEvent_t e1);
template <typename R>
hipError_t launch_scatter(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* scatter_blocks,
                          int nscatter, long long launch_idx, hipStream_t st, hipEvent_t e0, hipEvent_t e1);
template <typename R, bool KAHAN>
hipError_t launch_field_scatter(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                                long long gstep0, int nsteps, int K, int S, int kind, const BeamDev<R>* beams,
                                const int2* sblocks, int nsblocks, int nfirst, hipStream_t st, hipEvent_t ev0,
                                hipEvent_t ev1);
}

using namespace fdtd;

static thread_local std::string g_err;

int fdtd_fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------------------
// simulation block layout
// ------------------------------------------------------------------------
struct Layout {
    size_t off[64];
    int n = 0;
    size_t total = 0;
    size_t add(size_t bytes) {
        size_t o = total;
        off[n++] = o;
        total = align_up(total + bytes, 256);
        return o;
    }
};

template <typename R>
static int upload_beam(BeamHost& b, int N, hipStream_t st);

template <typename R>
static void fill_dev(SimHost& s, SimDev<R>& d) {
    const fdtd_sim_desc_t& q = s.d;
    char* blk = (char*)s.block;
    const int N = q.N, ns = s.ns, nstrip = s.nstrip, T = q.steps;
    std::memset(&d, 0, sizeof(d));
    d.N = N; d.T = T; d.Ls = q.Ls; d.Rs = q.Rs; d.F = q.F; d.Rr = q.R;
    d.slab_kind = q.slab_kind; d.ade = q.ade; d.src_cell = q.src_cell; d.tfsf = q.tfsf;
    d.nprobe = q.nprobe; d.snap_interval = q.snap_interval; d.snap_rows = q.snap_rows;
    d.has_current = q.has_current;
    d.launch_begin = s.launch_begin; d.t_begin = s.t_begin;
    for (int p = 0; p < FDTD_MAX_PROBES; ++p) {
        d.probe_cell[p] = q.probe_cell[p]; d.probe_t0[p] = q.probe_t0[p]; d.probe_t1[p] = q.probe_t1[p];
    }
    d.ke0 = (R)q.ke0; d.kh0 = (R)q.kh0; d.keS = (R)q.keS; d.khS = (R)q.khS; d.kD = (R)q.kD; d.jD = (R)q.jD;
    d.kerr_a = q.kerr_a; d.kerr_b = q.kerr_b; d.kerr_c = q.kerr_c;
    d.kerr_e0 = q.kerr_e0; d.kerr_lin = q.kerr_lin; d.kerr_chi3 = q.kerr_chi3;
    auto P = [&](int k) { return (R*)(blk + s.lay_off[k]); };
    int k = 0;
    const R* cs[8];
    for (int i = 0; i < 8; ++i) cs[i] = P(k++);
    d.s_ke = cs[0]; d.s_kh = cs[1]; d.s_bE = cs[2]; d.s_cE = cs[3];
    d.s_CbE = cs[4]; d.s_bH = cs[5]; d.s_cH = cs[6]; d.s_C2H = cs[7];
    d.sl_A = P(k++); d.sl_B = P(k++); d.sl_C = P(k++); d.sl_e = P(k++);
    d.srcE = P(k++); d.srcH = P(k++);
    for (int b = 0; b < 2; ++b) {
        d.E[b] = P(k++); d.H[b] = P(k++); d.cE[b] = P(k++); d.cH[b] = P(k++);
        d.D[b] = P(k++); d.P[b] = P(k++); d.Pm[b] = P(k++); d.cD[b] = P(k++); d.cP[b] = P(k++);
        d.psiE[b] = P(k++); d.psiH[b] = P(k++);
    }
    d.probe_out = (double*)P(k++);
    d.snap = P(k++);
    d.J = q.has_current ? P(k++) : nullptr;
    d.Hprev = s.e_every > 0 ? (double*)s.eblock : nullptr;
    (void)N; (void)ns; (void)nstrip; (void)T;
}

template <typename R>
static int build_sim(fdtd_ctx* ctx, SimHost& s, const fdtd_sim_desc_t* q) {
    const int N = q->N, T = q->steps;
    const int ns = std::max(q->R - q->F, 0);
    const int nstrip = q->Ls + (N + 1 - q->Rs);
    s.ns = ns; s.nstrip = nstrip;
    Layout L;
    const size_t rs = sizeof(R);
    for (int i = 0; i < 8; ++i) L.add(rs * std::max(nstrip, 1));      // strip coefficients
    for (int i = 0; i < 4; ++i) L.add(rs * std::max(ns, 1));          // slab coefficients
    L.add(rs * T); L.add(rs * T);                                      // srcE, srcH
    const size_t coef_end = L.total;
    for (int b = 0; b < 2; ++b) {
        for (int i = 0; i < 4; ++i) L.add(rs * (N + 1));              // E H cE cH
        for (int i = 0; i < 5; ++i) L.add(rs * std::max(ns, 1));      // D P Pm cD cP
        for (int i = 0; i < 2; ++i) L.add(rs * std::max(nstrip, 1));  // psiE psiH
    }
    L.add(sizeof(double) * (size_t)std::max(q->nprobe, 1) * T);       // probes (FP64 in both precisions)
    L.add(rs * (size_t)std::max(q->snap_rows, 1) * (N + 1));           // snapshots
    if (q->has_current) L.add(rs * (N + 1));                           // J
    for (int i = 0; i < L.n; ++i) s.lay_off[i] = L.off[i];
    s.bytes = L.total;
    if (hipMalloc(&s.block, s.bytes) != hipSuccess) {
        s.block = nullptr;
        return fdtd_fail(FDTD_E_NOMEM, "hipMalloc of " + std::to_string(s.bytes) + " bytes failed");
    }
    // coefficient image (host) -> one copy; state region zeroed.
    std::vector<R> img(coef_end / rs, R(0));
    auto put = [&](int k, const double* src, int n) {
        if (!src) return;
        R* dst = img.data() + L.off[k] / rs;
        for (int i = 0; i < n; ++i) dst[i] = (R)src[i];
    };
    const double* sc[8] = {q->s_ke, q->s_kh, q->s_bE, q->s_cE, q->s_CbE, q->s_bH, q->s_cH, q->s_C2H};
    for (int i = 0; i < 8; ++i) put(i, sc[i], nstrip);
    if (ns > 0 && q->slab_kind >= FDTD_SLAB_LINEAR) {
        R* A = img.data() + L.off[8] / rs;
        R* B = img.data() + L.off[9] / rs;
        R* C = img.data() + L.off[10] / rs;
        for (int i = 0; i < ns; ++i) {
            const double a = q->ade ? q->sl_A[i] : 0.0, b = q->ade ? q->sl_B[i] : 0.0, c = q->ade ? q->sl_C[i] : 0.0;
            if (sizeof(R) == 4) {   // velocity form coefficients: -B, 1-A-B, C
                A[i] = (R)(-b); B[i] = (R)(1.0 - a - b); C[i] = (R)c;
            } else {
                A[i] = (R)a; B[i] = (R)b; C[i] = (R)c;
            }
        }
        // the kernels multiply by 1/e (LINEAR: 1/eps0, KERR_E: 1/eps_inf)
        R* ie = img.data() + L.off[11] / rs;
        for (int i = 0; i < ns; ++i) ie[i] = (R)(1.0 / q->sl_e[i]);
    }
    put(12, q->srcE, T);
    put(13, q->srcH, T);
    // a new block is not referenced by queued launches: upload on the upload
    // stream, so adding a simulation does not wait for running ones
    HIPCHK(hipMemsetAsync(s.block, 0, s.bytes, ctx->up));
    HIPCHK(hipMemcpyAsync(s.block, img.data(), coef_end, hipMemcpyHostToDevice, ctx->up));
    HIPCHK(hipStreamSynchronize(ctx->up));
    return FDTD_OK;
}

static int validate(const fdtd_sim_desc_t* q, int prec) {
    if (!q) return fdtd_fail(FDTD_E_ARG, "null descriptor");
    if (q->N < 2 || q->steps < 1) return fdtd_fail(FDTD_E_ARG, "N must be >= 2 and steps >= 1");
    if (q->Ls < 0 || q->Rs > q->N + 1 || q->Ls > q->Rs)
        return fdtd_fail(FDTD_E_ARG, "strips must satisfy 0 <= Ls <= Rs <= N+1");
    if (q->slab_kind < 0 || q->slab_kind > FDTD_SLAB_KERR_E) return fdtd_fail(FDTD_E_ARG, "bad slab_kind");
    if (q->slab_kind != FDTD_SLAB_NONE && (q->F < 1 || q->R > q->N + 1 || q->F >= q->R))
        return fdtd_fail(FDTD_E_ARG, "slab must satisfy 1 <= F < R <= N+1");
    if (q->slab_kind >= FDTD_SLAB_LINEAR && (!q->sl_e || (q->ade && (!q->sl_A || !q->sl_B || !q->sl_C))))
        return fdtd_fail(FDTD_E_ARG, "slab coefficient arrays missing");
    if (q->slab_kind == FDTD_SLAB_KERR_REF && prec != FDTD_F64)
        return fdtd_fail(FDTD_E_ARG, "the reference Kerr path runs in FP64 only");
    if (q->slab_kind == FDTD_SLAB_KERR_E && q->kerr_chi3 < 0)
        return fdtd_fail(FDTD_E_ARG, "KERR_E needs chi3 >= 0");
    if (q->src_cell < 1 || q->src_cell > q->N) return fdtd_fail(FDTD_E_ARG, "source cell out of range");
    if (!q->srcE) return fdtd_fail(FDTD_E_ARG, "srcE missing");
    if (q->tfsf && !q->srcH) return fdtd_fail(FDTD_E_ARG, "srcH missing with tfsf");
    if (q->slab_kind >= FDTD_SLAB_LINEAR && q->src_cell >= q->F - (q->tfsf ? 1 : 0) && q->src_cell < q->R + (q->tfsf ? 1 : 0))
        return fdtd_fail(FDTD_E_ARG, "source inside (or, with TFSF, adjacent to) the slab");
    const int nstrip = q->Ls + (q->N + 1 - q->Rs);
    if (nstrip > 0 && (!q->s_ke || !q->s_kh || !q->s_bE || !q->s_cE || !q->s_CbE || !q->s_bH || !q->s_cH || !q->s_C2H))
        return fdtd_fail(FDTD_E_ARG, "strip coefficient arrays missing");
    if (q->nprobe < 0 || q->nprobe > FDTD_MAX_PROBES) return fdtd_fail(FDTD_E_ARG, "nprobe out of range");
    for (int p = 0; p < q->nprobe; ++p)
        if (q->probe_cell[p] < 0 || q->probe_cell[p] > q->N) return fdtd_fail(FDTD_E_ARG, "probe cell out of range");
    if (q->snap_interval < 0 || q->snap_rows < 0) return fdtd_fail(FDTD_E_ARG, "bad snapshot settings");
    return FDTD_OK;
}

// ------------------------------------------------------------------------
// One upload stream per device, shared by every context on it (created on
// first use, kept until the process ends).  A batch of two contexts then runs
// three streams (two compute, one upload), which HIP's default of four
// hardware queues per process gives a queue each, so an upload never waits
// behind another stream's running launches on a shared queue.
// ------------------------------------------------------------------------
constexpr int kMaxDevices = 64;
static std::mutex g_dev_mu;
static hipStream_t g_upload[kMaxDevices] = {nullptr};

static int shared_upload_stream(int dev, hipStream_t* out) {
    if (dev < 0 || dev >= kMaxDevices) return fdtd_fail(FDTD_E_ARG, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (!g_upload[dev]) {
        hipStream_t st = nullptr;
        const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        g_upload[dev] = st;
    }
    *out = g_upload[dev];
    return FDTD_OK;
}

// ------------------------------------------------------------------------
// Scratch of fdtd_probe_spectra_noalloc: a file-scope device array, one copy
// per GPU, that holds the pointer table, the series lengths, the frequencies
// and the staged results of one call.  The host never passes the array's
// name to a kernel (in host code it is a stand-in, not a device address):
// fdtd_ctx_create resolves each device's address once with
// hipGetSymbolAddress, with that device current, and the entry point uses it.
// ------------------------------------------------------------------------
constexpr int kNaMaxSims = 4096;                 // series per call
constexpr int kNaMaxFreqs = 1024;                // frequencies per call
constexpr int64_t kNaMaxPairs = 262144;          // series x frequencies per call (4 MiB of results)
constexpr size_t kNaOffLens = sizeof(void*) * kNaMaxSims;
constexpr size_t kNaOffFreqs = kNaOffLens + sizeof(int) * kNaMaxSims;
constexpr size_t kNaOffOut = kNaOffFreqs + sizeof(double) * kNaMaxFreqs;
constexpr size_t kNaScratchBytes = kNaOffOut + 2 * sizeof(double) * kNaMaxPairs;
static_assert(kNaOffLens % 256 == 0 && kNaOffFreqs % 256 == 0 && kNaOffOut % 256 == 0, "aligned sub-arrays");

static __device__ __align__(256) char scratch[kNaScratchBytes];

static void* g_na_ptr[kMaxDevices] = {nullptr};

// Resolves the device address of `scratch` on device dev (current), once.
static int noalloc_scratch_init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return fdtd_fail(FDTD_E_ARG, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_na_ptr[dev]) return FDTD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fdtd_fail(FDTD_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_na_ptr[dev] = p;
    return FDTD_OK;
}

static void* noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_dev_mu);
    return g_na_ptr[dev];
}

// ------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------
extern "C" {

const char* fdtd_last_error(void) { return g_err.c_str(); }
const char* fdtd_version(void) { return "fdtdpic-gfx950 0.1"; }

int fdtd_ctx_create(int precision, fdtd_ctx** out) {
    if (!out) return fdtd_fail(FDTD_E_ARG, "null out");
    if (precision != FDTD_F64 && precision != FDTD_F32) return fdtd_fail(FDTD_E_ARG, "bad precision");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fdtd_fail(FDTD_E_HIP, "no HIP device visible");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fdtd_fail(FDTD_E_HIP, "hipGetDevice failed");
    int rc0 = noalloc_scratch_init(dev);
    if (rc0) return rc0;
    fdtd_ctx* c = new fdtd_ctx();
    c->prec = precision;
    c->device = dev;
    c->K = c->K_slab = precision == FDTD_F64 ? 2 : 4;
    c->S = default_S(c->K);
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fdtd_fail(FDTD_E_HIP, "hipStreamCreate failed");
    }
    if (int rc = shared_upload_stream(dev, &c->up)) {
        (void)hipStreamDestroy(c->st);
        delete c;
        return rc;
    }
    if (hipMalloc(&c->d_err, sizeof(int32_t)) != hipSuccess || hipMemset(c->d_err, 0, sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->d_perr, sizeof(int32_t)) != hipSuccess || hipMemset(c->d_perr, 0, sizeof(int32_t)) != hipSuccess) {
        if (c->d_err) (void)hipFree(c->d_err);
        if (c->d_perr) (void)hipFree(c->d_perr);
        (void)hipStreamDestroy(c->st);
        delete c;
        return fdtd_fail(FDTD_E_HIP, "allocation of the error flags failed");
    }
    if (hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        c->cu_count = 0;
    if (hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        c->lds_max = 0;
    *out = c;
    return FDTD_OK;
}

int fdtd_ctx_destroy(fdtd_ctx* c) {
    if (!c) return FDTD_OK;
    // best effort: teardown errors are not actionable here
    (void)hipStreamSynchronize(c->st);
    (void)hipStreamSynchronize(c->up);
    for (SimHost* s : c->sims)
        if (s) {
            if (s->block) (void)hipFree(s->block);
            if (s->eblock) (void)hipFree(s->eblock);
            delete s;
        }
    for (BeamHost* b : c->beams)
        if (b) { if (b->block) (void)hipFree(b->block); delete b; }
    for (void* p : {(void*)c->d_beams, (void*)c->d_pblocks, (void*)c->d_pitems, (void*)c->d_err, c->d_scratch,
                    (void*)c->d_perr, (void*)c->d_flags, (void*)c->d_nbr})
        if (p) (void)hipFree(p);
    if (c->d_sims) (void)hipFree(c->d_sims);
    if (c->d_tiles) (void)hipFree(c->d_tiles);
    for (auto& e : c->events) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (hipEvent_t e : c->throttle) (void)hipEventDestroy(e);
    if (c->stage) (void)hipHostFree(c->stage);
    if (c->stage_ev) (void)hipEventDestroy(c->stage_ev);
    (void)hipStreamDestroy(c->st);        // c->up is the device's shared upload stream
    delete c;
    return FDTD_OK;
}

int fdtd_set_tiling_mixed(fdtd_ctx* c, int K, int K_slab, int S) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (K != 1 && K != 2 && K != 4 && K != 8) return fdtd_fail(FDTD_E_ARG, "K must be 1, 2, 4 or 8");
    if (K_slab != K && !(K >= 2 && K_slab == K / 2)) return fdtd_fail(FDTD_E_ARG, "K_slab must be K or K/2");
    if (S == 0) S = default_S(K_slab);
    // S <= 32 K_slab - 1 for the default tiles, checked when the tile list is
    // built: with the workgroup halo exchange S may reach 64 K - h - 1
    if (S < 1 || S > 255 || 64 * K_slab - 2 - S < 1)
        return fdtd_fail(FDTD_E_ARG, "S must satisfy 1 <= S <= min(255, 64*K_slab-3)");
    if ((int64_t)c->sims.size() >= (1 << 24)) return fdtd_fail(FDTD_E_ARG, "too many simulations");
    c->K = K; c->K_slab = K_slab; c->S = S; c->dirty = true;
    return FDTD_OK;
}

int fdtd_set_tiling(fdtd_ctx* c, int K, int S) { return fdtd_set_tiling_mixed(c, K, K, S); }

int fdtd_add_sim(fdtd_ctx* c, const fdtd_sim_desc_t* q, int* sim_id) {
    if (!c || !sim_id) return fdtd_fail(FDTD_E_ARG, "null argument");
    int rc = validate(q, c->prec);
    if (rc) return rc;
    SimHost* s = new SimHost();
    s->d = *q;
    s->launch_begin = c->launch_idx;
    s->t_begin = c->gstep;
    rc = c->prec == FDTD_F64 ? build_sim<double>(c, *s, q) : build_sim<float>(c, *s, q);
    if (rc) { delete s; return rc; }
    // host pointers are not kept
    s->d.s_ke = s->d.s_kh = s->d.s_bE = s->d.s_cE = s->d.s_CbE = s->d.s_bH = s->d.s_cH = s->d.s_C2H = nullptr;
    s->d.sl_A = s->d.sl_B = s->d.sl_C = s->d.sl_e = s->d.srcE = s->d.srcH = nullptr;
    c->dev_bytes += s->bytes;
    c->sims.push_back(s);
    *sim_id = (int)c->sims.size() - 1;
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_remove_sim(fdtd_ctx* c, int id) {
    if (!c || id < 0 || id >= (int)c->sims.size() || !c->sims[id]) return fdtd_fail(FDTD_E_ARG, "bad sim id");
    HIPCHK(hipStreamSynchronize(c->st));
    SimHost* s = c->sims[id];
    if (s->block) (void)hipFree(s->block);
    if (s->eblock) (void)hipFree(s->eblock);
    c->dev_bytes -= s->bytes + s->ebytes;
    delete s;
    c->sims[id] = nullptr;
    if (id < (int)c->beams.size() && c->beams[id]) {
        BeamHost* b = c->beams[id];
        if (b->block) (void)hipFree(b->block);
        c->dev_bytes -= b->bytes;
        delete b;
        c->beams[id] = nullptr;
    }
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_reset_sim(fdtd_ctx* c, int id) {
    if (!c || id < 0 || id >= (int)c->sims.size() || !c->sims[id]) return fdtd_fail(FDTD_E_ARG, "bad sim id");
    SimHost* s = c->sims[id];
    const size_t state0 = s->lay_off[14];   // first state sub-array (after coefficients and sources)
    HIPCHK(hipMemsetAsync((char*)s->block + state0, 0, s->bytes - state0, c->st));
    if (id < (int)c->beams.size() && c->beams[id]) {
        // ordered after the simulation's queued launches
        int rc = c->prec == FDTD_F64 ? upload_beam<double>(*c->beams[id], s->d.N, c->st)
                                     : upload_beam<float>(*c->beams[id], s->d.N, c->st);
        if (rc) return rc;
    }
    s->steps_done = 0;
    s->launches = 0;
    s->launch_begin = c->launch_idx;
    s->t_begin = c->gstep;
    s->e_count = 0;
    s->e_launches = 0;
    s->e_steps.clear();
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_sim_progress(fdtd_ctx* c, int id, int64_t* done) {
    if (!c || !done || id < 0 || id >= (int)c->sims.size() || !c->sims[id]) return fdtd_fail(FDTD_E_ARG, "bad sim id");
    *done = c->sims[id]->steps_done;
    return FDTD_OK;
}

}  // extern "C"

// Grow-only device buffer, then an async upload on the context stream.
// Device arrays rebuilt from host vectors (descriptors, tile and work lists).
struct Upload {
    void** dst;
    size_t* cap;         // bytes allocated at *dst
    const void* src;
    size_t bytes;
};

// Copies the sources into the context's pinned staging buffer and queues
// the uploads on the compute stream, after the launches already queued
// (which keep reading the old contents until then).  The staging buffer is
// reused only after the previous upload has run; a device array that must
// grow is reallocated after the queued launches have finished with it.
static int staged_upload(fdtd_ctx* c, const std::vector<Upload>& ups) {
    auto pad = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t total = 0;
    bool grow = false;
    for (const Upload& u : ups) {
        total += pad(u.bytes);
        grow = grow || u.bytes > *u.cap;
    }
    if (c->stage_pending) {
        HIPCHK(hipEventSynchronize(c->stage_ev));
        c->stage_pending = false;
    }
    if (total > c->stage_cap) {
        if (c->stage) HIPCHK(hipHostFree(c->stage));
        c->stage = nullptr;
        c->stage_cap = 0;
        HIPCHK(hipHostMalloc(&c->stage, total * 2, hipHostMallocDefault));
        c->stage_cap = total * 2;
    }
    if (grow) HIPCHK(hipStreamSynchronize(c->st));
    size_t off = 0;
    for (const Upload& u : ups) {
        if (u.bytes > *u.cap) {
            if (*u.dst) HIPCHK(hipFree(*u.dst));
            *u.dst = nullptr;
            *u.cap = 0;
            HIPCHK(hipMalloc(u.dst, u.bytes * 2));
            *u.cap = u.bytes * 2;
        }
        if (u.bytes) {
            char* s = (char*)c->stage + off;
            std::memcpy(s, u.src, u.bytes);
            HIPCHK(hipMemcpyAsync(*u.dst, s, u.bytes, hipMemcpyHostToDevice, c->st));
        }
        off += pad(u.bytes);
    }
    if (!c->stage_ev) HIPCHK(hipEventCreateWithFlags(&c->stage_ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->stage_ev, c->st));
    c->stage_pending = true;
    return FDTD_OK;
}

// ------------------------------------------------------------------------
// particle beams
// ------------------------------------------------------------------------
enum { BO_CELL = 0, BO_FRAC = 2, BO_UX = 4, BO_UY = 6, BO_UZ = 8, BO_ID = 10, BO_START = 12, BO_RANGE = 14,
       BO_URANGE = 16, BO_CUP = 18, BO_CDOWN, BO_S, BO_PARK, BO_END };

template <typename R>
static BeamDev<R> fill_beam_dev(BeamHost& b, int32_t* err, int64_t launch_begin) {
    BeamDev<R> d;
    std::memset(&d, 0, sizeof(d));
    char* blk = (char*)b.block;
    d.launch_begin = launch_begin;
    d.sim = b.sim; d.N = b.N; d.np = (int32_t)b.np; d.G = b.G;
    d.nwaves = b.nwaves; d.nnode = b.nnode; d.nmove = b.nmove;
    d.a = (R)b.a; d.jscale = (R)b.jscale; d.Sc = (R)b.Sc;
    for (int k = 0; k < 2; ++k) {
        d.cell[k] = (int32_t*)(blk + b.off[BO_CELL + k]);
        d.frac[k] = (R*)(blk + b.off[BO_FRAC + k]);
        d.ux[k] = (R*)(blk + b.off[BO_UX + k]);
        d.uy[k] = (R*)(blk + b.off[BO_UY + k]);
        d.uz[k] = (R*)(blk + b.off[BO_UZ + k]);
        d.id[k] = (int32_t*)(blk + b.off[BO_ID + k]);
        d.start[k] = (int32_t*)(blk + b.off[BO_START + k]);
        d.range[k] = (int32_t*)(blk + b.off[BO_RANGE + k]);
        d.urange[k] = (int32_t*)(blk + b.off[BO_URANGE + k]);
    }
    d.park = (int32_t*)(blk + b.off[BO_PARK]);
    d.cup = (int32_t*)(blk + b.off[BO_CUP]);
    d.cdown = (int32_t*)(blk + b.off[BO_CDOWN]);
    d.S = (double*)(blk + b.off[BO_S]);
    d.err = err;
    return d;
}

// Upload the initial (sorted) particle state into buffer 0 and start[] into both.
template <typename R>
static int upload_beam(BeamHost& b, int N, hipStream_t st) {
    const int64_t n = b.np;
    char* blk = (char*)b.block;
    std::vector<R> fr(n), ux(n), uy(n), uz(n);
    for (int64_t i = 0; i < n; ++i) {
        fr[i] = (R)b.frac0[i];
        ux[i] = (R)b.u0[3 * i]; uy[i] = (R)b.u0[3 * i + 1]; uz[i] = (R)b.u0[3 * i + 2];
    }
    const size_t rn = sizeof(R) * n, in = sizeof(int32_t) * n, sn = sizeof(int32_t) * (N + 2);
    if (b.G != 1)   // one-cell beams keep no cell array
        HIPCHK(hipMemcpyAsync(blk + b.off[BO_CELL], b.cell0.data(), in, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(blk + b.off[BO_FRAC], fr.data(), rn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(blk + b.off[BO_UX], ux.data(), rn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(blk + b.off[BO_UY], uy.data(), rn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(blk + b.off[BO_UZ], uz.data(), rn, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(blk + b.off[BO_ID], b.id0.data(), in, hipMemcpyHostToDevice, st));
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipMemcpyAsync(blk + b.off[BO_START + k], b.start0.data(), sn, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(blk + b.off[BO_RANGE + k], b.range0, sizeof(b.range0), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(blk + b.off[BO_URANGE + k], b.range0, sizeof(b.range0), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));   // the host vectors above go out of scope
    return FDTD_OK;
}

template <typename R>
static int attach_t(fdtd_ctx* c, int id, SimHost& s, const fdtd_beam_desc_t* bd) {
    const int N = s.d.N;
    const int64_t n = bd->n;
    BeamHost* b = new BeamHost();
    b->sim = id;
    b->N = N;
    b->np = n;
    b->q = bd->q; b->m = bd->m; b->w = bd->w;
    // normalised coefficients (E in V/m, H' = Z0 H, u = gamma v / c)
    const double c0 = 299792458.0, z0 = 376.730313668;
    b->Sc = c0 * bd->dt / bd->dz;
    b->a = bd->q * bd->dt / (2.0 * bd->m * c0);
    b->jscale = z0 * bd->q * bd->w * c0;
    // sort by (cell, frac, ux, uy, uz, input index), a total order of the
    // input: a counting sort by cell, then each cell's run by the rest
    // (contiguous (frac, index) keys; u and the index only break frac ties)
    std::vector<int32_t> cl(n);
    std::vector<double> fr(n);
    for (int64_t i = 0; i < n; ++i) {
        const double x = bd->x[i];
        const double fl = std::floor(x);
        cl[i] = (int32_t)fl;
        fr[i] = x - fl;
    }
    const double* u = bd->u;
    std::vector<int64_t> first(N + 3, 0);
    for (int64_t i = 0; i < n; ++i) first[cl[i] + 1]++;
    for (int cc = 0; cc <= N + 1; ++cc) first[cc + 1] += first[cc];
    struct Key { double f; int64_t i; };
    std::vector<Key> keys(n);
    {
        std::vector<int64_t> pos(first.begin(), first.end());
        for (int64_t i = 0; i < n; ++i) keys[pos[cl[i]]++] = Key{fr[i], i};
    }
    for (int cc = 0; cc <= N + 1; ++cc) {
        if (first[cc + 1] - first[cc] < 2) continue;
        std::sort(keys.begin() + first[cc], keys.begin() + first[cc + 1], [&](const Key& p, const Key& q) {
            if (p.f != q.f) return p.f < q.f;
            for (int k = 0; k < 3; ++k)
                if (u[3 * p.i + k] != u[3 * q.i + k]) return u[3 * p.i + k] < u[3 * q.i + k];
            return p.i < q.i;
        });
    }
    std::vector<int64_t> ord(n);
    for (int64_t k = 0; k < n; ++k) ord[k] = keys[k].i;
    b->cell0.resize(n); b->id0.resize(n); b->frac0.resize(n); b->u0.resize(3 * n);
    b->start0.assign(N + 2, 0);
    std::vector<int64_t> occ(N + 2, 0);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = ord[k];
        b->cell0[k] = cl[i]; b->frac0[k] = fr[i]; b->id0[k] = (int32_t)i;
        for (int j = 0; j < 3; ++j) b->u0[3 * k + j] = u[3 * i + j];
        occ[cl[i]]++;
    }
    int64_t acc = 0, occupied = 0, fullest = 0;
    for (int cc = 0; cc <= N + 1; ++cc) {
        b->start0[cc] = (int32_t)acc;
        acc += occ[cc];
        occupied += occ[cc] > 0;
        fullest = std::max(fullest, occ[cc]);
    }
    // LDS rows per cell that hold the fullest cell's chunks beyond the
    // register batch, with a quarter to spare as the beam evolves
    b->xr_want = (int)std::max<int64_t>(1, (fullest + fullest / 4 + 63) / 64 - kBatch);
    // live range (cells of the particles; all live at attach)
    b->range0[0] = b->cell0.front();
    b->range0[1] = b->cell0.back();
    // cells per push/count group; a wave loops over groups of the live range
    // in a grid stride.  Dense beams (>= 96 particles per occupied cell) take
    // one-cell groups (the scalar-carry path); sparser ones groups of about
    // 1,024 particles, which amortise the per-group work of the multi-cell
    // path (measured on config 4's 2,048 beams: push/count 1.17 ms per step
    // at ~96 particles per group, 0.85 ms with this rule; config 3, 100 per
    // cell: 13.0 us at one cell per group, 16.9 us at four)
    int G = bd->group;
    if (G <= 0) {
        const double per_cell = occupied ? (double)n / occupied : 1.0;
        G = per_cell >= 96.0 ? 1 : (int)std::lround(1024.0 / per_cell);
    }
    b->G = std::min(std::max(G, 1), kPicGMax);
    const int span = b->range0[1] - b->range0[0] + 3;
    // one-cell groups run two cells per wave, in workgroups of kPushWaves
    // waves that take blocks of 2 * kPushWaves cells together
    if (b->G == 1) {
        const int64_t cells_per_wg = 2 * kPushWaves;
        b->nwaves = kPushWaves * (int)std::min<int64_t>(32768 / kPushWaves,
                                                        std::max<int64_t>(1, (span + cells_per_wg - 1) / cells_per_wg));
    } else {
        b->nwaves = (int)std::min<int64_t>(32768, std::max<int64_t>(1, (span + b->G - 1) / b->G));
    }
    b->nnode = std::min(64, std::max(1, (span + 8 + 255) / 256 + 1));
    // move_block: kXMove threads per push/count block boundary (blocks of
    // 2 * kPushWaves cells for one-cell beams, groups of G cells otherwise)
    const int bcells = b->G == 1 ? 2 * kPushWaves : b->G;
    const int nbound = (span + bcells - 1) / bcells + 2;
    b->nmove = std::min(256, std::max(1, (nbound * kXMove + 255) / 256));
    // device block
    Layout L;
    const size_t rn = sizeof(R) * n, in = sizeof(int32_t) * n, cn = sizeof(int32_t) * (N + 2);
    const bool one = b->G == 1;
    for (int k = 0; k < 2; ++k) L.add(one ? 0 : in);             // cell (multi-cell beams)
    for (int k = 0; k < 2; ++k) L.add(rn);                       // frac
    for (int k = 0; k < 6; ++k) L.add(rn);                       // ux0 ux1 uy0 uy1 uz0 uz1 (BO_UX/UY/UZ + k)
    for (int k = 0; k < 2; ++k) L.add(in);                       // id
    for (int k = 0; k < 2; ++k) L.add(cn);                       // start
    for (int k = 0; k < 4; ++k) L.add(2 * sizeof(int32_t));      // range x2, urange x2
    L.add(cn); L.add(cn);                                        // cup cdown
    L.add(4 * sizeof(double) * (N + 2));                         // S
    L.add(2 * sizeof(int32_t));                                  // park
    for (int k = 0; k < L.n; ++k) b->off[k] = L.off[k];
    b->bytes = L.total;
    if (hipMalloc(&b->block, b->bytes) != hipSuccess) {
        delete b;
        return fdtd_fail(FDTD_E_NOMEM, "hipMalloc of " + std::to_string(L.total) + " bytes (beam) failed");
    }
    if (c->beams.size() < c->sims.size()) c->beams.resize(c->sims.size(), nullptr);
    c->beams[id] = b;
    c->dev_bytes += b->bytes;
    int rc = upload_beam<R>(*b, N, c->up);     // a new beam: no queued launch reads it
    if (rc) return rc;
    c->dirty = true;
    return FDTD_OK;
}

// Rebuild the device SimDev array and the tile list for active, unfinished sims.
template <typename R>
static int rebuild_t(fdtd_ctx* c) {
    const int nsim = (int)c->sims.size();
    std::vector<SimDev<R>> devs(std::max(nsim, 1));
    std::vector<int2> by_kind[5];
    // a coupled beam forces one step per launch: tiles then carry a 1-cell halo
    bool pic = false;
    for (int i = 0; i < nsim; ++i)
        if (c->sims[i] && c->sims[i]->steps_done < c->sims[i]->d.steps && i < (int)c->beams.size() && c->beams[i])
            pic = true;
    c->S_run = pic ? 1 : c->S;
    const int U = 64 * c->K - 2 * c->S_run;              // plain tiles
    const int Un = 64 * (c->K / 2) - 2 * c->S_run;       // narrow (slab) tiles, when K_slab = K/2
    // workgroup halo exchange: four adjacent windows per workgroup, 64K - 2h
    // apart, each workgroup keeping L - 2S of its L = 4 * 64K - 6h cells
    const int hx = c->halo;
    const int Lg = 4 * 64 * c->K - 6 * hx;
    const int Ug = Lg - 2 * c->S_run;
    c->grouped = hx > 0 && !pic && !c->persist && c->K_slab == c->K && (c->K == 2 || c->K == 4) && hx % c->K == 0 &&
                 hx / c->K <= 8 && c->S_run >= hx && 64 * c->K - hx - c->S_run >= 1 && Ug >= 1;
    if (!c->grouped && (U < 1 || Un < 1 && c->K_slab != c->K))
        return fdtd_fail(FDTD_E_ARG, "S = " + std::to_string(c->S_run) + " needs S <= 32*K_slab-1 without the "
                                     "workgroup halo exchange");
    c->tile_sims.clear();
    for (int i = 0; i < nsim; ++i) {
        SimHost* s = c->sims[i];
        if (!s) continue;
        fill_dev<R>(*s, devs[i]);
        if (i < (int)c->beams.size() && c->beams[i]) {     // J' formed in the field step from the CIC sums
            const BeamHost& bh = *c->beams[i];
            char* bb = (char*)bh.block;
            devs[i].pS = (const double*)(bb + bh.off[BO_S]);
            for (int k = 0; k < 2; ++k) devs[i].prange[k] = (const int32_t*)(bb + bh.off[BO_RANGE + k]);
        }
        if (s->steps_done >= s->d.steps) continue;
        c->tile_sims.push_back(i);
        const int kind = s->d.slab_kind == FDTD_SLAB_NONE ? FDTD_SLAB_DIEL : s->d.slab_kind;
        // a tile whose plain-width window [o - S, o + U + S) meets a D-form
        // slab is narrow (K/2 cells per lane) when the tiling asks for it
        const bool slabby = c->K_slab != c->K && s->d.slab_kind >= FDTD_SLAB_LINEAR && s->d.R > s->d.F;
        if (c->grouped) {
            // entries hold window starts; a workgroup's four entries stay together
            for (int ws0 = -c->S_run; ws0 + c->S_run <= s->d.N; ws0 += Ug)
                for (int w = 0; w < 4; ++w) by_kind[kind].push_back(make_int2(i, ws0 + w * (64 * c->K - 2 * hx)));
            continue;
        }
        for (int o = 0; o <= s->d.N;) {
            const bool narrow = slabby && o - c->S_run < s->d.R && o + U + c->S_run > s->d.F;
            by_kind[kind].push_back(make_int2(i | (narrow ? 1 << 24 : 0), o));
            o += narrow ? Un : U;
        }
    }
    // particle beams of the active simulations
    std::vector<BeamDev<R>> bdev;
    std::vector<int2> pblocks, pitems, citems;
    for (int i : c->tile_sims) {
        BeamHost* bh = i < (int)c->beams.size() ? c->beams[i] : nullptr;
        if (!bh) continue;
        const int bi = (int)bdev.size();
        bdev.push_back(fill_beam_dev<R>(*bh, c->d_err, c->sims[i]->launch_begin));
        if (bh->G == 1) {
            // one item (beam, workgroup index) per workgroup of pic_push_cell
            for (int g = 0; g < bh->nwaves / kPushWaves; ++g) citems.push_back(make_int2(bi, g));
        } else {
            for (int w = 0; w < bh->nwaves; ++w) pitems.push_back(make_int2(bi, w));
        }
        // node blocks (push/count writes every particle that stays within
        // its push/count block): nnode for J', start[] and the range, nmove
        // for the staged movers and the parked particles
        for (int k = 0; k < bh->nnode + bh->nmove; ++k) pblocks.push_back(make_int2(bi, -(k + 1)));
    }
    while (pitems.size() % kPushWaves) pitems.push_back(make_int2(-1, 0));   // whole push/count workgroups
    c->has_pic = !bdev.empty();
    c->npnode = (int)pblocks.size();
    c->npblocks = (int)pblocks.size();
    c->npitems = (int)pitems.size();
    c->npcell = (int)citems.size();
    // the one-cell items of a single beam are (beam, 0 .. npcell-1): the kernel forms them itself
    c->pc_single = -1;
    if (!citems.empty()) {
        bool one_beam = true;
        for (size_t k = 0; k < citems.size() && one_beam; ++k)
            one_beam = citems[k].x == citems[0].x && citems[k].y == (int)k;
        if (one_beam) c->pc_single = citems[0].x;
    }
    // LDS rows of the one-cell kernel: what the launch's fullest beam wants
    // when its workgroups fit one per CU (a lone dense beam), else one (a
    // large launch keeps its occupancy); capped by the device's shared memory
    {
        int want = 1;
        for (int i : c->tile_sims) {
            const BeamHost* bh = i < (int)c->beams.size() ? c->beams[i] : nullptr;
            if (bh && bh->G == 1) want = std::max(want, bh->xr_want);
        }
        if (c->cu_count <= 0 || c->npcell > c->cu_count) want = 1;
        if (c->pc_rows > 0) want = c->pc_rows;
        const long long per_row = (long long)kPushWaves * 2 * (long long)kOvfRowBytes<R>;
        const long long room = c->lds_max > 0 ? (long long)c->lds_max - 20 * 1024 : per_row;
        const int cap = (int)std::max<long long>(1, std::min<long long>(kPushCellMaxRows<R>, room / per_row));
        c->pc_xr = std::min(want, cap);
    }
    pitems.insert(pitems.end(), citems.begin(), citems.end());             // the one-cell items follow
    std::vector<Upload> ups;
    if (c->has_pic) {
        ups.push_back({&c->d_beams, &c->d_beams_cap, bdev.data(), sizeof(BeamDev<R>) * bdev.size()});
        ups.push_back({(void**)&c->d_pblocks, &c->d_pblocks_cap, pblocks.data(), sizeof(int2) * pblocks.size()});
        ups.push_back({(void**)&c->d_pitems, &c->d_pitems_cap, pitems.data(), sizeof(int2) * pitems.size()});
    }
    // Costly tiles first: a tile whose window meets a D-form slab, a CPML
    // strip, the source or a probe is dispatched before the plain ones, so
    // the launch's long tiles start at once and the short plain tiles fill
    // the remaining wave slots (measured on config 2 at K=2, S=48: 16.5 us
    // launches against 19.8 us in grid order).
    for (int k = 0; k < 5; ++k) {
        std::vector<int2> heavy, light;
        const size_t step = c->grouped ? 4 : 1;       // whole workgroups in grouped lists
        for (size_t t0 = 0; t0 < by_kind[k].size(); t0 += step) {
            const int2 t = by_kind[k][t0];
            const SimHost* sh = c->sims[t.x & 0xffffff];
            const int Ut = (t.x >> 24) ? Un : U;
            const int lo = c->grouped ? t.y : t.y - c->S_run;
            const int hi = c->grouped ? t.y + Lg : t.y + Ut + c->S_run;
            bool h = (sh->d.slab_kind >= FDTD_SLAB_LINEAR && lo < sh->d.R && hi > sh->d.F) || lo < sh->d.Ls ||
                     hi > sh->d.Rs || (sh->d.src_cell >= lo - 1 && sh->d.src_cell < hi);
            for (int p = 0; p < sh->d.nprobe && !h; ++p) h = sh->d.probe_cell[p] >= lo && sh->d.probe_cell[p] < hi;
            for (size_t q = t0; q < t0 + step; ++q) (h ? heavy : light).push_back(by_kind[k][q]);
        }
        by_kind[k].swap(heavy);
        by_kind[k].insert(by_kind[k].end(), light.begin(), light.end());
    }
    std::vector<int2> tiles;
    c->groups.clear();
    for (int k = 0; k < 5; ++k) {
        if (by_kind[k].empty()) continue;
        c->groups.push_back({k, (int)tiles.size(), (int)by_kind[k].size()});
        tiles.insert(tiles.end(), by_kind[k].begin(), by_kind[k].end());
    }
    ups.push_back({&c->d_sims, &c->d_sims_cap, devs.data(), sizeof(SimDev<R>) * devs.size()});
    ups.push_back({(void**)&c->d_tiles, &c->d_tiles_cap, tiles.data(), sizeof(int2) * tiles.size()});
    // persistent launches: one tile group of uniform tiles whose halos lie in
    // the neighbouring interiors (S <= U), no beam, and every tile resident at
    // once; the neighbour table maps each tile to the tiles on either side
    c->persist_ok = false;
    std::vector<int2> nbr;
    bool energy = false;
    for (int i : c->tile_sims) energy = energy || c->sims[i]->e_every > 0;
    if (c->persist && !pic && !energy && c->groups.size() == 1 && c->K_slab == c->K && c->S_run <= U &&
        c->cu_count > 0) {
        const int per_cu = field_capacity_per_cu<R, sizeof(R) == 4>(c->K, c->groups[0].kind);
        const long long wgs = ((long long)tiles.size() + 3) / 4;
        if (per_cu > 0 && wgs <= (long long)per_cu * c->cu_count) {
            std::vector<std::pair<long long, int>> key(tiles.size());
            for (size_t t = 0; t < tiles.size(); ++t)
                key[t] = {((long long)(tiles[t].x & 0xffffff) << 32) | (unsigned)tiles[t].y, (int)t};
            std::sort(key.begin(), key.end());
            auto find = [&](int sim, long long o) {
                const long long k = ((long long)sim << 32) | (unsigned)o;
                auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(k, INT32_MIN));
                return (it != key.end() && it->first == k) ? it->second : -1;
            };
            nbr.resize(tiles.size());
            for (size_t t = 0; t < tiles.size(); ++t) {
                const int sim = tiles[t].x & 0xffffff, o = tiles[t].y;
                nbr[t] = make_int2(o > 0 ? find(sim, (long long)o - U) : -1,
                                   o + U <= c->sims[sim]->d.N ? find(sim, (long long)o + U) : -1);
            }
            ups.push_back({(void**)&c->d_nbr, &c->d_nbr_cap, nbr.data(), sizeof(int2) * nbr.size()});
            if (sizeof(unsigned long long) * tiles.size() > c->d_flags_cap) {
                // flags only grow: a fresh buffer starts at zero, below every block index
                HIPCHK(hipStreamSynchronize(c->st));
                if (c->d_flags) HIPCHK(hipFree(c->d_flags));
                c->d_flags = nullptr;
                c->d_flags_cap = 0;
                const size_t bytes = sizeof(unsigned long long) * tiles.size() * 2;
                HIPCHK(hipMalloc(&c->d_flags, bytes));
                HIPCHK(hipMemsetAsync(c->d_flags, 0, bytes, c->st));
                c->d_flags_cap = bytes;
            }
            c->persist_ok = true;
        }
    }
    int rc = staged_upload(c, ups);
    if (rc) return rc;
    c->ntiles = (int)tiles.size();
    c->dirty = false;
    return FDTD_OK;
}

// Keeps at most about kThrottleEvery * kThrottleDepth launches queued on the
// stream: a long fdtd_run would otherwise enqueue every kernel of the run
// (three per step with a beam) before the first one finishes.
static int throttle(fdtd_ctx* c) {
    if (++c->throttle_launches % kThrottleEvery) return FDTD_OK;
    if (c->throttle.size() < kThrottleDepth) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->throttle.push_back(e);
    }
    hipEvent_t& slot = c->throttle[c->throttle_head];
    if (c->throttle_launches > (long long)(kThrottleEvery * kThrottleDepth)) HIPCHK(hipEventSynchronize(slot));
    HIPCHK(hipEventRecord(slot, c->st));
    c->throttle_head = (c->throttle_head + 1) % kThrottleDepth;
    return FDTD_OK;
}

// Next profiling event pair (created on first use); nullptr on a HIP error.
static EvPair* next_event(fdtd_ctx* c, int kind) {
    if (c->ev_used == c->events.size()) {
        EvPair p;
        if (hipEventCreate(&p.first) != hipSuccess || hipEventCreate(&p.second) != hipSuccess) {
            fdtd_fail(FDTD_E_HIP, "hipEventCreate failed");
            return nullptr;
        }
        c->events.push_back(p);
    }
    EvPair* e = &c->events[c->ev_used++];
    e->kind = kind;
    e->launches = 1;
    return e;
}

// ev (profiling): ev->first is stamped at the start of the first group's
// kernel, ev->second at the end of the last one.  With fused, the last group
// runs in one launch with the particle scatter (stamped by fev instead).
static int launch_once(fdtd_ctx* c, int n, EvPair* ev, hipStream_t st, bool fused = false,
                       EvPair* fev = nullptr, int nblocks = 1) {
    PersistArgs pa{1, nullptr, nullptr, nullptr, 0, c->grouped ? c->halo : 0};
    if (nblocks > 1) pa = PersistArgs{nblocks, c->d_flags, c->d_nbr, c->d_perr, kPersistSpinTicks, 0};
    const size_t ng = c->groups.size();
    const size_t nplain = fused ? ng - 1 : ng;
    for (size_t gi = 0; gi < ng; ++gi) {
        const TileGroup& g = c->groups[gi];
        hipError_t e;
        if (gi < nplain) {
            hipEvent_t e0 = (ev && gi == 0) ? ev->first : nullptr;
            hipEvent_t e1 = (ev && gi + 1 == nplain) ? ev->second : nullptr;
            if (c->prec == FDTD_F64)
                e = launch_stencil<double, false>((const SimDev<double>*)c->d_sims, c->d_tiles + g.offset, g.count,
                                                  c->launch_idx, c->gstep, n, c->K, c->S_run, g.kind, pa, st, e0, e1);
            else
                e = launch_stencil<float, true>((const SimDev<float>*)c->d_sims, c->d_tiles + g.offset, g.count,
                                                c->launch_idx, c->gstep, n, c->K, c->S_run, g.kind, pa, st, e0, e1);
        } else {
            hipEvent_t e0 = fev ? fev->first : nullptr, e1 = fev ? fev->second : nullptr;
            if (c->prec == FDTD_F64)
                e = launch_field_scatter<double, false>(
                    (const SimDev<double>*)c->d_sims, c->d_tiles + g.offset, g.count, c->launch_idx, c->gstep, n,
                    c->K, c->S_run, g.kind, (const BeamDev<double>*)c->d_beams, c->d_pblocks, c->npblocks,
                    c->npnode, st, e0, e1);
            else
                e = launch_field_scatter<float, true>(
                    (const SimDev<float>*)c->d_sims, c->d_tiles + g.offset, g.count, c->launch_idx, c->gstep, n,
                    c->K, c->S_run, g.kind, (const BeamDev<float>*)c->d_beams, c->d_pblocks, c->npblocks,
                    c->npnode, st, e0, e1);
            if (e != hipSuccess)
                return fdtd_fail(FDTD_E_HIP, std::string("field + scatter launch ") + std::to_string(c->launch_idx) +
                                                 " (" + std::to_string(g.count) + " tiles, " +
                                                 std::to_string(c->npblocks) + " scatter blocks): " + hipGetErrorString(e));
            continue;
        }
        if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("stencil launch: ") + hipGetErrorString(e));
    }
    return FDTD_OK;
}

// Energy sums of the simulations due one after this launch (the bookkeeping
// in fdtd_run has counted the launch): jobs on the context stream.
template <typename R>
static int record_energy(fdtd_ctx* c, const std::vector<int>& due) {
    std::vector<EnergyJob<R>> jobs;
    for (int i : due) {
        SimHost* s = c->sims[i];
        SimDev<R> d;
        fill_dev<R>(*s, d);
        const int par = (int)(s->launches & 1);
        EnergyJob<R> j;
        j.E = d.E[par];
        j.H = d.H[par];
        j.cE = sizeof(R) == 4 ? d.cE[par] : nullptr;
        j.cH = sizeof(R) == 4 ? d.cH[par] : nullptr;
        j.Hprev = (const double*)s->eblock;
        j.out = (double*)((char*)s->eblock + sizeof(double) * (s->d.N + 1)) + s->e_count;
        j.N = s->d.N;
        jobs.push_back(j);
        s->e_steps.push_back(s->steps_done - 1);
        s->e_count++;
    }
    if (jobs.empty()) return FDTD_OK;
    const hipError_t e = launch_energy<R>(jobs.data(), (int)jobs.size(), c->st);
    if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("energy launch: ") + hipGetErrorString(e));
    return FDTD_OK;
}

// Field-only profiling brackets a run of consecutive launches with stream
// markers; this closes the open run (before a rebuild, at the end of a run
// call, before the timings are read).
static int close_bracket(fdtd_ctx* c) {
    if (c->bracket < 0) return FDTD_OK;
    EvPair& b = c->events[(size_t)c->bracket];
    c->bracket = -1;
    if (b.launches > 0) HIPCHK(hipEventRecord(b.second, c->st));
    return FDTD_OK;
}

extern "C" {

int fdtd_run(fdtd_ctx* c, int64_t nsteps) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (nsteps < 0) return fdtd_fail(FDTD_E_ARG, "nsteps < 0");
    int64_t remaining = nsteps;
    while (remaining > 0) {
        if (c->dirty) {
            int rc = close_bracket(c);
            if (rc) return rc;
            rc = c->prec == FDTD_F64 ? rebuild_t<double>(c) : rebuild_t<float>(c);
            if (rc) return rc;
        }
        if (c->ntiles == 0) break;
        if (c->persist_ok && remaining >= 2 * (int64_t)c->S_run) {
            // one persistent launch of nb blocks of S steps (see PersistArgs)
            const int nb = (int)std::min<int64_t>(remaining / c->S_run, kPersistMaxBlocks);
            int rc = close_bracket(c);
            if (rc) return rc;
            hipEvent_t pe0 = nullptr, pe1 = nullptr;
            if (c->profile > 0) {
                EvPair* ev = next_event(c, FDTD_KERNEL_STENCIL);
                if (!ev) return FDTD_E_HIP;
                ev->launches = nb;      // the timing is per block
                pe0 = ev->first;
                pe1 = ev->second;
                HIPCHK(hipEventRecord(pe0, c->st));
            }
            rc = launch_once(c, c->S_run, nullptr, c->st, false, nullptr, nb);
            if (rc) return rc;
            if (pe1) HIPCHK(hipEventRecord(pe1, c->st));
            rc = throttle(c);
            if (rc) return rc;
            c->launches += nb;
            bool finished = false;
            for (int i : c->tile_sims) {
                SimHost* s = c->sims[i];
                s->launches += nb;
                const int64_t done = std::min<int64_t>(s->d.steps, c->gstep + (int64_t)nb * c->S_run - s->t_begin);
                s->steps_done = std::max<int64_t>(s->steps_done, done);
                if (s->steps_done >= s->d.steps) finished = true;
            }
            c->launch_idx += nb;
            c->gstep += (int64_t)nb * c->S_run;
            remaining -= (int64_t)nb * c->S_run;
            if (finished) c->dirty = true;
            continue;
        }
        // steps this launch: S, bounded by the remaining request; one step when
        // a beam is coupled (the deposit feeds every E update)
        int n = (int)std::min<int64_t>(remaining, c->S_run);
        // profiling samples every profile-th launch (events cost host time)
        const bool prof = c->profile > 0 && (c->launches % c->profile) == 0;
        // event handles are copied out at once: next_event may grow (move) the vector
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
        // the fused launch takes field tiles of 1 or 2 cells per lane, and
        // only a small field launch: the fused kernel has the field tile's
        // register count, which cuts the scatter's occupancy (FP32: 8 -> 3
        // waves per SIMD), so a large field launch keeps its own (config 4's
        // 2,048 beam passes: 2.11 ms per step fused against 1.74 separate;
        // config 3's 1,024 tiles: 13.8 us fused against 4.8 + 11.9)
        const bool fused = c->has_pic && c->fuse_pic && (c->K == 1 || c->K == 2) && !c->groups.empty() &&
                           c->groups.back().count <= kFuseMaxTiles;
        if (c->has_pic) {
            if (prof) {
                const EvPair* pe = next_event(c, FDTD_KERNEL_PUSH);
                if (!pe) return FDTD_E_HIP;
                e0 = pe->first; e1 = pe->second;
                const EvPair* se = next_event(c, fused ? FDTD_KERNEL_FIELD_SORT : FDTD_KERNEL_SORT);
                if (!se) return FDTD_E_HIP;
                e2 = se->first; e3 = se->second;
            }
            const hipError_t stale = hipGetLastError();
            if (stale != hipSuccess)
                return fdtd_fail(FDTD_E_HIP, std::string("error pending before the pic launch at launch ") +
                                                 std::to_string(c->launch_idx) + ": " + hipGetErrorString(stale));
            hipError_t e = c->prec == FDTD_F64
                ? launch_push<double>((const SimDev<double>*)c->d_sims, (const BeamDev<double>*)c->d_beams,
                                      c->d_pitems + c->npitems, c->npcell, c->d_pitems, c->npitems,
                                      c->launch_idx, c->pc_xr, c->pc_single, c->st, e0, e1)
                : launch_push<float>((const SimDev<float>*)c->d_sims, (const BeamDev<float>*)c->d_beams,
                                     c->d_pitems + c->npitems, c->npcell, c->d_pitems, c->npitems,
                                     c->launch_idx, c->pc_xr, c->pc_single, c->st, e0, e1);
            if (e != hipSuccess)
                return fdtd_fail(FDTD_E_HIP, std::string("push/count launch ") + std::to_string(c->launch_idx) +
                                                 " (" + std::to_string(c->npcell) + " one-cell workgroups, " +
                                                 std::to_string(c->npitems) + " waves): " + hipGetErrorString(e));
        }
        int rc;
        if (!c->has_pic && c->profile > 0) {
            // field only: runs of `profile` consecutive launches between two
            // stream markers, so the average launch duration includes the
            // gaps between launches and no launch carries an event of its
            // own (a launch stamped by its dispatch packet runs ~1 us longer)
            if (c->bracket < 0) {
                EvPair* b = next_event(c, FDTD_KERNEL_STENCIL);
                if (!b) return FDTD_E_HIP;
                b->launches = 0;
                HIPCHK(hipEventRecord(b->first, c->st));
                c->bracket = (long long)(b - c->events.data());
            }
            rc = launch_once(c, n, nullptr, c->st);
            if (rc) return rc;
            if (++c->events[(size_t)c->bracket].launches == c->profile) {
                rc = close_bracket(c);
                if (rc) return rc;
            }
        } else {
            // a stencil event only when some field launch runs by itself
            EvPair* ev = (prof && !(fused && c->groups.size() == 1)) ? next_event(c, FDTD_KERNEL_STENCIL) : nullptr;
            if (prof && !ev && !(fused && c->groups.size() == 1)) return FDTD_E_HIP;
            EvPair fev{e2, e3, FDTD_KERNEL_FIELD_SORT, 1};
            rc = launch_once(c, n, ev, c->st, fused, prof ? &fev : nullptr);
            if (rc) return rc;
        }
        if (c->has_pic && !fused) {
            hipError_t e = c->prec == FDTD_F64
                ? launch_scatter<double>((const SimDev<double>*)c->d_sims, (const BeamDev<double>*)c->d_beams,
                                         c->d_pblocks, c->npblocks, c->launch_idx, c->st, e2, e3)
                : launch_scatter<float>((const SimDev<float>*)c->d_sims, (const BeamDev<float>*)c->d_beams,
                                        c->d_pblocks, c->npblocks, c->launch_idx, c->st, e2, e3);
            if (e != hipSuccess)
                return fdtd_fail(FDTD_E_HIP, std::string("scatter launch ") + std::to_string(c->launch_idx) + " (" +
                                                 std::to_string(c->npblocks) + " blocks): " + hipGetErrorString(e));
        }
        rc = throttle(c);
        if (rc) return rc;
        c->launches++;
        bool finished = false;
        std::vector<int> due;      // simulations whose energy is taken after this launch
        for (int i : c->tile_sims) {
            SimHost* s = c->sims[i];
            s->launches++;
            const int64_t done = std::min<int64_t>(s->d.steps, c->gstep + n - s->t_begin);
            const bool stepped = done > s->steps_done;
            s->steps_done = std::max<int64_t>(s->steps_done, done);
            if (s->steps_done >= s->d.steps) finished = true;
            if (s->e_every > 0 && stepped && ++s->e_launches % s->e_every == 0 && s->e_count < s->e_rows)
                due.push_back(i);
        }
        if (!due.empty()) {
            rc = c->prec == FDTD_F64 ? record_energy<double>(c, due) : record_energy<float>(c, due);
            if (rc) return rc;
        }
        c->launch_idx++;
        c->gstep += n;
        remaining -= n;
        if (finished) c->dirty = true;
    }
    return close_bracket(c);
}

int fdtd_run_all(fdtd_ctx* c) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    int64_t most = 0;
    for (SimHost* s : c->sims)
        if (s) most = std::max<int64_t>(most, s->d.steps - s->steps_done + (s->t_begin - c->gstep > 0 ? s->t_begin - c->gstep : 0));
    return fdtd_run(c, most);
}

static int check_pic_flag(fdtd_ctx* c) {
    if (!c->d_err) return FDTD_OK;
    int32_t e = 0;
    HIPCHK(hipMemcpy(&e, c->d_err, sizeof(e), hipMemcpyDeviceToHost));
    if (e) return fdtd_fail(FDTD_E_STATE, "particle sort: an index check failed (inconsistent cell starts)");
    return FDTD_OK;
}

// A persistent field launch whose neighbour wait timed out (the flag is
// cleared once reported).  Call after the stream has drained.
static int check_persist_flag(fdtd_ctx* c) {
    if (!c->d_perr || !c->persist) return FDTD_OK;
    int32_t e = 0;
    HIPCHK(hipMemcpy(&e, c->d_perr, sizeof(e), hipMemcpyDeviceToHost));
    if (!e) return FDTD_OK;
    HIPCHK(hipMemset(c->d_perr, 0, sizeof(int32_t)));
    return fdtd_fail(FDTD_E_STATE, "persistent field launch: a neighbour wait timed out (tiles not all resident "
                                   "at once?); the fields of that run are invalid");
}

int fdtd_sync(fdtd_ctx* c) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    HIPCHK(hipStreamSynchronize(c->st));
    int rc = check_persist_flag(c);
    if (rc) return rc;
    return c->has_pic ? check_pic_flag(c) : FDTD_OK;
}

}  // extern "C"

// Checkpoint of one simulation: a header, the energy rows' step indices, then
// the device bytes of the simulation block's state region (from the first
// state sub-array: fields of both ping-pong buffers, ADE and CPML state,
// probes, snapshot rows, J), of the beam block and of the energy block, as
// the device holds them.  The coefficients are not saved: a checkpoint is
// loaded into a simulation built from the same description.
namespace {
struct CkptHeader {
    char magic[8];
    int32_t prec, N, steps, nprobe;
    int64_t state_bytes, beam_bytes, e_bytes, beam_np;
    int64_t steps_done, launches, rel_launch, rel_step;
    int64_t e_count, e_launches, n_esteps;
};
constexpr char kCkptMagic[8] = {'F', 'D', 'T', 'D', 'C', 'K', 'P', '1'};

struct CkptParts {
    size_t state0 = 0, state_bytes = 0, beam_bytes = 0, e_bytes = 0;
    const BeamHost* b = nullptr;
    int64_t total = 0;
};

CkptParts ckpt_parts(fdtd_ctx* c, int id, const SimHost* s) {
    CkptParts p;
    p.state0 = s->lay_off[14];
    p.state_bytes = s->bytes - p.state0;
    p.b = (id < (int)c->beams.size()) ? c->beams[id] : nullptr;
    p.beam_bytes = p.b ? p.b->bytes : 0;
    p.e_bytes = s->eblock ? s->ebytes : 0;
    p.total = (int64_t)(sizeof(CkptHeader) + sizeof(int64_t) * s->e_steps.size() + p.state_bytes + p.beam_bytes +
                        p.e_bytes);
    return p;
}
}  // namespace

extern "C" {

static SimHost* ckpt_sim(fdtd_ctx* c, int id) {
    return (c && id >= 0 && id < (int)c->sims.size()) ? c->sims[id] : nullptr;
}

int fdtd_state_size(fdtd_ctx* c, int id, int64_t* bytes) {
    SimHost* s = ckpt_sim(c, id);
    if (!s || !bytes) return fdtd_fail(FDTD_E_ARG, "bad sim id or null output");
    *bytes = ckpt_parts(c, id, s).total;
    return FDTD_OK;
}

int fdtd_state_save(fdtd_ctx* c, int id, void* buf, int64_t len) {
    SimHost* s = ckpt_sim(c, id);
    if (!s || !buf) return fdtd_fail(FDTD_E_ARG, "bad sim id or null buffer");
    const CkptParts p = ckpt_parts(c, id, s);
    if (len < p.total) return fdtd_fail(FDTD_E_ARG, "checkpoint buffer too small (fdtd_state_size)");
    if (int rc = fdtd_sync(c)) return rc;           // the state after every queued launch
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, sizeof(h.magic));
    h.prec = c->prec;
    h.N = s->d.N;
    h.steps = s->d.steps;
    h.nprobe = s->d.nprobe;
    h.state_bytes = (int64_t)p.state_bytes;
    h.beam_bytes = (int64_t)p.beam_bytes;
    h.e_bytes = (int64_t)p.e_bytes;
    h.beam_np = p.b ? p.b->np : -1;
    h.steps_done = s->steps_done;
    h.launches = s->launches;
    h.rel_launch = c->launch_idx - s->launch_begin;    // ping-pong parity of the next launch
    h.rel_step = c->gstep - s->t_begin;               // the simulation's time index of the next step
    h.e_count = s->e_count;
    h.e_launches = s->e_launches;
    h.n_esteps = (int64_t)s->e_steps.size();
    char* o = (char*)buf;
    std::memcpy(o, &h, sizeof(h));
    o += sizeof(h);
    if (!s->e_steps.empty()) std::memcpy(o, s->e_steps.data(), sizeof(int64_t) * s->e_steps.size());
    o += sizeof(int64_t) * s->e_steps.size();
    HIPCHK(hipMemcpy(o, (const char*)s->block + p.state0, p.state_bytes, hipMemcpyDeviceToHost));
    o += p.state_bytes;
    if (p.beam_bytes) HIPCHK(hipMemcpy(o, p.b->block, p.beam_bytes, hipMemcpyDeviceToHost));
    o += p.beam_bytes;
    if (p.e_bytes) HIPCHK(hipMemcpy(o, s->eblock, p.e_bytes, hipMemcpyDeviceToHost));
    return FDTD_OK;
}

int fdtd_state_load(fdtd_ctx* c, int id, const void* buf, int64_t len) {
    SimHost* s = ckpt_sim(c, id);
    if (!s || !buf) return fdtd_fail(FDTD_E_ARG, "bad sim id or null buffer");
    CkptHeader h;
    if (len < (int64_t)sizeof(h)) return fdtd_fail(FDTD_E_ARG, "checkpoint: buffer shorter than its header");
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kCkptMagic, sizeof(h.magic)) != 0) return fdtd_fail(FDTD_E_ARG, "checkpoint: bad magic");
    const CkptParts p = ckpt_parts(c, id, s);
    if (h.prec != c->prec || h.N != s->d.N || h.steps != s->d.steps || h.nprobe != s->d.nprobe ||
        h.state_bytes != (int64_t)p.state_bytes)
        return fdtd_fail(FDTD_E_ARG, "checkpoint: precision or simulation shape differs from the target");
    if (h.beam_bytes != (int64_t)p.beam_bytes || h.beam_np != (p.b ? p.b->np : -1))
        return fdtd_fail(FDTD_E_ARG, "checkpoint: beam differs from the target's (attach the same beam first)");
    if (h.e_bytes != (int64_t)p.e_bytes)
        return fdtd_fail(FDTD_E_ARG, "checkpoint: energy recording differs from the target's (fdtd_set_energy)");
    if (h.n_esteps != h.e_count || h.n_esteps < 0 ||
        len < (int64_t)(sizeof(h) + sizeof(int64_t) * h.n_esteps + p.state_bytes + p.beam_bytes + p.e_bytes))
        return fdtd_fail(FDTD_E_ARG, "checkpoint: buffer shorter than the checkpoint it describes");
    if (h.steps_done < 0 || h.steps_done > h.steps) return fdtd_fail(FDTD_E_ARG, "checkpoint: bad step count");
    const char* o = (const char*)buf + sizeof(h);
    std::vector<int64_t> esteps((size_t)h.n_esteps);
    if (h.n_esteps) std::memcpy(esteps.data(), o, sizeof(int64_t) * h.n_esteps);
    o += sizeof(int64_t) * h.n_esteps;
    // ordered after the target's queued launches; the host buffer is
    // pageable, so the copies finish before returning
    HIPCHK(hipMemcpyAsync((char*)s->block + p.state0, o, p.state_bytes, hipMemcpyHostToDevice, c->st));
    o += p.state_bytes;
    if (p.beam_bytes) HIPCHK(hipMemcpyAsync(p.b->block, o, p.beam_bytes, hipMemcpyHostToDevice, c->st));
    o += p.beam_bytes;
    if (p.e_bytes) HIPCHK(hipMemcpyAsync(s->eblock, o, p.e_bytes, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    s->steps_done = h.steps_done;
    s->launches = h.launches;
    s->launch_begin = c->launch_idx - h.rel_launch;
    s->t_begin = c->gstep - h.rel_step;
    s->e_count = (int)h.e_count;
    s->e_launches = h.e_launches;
    s->e_steps = std::move(esteps);
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_set_energy(fdtd_ctx* c, int id, int every, int rows) {
    if (!c || id < 0 || id >= (int)c->sims.size() || !c->sims[id]) return fdtd_fail(FDTD_E_ARG, "bad sim id");
    if (every < 0 || rows < 0 || (every > 0 && rows < 1))
        return fdtd_fail(FDTD_E_ARG, "energy: every >= 0 and rows >= 1 when on");
    SimHost* s = c->sims[id];
    HIPCHK(hipStreamSynchronize(c->st));
    if (s->eblock) {
        (void)hipFree(s->eblock);
        c->dev_bytes -= s->ebytes;
        s->eblock = nullptr;
        s->ebytes = 0;
    }
    s->e_every = 0;
    s->e_rows = s->e_count = 0;
    s->e_launches = 0;
    s->e_steps.clear();
    if (every > 0) {
        const size_t bytes = sizeof(double) * ((size_t)s->d.N + 1 + (size_t)rows);
        if (hipMalloc(&s->eblock, bytes) != hipSuccess) {
            s->eblock = nullptr;
            return fdtd_fail(FDTD_E_NOMEM, "hipMalloc of the energy rows failed");
        }
        HIPCHK(hipMemsetAsync(s->eblock, 0, bytes, c->st));
        s->ebytes = bytes;
        c->dev_bytes += bytes;
        s->e_every = every;
        s->e_rows = rows;
    }
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_read_energy(fdtd_ctx* c, int id, double* values, int64_t* steps, int64_t cap, int64_t* count) {
    if (!c || id < 0 || id >= (int)c->sims.size() || !c->sims[id] || !count || cap < 0 ||
        (cap > 0 && (!values || !steps)))
        return fdtd_fail(FDTD_E_ARG, "bad argument");
    SimHost* s = c->sims[id];
    HIPCHK(hipStreamSynchronize(c->st));
    const int64_t n = std::min<int64_t>(cap, s->e_count);
    *count = s->e_count;
    if (n > 0) {
        HIPCHK(hipMemcpy(values, (char*)s->eblock + sizeof(double) * (s->d.N + 1), sizeof(double) * n,
                         hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < n; ++k) steps[k] = s->e_steps[k];
    }
    return FDTD_OK;
}

int fdtd_set_persistent(fdtd_ctx* c, int on) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (on != 0 && on != 1) return fdtd_fail(FDTD_E_ARG, "persistent flag must be 0 or 1");
    c->persist = on != 0;
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_persistent_active(fdtd_ctx* c, int* active) {
    if (!c || !active) return fdtd_fail(FDTD_E_ARG, "null argument");
    if (c->dirty) {
        int rc = close_bracket(c);
        if (rc) return rc;
        rc = c->prec == FDTD_F64 ? rebuild_t<double>(c) : rebuild_t<float>(c);
        if (rc) return rc;
    }
    *active = c->persist_ok ? 1 : 0;
    return FDTD_OK;
}

int fdtd_set_pic_lds_rows(fdtd_ctx* c, int rows) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (rows < 0 || rows > 8) return fdtd_fail(FDTD_E_ARG, "LDS rows must be 0 (chosen per launch) .. 8");
    c->pc_rows = rows;
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_set_halo_exchange(fdtd_ctx* c, int h) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (h < 0 || h > 32) return fdtd_fail(FDTD_E_ARG, "halo exchange period must be 0 (off) .. 32");
    c->halo = h;
    c->dirty = true;
    return FDTD_OK;
}

int fdtd_halo_exchange_active(fdtd_ctx* c, int* active) {
    if (!c || !active) return fdtd_fail(FDTD_E_ARG, "null argument");
    if (c->dirty) {
        int rc = close_bracket(c);
        if (rc) return rc;
        rc = c->prec == FDTD_F64 ? rebuild_t<double>(c) : rebuild_t<float>(c);
        if (rc) return rc;
    }
    *active = c->grouped ? 1 : 0;
    return FDTD_OK;
}

int fdtd_busy(fdtd_ctx* c, int* busy) {
    if (!c || !busy) return fdtd_fail(FDTD_E_ARG, "null argument");
    const hipError_t e = hipStreamQuery(c->st);
    if (e == hipErrorNotReady) { *busy = 1; return FDTD_OK; }
    if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
    *busy = 0;
    return FDTD_OK;
}

int fdtd_kernel_timing(fdtd_ctx* c, int kind, double* ms, int64_t* launches) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (kind < FDTD_KERNEL_STENCIL || kind > FDTD_KERNEL_FIELD_SORT) return fdtd_fail(FDTD_E_ARG, "bad kernel kind");
    int rc = close_bracket(c);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    double tot = 0;
    int64_t n = 0;
    for (size_t i = 0; i < c->ev_used; ++i) {
        if (c->events[i].kind != kind || c->events[i].launches <= 0) continue;
        float m = 0;
        HIPCHK(hipEventElapsedTime(&m, c->events[i].first, c->events[i].second));
        tot += m;
        n += c->events[i].launches;
    }
    if (ms) *ms = tot;
    if (launches) *launches = n;
    return FDTD_OK;
}

int fdtd_timing(fdtd_ctx* c, double* ms, int64_t* launches, int reset) {
    int rc = fdtd_kernel_timing(c, FDTD_KERNEL_STENCIL, ms, launches);
    if (rc) return rc;
    if (reset) {
        c->ev_used = 0;
        c->bracket = -1;
    }
    return FDTD_OK;
}

int fdtd_set_profile(fdtd_ctx* c, int every) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (every < 0) return fdtd_fail(FDTD_E_ARG, "negative sampling interval");
    c->profile = every;
    c->ev_used = 0;
    c->bracket = -1;
    return FDTD_OK;
}

int fdtd_set_pic_fusion(fdtd_ctx* c, int on) {
    if (!c) return fdtd_fail(FDTD_E_ARG, "null ctx");
    if (on != 0 && on != 1) return fdtd_fail(FDTD_E_ARG, "pic fusion flag must be 0 or 1");
    c->fuse_pic = on != 0;
    return FDTD_OK;
}

int fdtd_device_bytes(fdtd_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return fdtd_fail(FDTD_E_ARG, "null argument");
    *bytes = c->dev_bytes;
    return FDTD_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------
// reads / writes
// ------------------------------------------------------------------------
template <typename R>
static int xfer_field(fdtd_ctx* c, SimHost& s, int field, double* out, const double* in, int64_t len) {
    SimDev<R> d;
    fill_dev<R>(s, d);
    const int par = (int)(s.launches & 1);
    const int N = s.d.N;
    const R* src = nullptr;
    int64_t n = 0;
    bool strip = false;
    switch (field) {
        case FDTD_FIELD_E: src = d.E[par]; n = N + 1; break;
        case FDTD_FIELD_H: src = d.H[par]; n = N + 1; break;
        case FDTD_FIELD_D: src = d.D[par]; n = s.ns; break;
        case FDTD_FIELD_P: src = d.P[par]; n = s.ns; break;
        case FDTD_FIELD_PM:
            if (sizeof(R) == 4 && out) {   // FP32 keeps the velocity V = P^n - P^{n-1}
                std::vector<double> p(s.ns), v(s.ns);
                int rc = xfer_field<R>(c, s, FDTD_FIELD_P, p.data(), nullptr, s.ns);
                if (rc) return rc;
                if (len != s.ns) return fdtd_fail(FDTD_E_ARG, "length mismatch");
                std::vector<R> hv(s.ns);
                if (s.ns) HIPCHK(hipMemcpy(hv.data(), d.Pm[par], sizeof(R) * s.ns, hipMemcpyDeviceToHost));
                for (int i = 0; i < s.ns; ++i) out[i] = p[i] - (double)hv[i];
                return FDTD_OK;
            }
            if (sizeof(R) == 4) return fdtd_fail(FDTD_E_ARG, "PM is derived in FP32 and cannot be written");
            src = d.Pm[par]; n = s.ns; break;
        case FDTD_FIELD_PSIE: src = d.psiE[par]; n = s.nstrip; strip = true; break;
        case FDTD_FIELD_PSIH: src = d.psiH[par]; n = s.nstrip; strip = true; break;
        case FDTD_FIELD_J: src = d.J; n = d.J ? N + 1 : 0; break;
        default: return fdtd_fail(FDTD_E_ARG, "bad field id");
    }
    const int64_t want = strip ? (N + 1) : n;
    if (len != want) return fdtd_fail(FDTD_E_ARG, "length mismatch: expected " + std::to_string(want));
    if (n == 0) return FDTD_OK;
    std::vector<R> h(n);
    HIPCHK(hipStreamSynchronize(c->st));
    if (int rc = check_persist_flag(c)) return rc;
    if (out) {
        HIPCHK(hipMemcpy(h.data(), src, sizeof(R) * n, hipMemcpyDeviceToHost));
        if (!strip) {
            for (int64_t i = 0; i < n; ++i) out[i] = (double)h[i];
            // FP32 Kahan state: report the compensated value for E/H/D/P
            if (sizeof(R) == 4 && (field <= FDTD_FIELD_P)) {
                const R* comp = field == FDTD_FIELD_E ? d.cE[par] : field == FDTD_FIELD_H ? d.cH[par]
                              : field == FDTD_FIELD_D ? d.cD[par] : d.cP[par];
                std::vector<R> cc(n);
                HIPCHK(hipMemcpy(cc.data(), comp, sizeof(R) * n, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; ++i) out[i] = (double)h[i] - (double)cc[i];
            }
        } else {
            std::fill(out, out + N + 1, 0.0);
            for (int i = 0; i < s.d.Ls; ++i) out[i] = (double)h[i];
            for (int i = s.d.Rs; i <= N; ++i) out[i] = (double)h[s.d.Ls + (i - s.d.Rs)];
        }
    } else {
        if (!strip) {
            for (int64_t i = 0; i < n; ++i) h[i] = (R)in[i];
        } else {
            for (int i = 0; i < s.d.Ls; ++i) h[i] = (R)in[i];
            for (int i = s.d.Rs; i <= N; ++i) h[s.d.Ls + (i - s.d.Rs)] = (R)in[i];
        }
        HIPCHK(hipMemcpy((void*)src, h.data(), sizeof(R) * n, hipMemcpyHostToDevice));
        if (sizeof(R) == 4 && field <= FDTD_FIELD_P) {
            R* comp = field == FDTD_FIELD_E ? d.cE[par] : field == FDTD_FIELD_H ? d.cH[par]
                    : field == FDTD_FIELD_D ? d.cD[par] : d.cP[par];
            HIPCHK(hipMemset(comp, 0, sizeof(R) * n));
        }
    }
    return FDTD_OK;
}

template <typename R>
static int read_plain(fdtd_ctx* c, const R* src, int64_t n, double* out) {
    std::vector<R> h(n);
    HIPCHK(hipStreamSynchronize(c->st));
    if (int rc = check_persist_flag(c)) return rc;
    if (n) HIPCHK(hipMemcpy(h.data(), src, sizeof(R) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) out[i] = (double)h[i];
    return FDTD_OK;
}

static SimHost* get_sim(fdtd_ctx* c, int id) {
    if (!c || id < 0 || id >= (int)c->sims.size()) return nullptr;
    return c->sims[id];
}

extern "C" {

int fdtd_read_field(fdtd_ctx* c, int id, int field, double* out, int64_t len) {
    SimHost* s = get_sim(c, id);
    if (!s || !out) return fdtd_fail(FDTD_E_ARG, "bad sim id or null out");
    return c->prec == FDTD_F64 ? xfer_field<double>(c, *s, field, out, nullptr, len)
                               : xfer_field<float>(c, *s, field, out, nullptr, len);
}

int fdtd_write_field(fdtd_ctx* c, int id, int field, const double* in, int64_t len) {
    SimHost* s = get_sim(c, id);
    if (!s || !in) return fdtd_fail(FDTD_E_ARG, "bad sim id or null input");
    if (field == FDTD_FIELD_J) return fdtd_fail(FDTD_E_ARG, "J is written by the deposit");
    return c->prec == FDTD_F64 ? xfer_field<double>(c, *s, field, nullptr, in, len)
                               : xfer_field<float>(c, *s, field, nullptr, in, len);
}

int fdtd_read_probes(fdtd_ctx* c, int id, double* out, int64_t len) {
    SimHost* s = get_sim(c, id);
    if (!s || !out) return fdtd_fail(FDTD_E_ARG, "bad sim id or null out");
    const int64_t n = (int64_t)s->d.nprobe * s->d.steps;
    if (len != n) return fdtd_fail(FDTD_E_ARG, "length mismatch: expected " + std::to_string(n));
    if (c->prec == FDTD_F64) { SimDev<double> d; fill_dev<double>(*s, d); return read_plain<double>(c, d.probe_out, n, out); }
    SimDev<float> d; fill_dev<float>(*s, d); return read_plain<double>(c, d.probe_out, n, out);
}

int fdtd_read_snapshots(fdtd_ctx* c, int id, double* out, int64_t len) {
    SimHost* s = get_sim(c, id);
    if (!s || !out) return fdtd_fail(FDTD_E_ARG, "bad sim id or null out");
    const int64_t n = (int64_t)s->d.snap_rows * (s->d.N + 1);
    if (len != n) return fdtd_fail(FDTD_E_ARG, "length mismatch: expected " + std::to_string(n));
    if (c->prec == FDTD_F64) { SimDev<double> d; fill_dev<double>(*s, d); return read_plain<double>(c, d.snap, n, out); }
    SimDev<float> d; fill_dev<float>(*s, d); return read_plain<float>(c, d.snap, n, out);
}

}  // extern "C"

// ------------------------------------------------------------------------
// particle beams: C ABI
// ------------------------------------------------------------------------
template <typename R>
static int read_beam(fdtd_ctx* c, SimHost& s, BeamHost& b, double* x, double* u, int8_t* alive) {
    const int par = (int)(s.launches & 1);
    const int64_t n = b.np;
    const int N = s.d.N;
    std::vector<int32_t> cl(n), id(n);
    std::vector<R> fr(n), ux(n), uy(n), uz(n);
    HIPCHK(hipStreamSynchronize(c->st));
    int rc = check_pic_flag(c);
    if (rc) return rc;
    const char* blk = (const char*)b.block;
    const size_t rn = sizeof(R) * n, in = sizeof(int32_t) * n;
    if (b.G == 1) {
        // one-cell beams keep no cell array: cell c holds [start[c], start[c+1])
        std::vector<int32_t> st(N + 2);
        HIPCHK(hipMemcpy(st.data(), blk + b.off[BO_START + par], sizeof(int32_t) * (N + 2), hipMemcpyDeviceToHost));
        for (int cc = 0; cc <= N + 1; ++cc) {
            const int64_t e = cc <= N ? st[cc + 1] : n;
            if (st[cc] < 0 || e > n || e < st[cc]) return fdtd_fail(FDTD_E_STATE, "corrupt start[] table");
            for (int64_t k = st[cc]; k < e; ++k) cl[k] = cc;
        }
    } else {
        HIPCHK(hipMemcpy(cl.data(), blk + b.off[BO_CELL + par], in, hipMemcpyDeviceToHost));
    }
    HIPCHK(hipMemcpy(id.data(), blk + b.off[BO_ID + par], in, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(fr.data(), blk + b.off[BO_FRAC + par], rn, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ux.data(), blk + b.off[BO_UX + par], rn, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(uy.data(), blk + b.off[BO_UY + par], rn, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(uz.data(), blk + b.off[BO_UZ + par], rn, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = id[k];
        if (i < 0 || i >= n) return fdtd_fail(FDTD_E_STATE, "particle id out of range");
        x[i] = (double)cl[k] + (double)fr[k];
        u[3 * i] = (double)ux[k]; u[3 * i + 1] = (double)uy[k]; u[3 * i + 2] = (double)uz[k];
        if (alive) alive[i] = (int8_t)(cl[k] >= 2 && cl[k] <= N - 3);
    }
    return FDTD_OK;
}

extern "C" {

int pic_attach(fdtd_ctx* c, int id, const fdtd_beam_desc_t* bd) {
    SimHost* s = get_sim(c, id);
    if (!s || !bd) return fdtd_fail(FDTD_E_ARG, "bad sim id or null beam");
    if (!s->d.has_current) return fdtd_fail(FDTD_E_ARG, "the simulation was created without a current array");
    if (s->steps_done > 0 || s->launches > 0) return fdtd_fail(FDTD_E_STATE, "the simulation has already run");
    if (id < (int)c->beams.size() && c->beams[id]) return fdtd_fail(FDTD_E_STATE, "a beam is already attached");
    if (bd->n < 1 || bd->n > (int64_t)INT32_MAX / 2) return fdtd_fail(FDTD_E_ARG, "particle count out of range");
    if (!bd->x || !bd->u) return fdtd_fail(FDTD_E_ARG, "null particle arrays");
    if (!(bd->m > 0) || !(bd->dt > 0) || !(bd->dz > 0)) return fdtd_fail(FDTD_E_ARG, "m, dt and dz must be > 0");
    if (!(299792458.0 * bd->dt / bd->dz < 1.0)) return fdtd_fail(FDTD_E_ARG, "c dt / dz must be < 1");
    if (bd->group < 0 || bd->group > kPicGMax) return fdtd_fail(FDTD_E_ARG, "group must be 0..64");
    const int N = s->d.N;
    if (N < 8) return fdtd_fail(FDTD_E_ARG, "grid too small for particles");
    for (int64_t i = 0; i < bd->n; ++i) {
        const double x = bd->x[i];
        if (!(x >= 2.0 && x < (double)(N - 2))) return fdtd_fail(FDTD_E_ARG, "particle position outside [2, N-2)");
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(bd->u[3 * i + k])) return fdtd_fail(FDTD_E_ARG, "non-finite particle momentum");
    }
    return c->prec == FDTD_F64 ? attach_t<double>(c, id, *s, bd) : attach_t<float>(c, id, *s, bd);
}

int pic_read(fdtd_ctx* c, int id, double* x, double* u, int8_t* alive, int64_t n) {
    SimHost* s = get_sim(c, id);
    if (!s || !x || !u) return fdtd_fail(FDTD_E_ARG, "bad sim id or null output");
    BeamHost* b = id < (int)c->beams.size() ? c->beams[id] : nullptr;
    if (!b) return fdtd_fail(FDTD_E_ARG, "no beam attached");
    if (n != b->np) return fdtd_fail(FDTD_E_ARG, "length mismatch: expected " + std::to_string(b->np));
    return c->prec == FDTD_F64 ? read_beam<double>(c, *s, *b, x, u, alive) : read_beam<float>(c, *s, *b, x, u, alive);
}

}  // extern "C"

// ------------------------------------------------------------------------
// spectra of probe series (FP64 DFT on the device)
// ------------------------------------------------------------------------
// The series of probe `probe` of each listed simulation (device pointers to
// its FP64 samples, and lengths; host side).
template <typename R>
static int spectra_series(fdtd_ctx* c, int nsims, const int* ids, int probe, const double** ptrs, int* lens) {
    for (int k = 0; k < nsims; ++k) {
        SimHost* s = get_sim(c, ids[k]);
        if (!s) return fdtd_fail(FDTD_E_ARG, "bad sim id " + std::to_string(ids[k]));
        if (probe < 0 || probe >= s->d.nprobe) return fdtd_fail(FDTD_E_ARG, "probe index out of range");
        SimDev<R> d;
        fill_dev<R>(*s, d);
        ptrs[k] = d.probe_out + (size_t)probe * s->d.steps;
        lens[k] = s->d.steps;
    }
    return FDTD_OK;
}

// Uploads the tables into the device buffers given, runs the DFT and copies
// the results to out; synchronous.
static int spectra_run(fdtd_ctx* c, int nsims, const double* const* ptrs, const int* lens, const double* freqs,
                       int nf, double dt, double* out, const double** d_ptrs, int* d_lens, double* d_freqs,
                       double* d_out) {
    HIPCHK(hipMemcpyAsync(d_ptrs, ptrs, sizeof(void*) * nsims, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d_lens, lens, sizeof(int) * nsims, hipMemcpyHostToDevice, c->st));
    HIPCHK(hipMemcpyAsync(d_freqs, freqs, sizeof(double) * nf, hipMemcpyHostToDevice, c->st));
    for (int k0 = 0; k0 < nsims; k0 += 32768) {
        const int nk = std::min(32768, nsims - k0);
        hipError_t e = launch_dft(d_ptrs + k0, d_lens + k0, nk, d_freqs, nf, dt, d_out + 2 * (size_t)k0 * nf, c->st);
        if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("dft launch: ") + hipGetErrorString(e));
    }
    HIPCHK(hipMemcpyAsync(out, d_out, sizeof(double) * 2 * (size_t)nsims * nf, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    return FDTD_OK;
}

template <typename R>
static int spectra_t(fdtd_ctx* c, int nsims, const int* ids, int probe, const double* freqs, int nf, double dt,
                     double* out) {
    std::vector<const double*> ptrs(nsims);
    std::vector<int> lens(nsims);
    int rc = spectra_series<R>(c, nsims, ids, probe, ptrs.data(), lens.data());
    if (rc) return rc;
    const size_t o_lens = align_up(sizeof(void*) * nsims, 256);
    const size_t o_freqs = o_lens + align_up(sizeof(int) * nsims, 256);
    const size_t o_out = o_freqs + align_up(sizeof(double) * nf, 256);
    const size_t need = o_out + sizeof(double) * 2 * (size_t)nsims * nf;
    if (need > c->d_scratch_cap) {
        if (c->d_scratch) HIPCHK(hipFree(c->d_scratch));
        c->d_scratch = nullptr;
        c->d_scratch_cap = 0;
        HIPCHK(hipMalloc(&c->d_scratch, need));
        c->d_scratch_cap = need;
    }
    char* base = (char*)c->d_scratch;
    return spectra_run(c, nsims, ptrs.data(), lens.data(), freqs, nf, dt, out, (const double**)base,
                       (int*)(base + o_lens), (double*)(base + o_freqs), (double*)(base + o_out));
}

// The same spectra through the library's static scratch array of this
// context's device: no device or host allocation (tables on the stack).
template <typename R>
static int spectra_noalloc_t(fdtd_ctx* c, int nsims, const int* ids, int probe, const double* freqs, int nf,
                             double dt, double* out) {
    char* base = (char*)noalloc_scratch(c->device);
    if (!base) return fdtd_fail(FDTD_E_STATE, "no-allocation scratch not resolved on this device");
    const double* ptrs[kNaMaxSims];
    int lens[kNaMaxSims];
    int rc = spectra_series<R>(c, nsims, ids, probe, ptrs, lens);
    if (rc) return rc;
    return spectra_run(c, nsims, ptrs, lens, freqs, nf, dt, out, (const double**)base, (int*)(base + kNaOffLens),
                       (double*)(base + kNaOffFreqs), (double*)(base + kNaOffOut));
}

extern "C" int fdtd_probe_spectra(fdtd_ctx* c, int nsims, const int* sim_ids, int probe, const double* freqs, int nf,
                                  double dt, double* out) {
    if (!c || nsims < 0 || (nsims > 0 && (!sim_ids || !out)) || nf < 0 || (nf > 0 && !freqs))
        return fdtd_fail(FDTD_E_ARG, "bad argument");
    if (nsims == 0 || nf == 0) return FDTD_OK;
    return c->prec == FDTD_F64 ? spectra_t<double>(c, nsims, sim_ids, probe, freqs, nf, dt, out)
                               : spectra_t<float>(c, nsims, sim_ids, probe, freqs, nf, dt, out);
}

extern "C" int fdtd_probe_spectra_noalloc(fdtd_ctx* c, int nsims, const int* sim_ids, int probe,
                                          const double* freqs, int nf, double dt, double* out) {
    if (!c || nsims < 0 || (nsims > 0 && (!sim_ids || !out)) || nf < 0 || (nf > 0 && !freqs))
        return fdtd_fail(FDTD_E_ARG, "bad argument");
    if (nsims > kNaMaxSims || nf > kNaMaxFreqs || (int64_t)nsims * nf > kNaMaxPairs)
        return fdtd_fail(FDTD_E_ARG, "fdtd_probe_spectra_noalloc: " + std::to_string(nsims) + " series x " +
                                         std::to_string(nf) + " frequencies exceeds its limits (" +
                                         std::to_string(kNaMaxSims) + " series, " + std::to_string(kNaMaxFreqs) +
                                         " frequencies, " + std::to_string(kNaMaxPairs) + " pairs)");
    if (nsims == 0 || nf == 0) return FDTD_OK;
    return c->prec == FDTD_F64 ? spectra_noalloc_t<double>(c, nsims, sim_ids, probe, freqs, nf, dt, out)
                               : spectra_noalloc_t<float>(c, nsims, sim_ids, probe, freqs, nf, dt, out);
}

// ------------------------------------------------------------------------
// single reference update kernels (compatibility path of BaseFDTD11)
// ------------------------------------------------------------------------
namespace {
struct OpSpec {
    int narr, nscal, nout;
    int prev, next;        // reads index n-1 / n+1
};
// indexed by op: EX_UPDATE HY_UPDATE PSI_E PSI_M DX_UPDATE EX_CREATE POLARISE ACUBIC NONLIN_EX,
// then BaseFDTD's B_EX B_PSI_E B_PSI_M B_POLARISE B_DX
constexpr int kNumOps = 14;
const OpSpec kOps[kNumOps] = {{5, 0, 1, 1, 0}, {5, 0, 1, 0, 1}, {6, 0, 2, 1, 0}, {6, 0, 2, 0, 1}, {3, 1, 1, 1, 0},
                              {3, 1, 1, 0, 0}, {3, 3, 1, 0, 0}, {2, 4, 1, 0, 0}, {3, 3, 1, 0, 0}, {5, 0, 1, 1, 0},
                              {6, 0, 2, 1, 0}, {6, 0, 2, 0, 1}, {4, 3, 1, 0, 0}, {4, 3, 2, 1, 0}};
}  // namespace

extern "C" int fdtd_step_op(fdtd_ctx* c, int op, int narr, double* const* arrays, const int64_t* lens, int nscal,
                            const double* scalars, const int32_t* ranges) {
    if (!c || op < 0 || op >= kNumOps || !arrays || !lens || !ranges) return fdtd_fail(FDTD_E_ARG, "bad argument");
    if (c->prec != FDTD_F64) return fdtd_fail(FDTD_E_ARG, "the single-op path runs in an FP64 context");
    const OpSpec& o = kOps[op];
    if (narr != o.narr || nscal != o.nscal || (nscal > 0 && !scalars))
        return fdtd_fail(FDTD_E_ARG, "op " + std::to_string(op) + " takes " + std::to_string(o.narr) + " arrays and " +
                                         std::to_string(o.nscal) + " scalars");
    int64_t minlen = INT64_MAX, total = 0;
    for (int k = 0; k < narr; ++k) {
        if (!arrays[k] || lens[k] < 0) return fdtd_fail(FDTD_E_ARG, "null or negative-length array");
        minlen = std::min(minlen, lens[k]);
        total += lens[k];
    }
    if (minlen > INT32_MAX) return fdtd_fail(FDTD_E_ARG, "arrays too long");
    for (int r = 0; r < 2; ++r) {
        const int lo = ranges[2 * r], hi = ranges[2 * r + 1];
        if (hi <= lo) continue;
        if (lo - o.prev < 0 || hi + o.next > minlen) return fdtd_fail(FDTD_E_ARG, "index range outside the arrays");
    }
    std::vector<size_t> off(narr);
    size_t need = 0;
    for (int k = 0; k < narr; ++k) { off[k] = need; need = align_up(need + sizeof(double) * lens[k], 256); }
    if (need > c->d_scratch_cap) {
        if (c->d_scratch) HIPCHK(hipFree(c->d_scratch));
        c->d_scratch = nullptr;
        c->d_scratch_cap = 0;
        HIPCHK(hipMalloc(&c->d_scratch, need));
        c->d_scratch_cap = need;
    }
    OpArgs g;
    std::memset(&g, 0, sizeof(g));
    for (int k = 0; k < narr; ++k) {
        g.a[k] = (double*)((char*)c->d_scratch + off[k]);
        if (lens[k]) HIPCHK(hipMemcpyAsync(g.a[k], arrays[k], sizeof(double) * lens[k], hipMemcpyHostToDevice, c->st));
    }
    for (int k = 0; k < nscal; ++k) g.s[k] = scalars[k];
    for (int r = 0; r < 2; ++r) { g.lo[r] = ranges[2 * r]; g.hi[r] = ranges[2 * r + 1]; }
    hipError_t e = launch_step_op(op, g, c->st);
    if (e != hipSuccess) return fdtd_fail(FDTD_E_HIP, std::string("op launch: ") + hipGetErrorString(e));
    for (int k = 0; k < o.nout; ++k)
        if (lens[k]) HIPCHK(hipMemcpyAsync(arrays[k], g.a[k], sizeof(double) * lens[k], hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    (void)total;
    return FDTD_OK;
}
