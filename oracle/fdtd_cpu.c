/*
 * Oracle: C restatement of the physical FDTD step loop.  TEST INFRASTRUCTURE ONLY.
 *
 * Same model and step order as oracle/physical.py (which restates
 * Solver_Engine.py:294-316 with the physics of the benchmark configs), in
 * FP64, built with -O3 -ffp-contract=off.  It is the CPU baseline that
 * bench.py times ("kind": "port") and a fast oracle for grids the NumPy
 * oracle cannot run in seconds.  Only tests/ and bench.py's cpu_baseline
 * leg load it.
 *
 * Units (normalised): E, H' = Z0 H, D' = D/eps0, P' = P/eps0, J' = Z0 Jx_ref.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t N, steps;
    double Sc;
    /* CPML: E cells [e_l0,e_l1) U [e_r0,e_r1), H cells [h_l0,h_l1) U [h_r0,h_r1) */
    int32_t e_l0, e_l1, e_r0, e_r1, h_l0, h_l1, h_r0, h_r1;
    const double *bE, *cE, *bH, *cH;        /* length N+1 */
    int32_t F, R;                            /* slab [F,R) */
    const double *A, *B, *C, *eps;          /* length R-F */
    double chi3;
    int32_t src_cell;
    const double *src;                       /* length steps */
    int32_t nprobe;
    const int32_t *probes;
    double *probe_out;                       /* nprobe x steps */
    const double *J;                         /* optional, length (N+1)*steps or NULL */
} cpu_phys_t;

typedef struct {
    double *E, *H, *psiE, *psiH;            /* length N+1 */
    double *D, *P, *Pm;                     /* length R-F */
} cpu_state_t;

/* Real root of chi3 E^3 + eps E = u continuing u/eps (chi3 >= 0). */
static double kerr_root(double u, double eps, double chi3)
{
    if (chi3 == 0.0 || u == 0.0) return u / eps;
    double p = eps / chi3, q = -u / chi3;
    double s = sqrt(p / 3.0);
    double arg = (1.5 * q / p) / s;
    return -2.0 * s * sinh(asinh(arg) / 3.0);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void one_step(const cpu_phys_t *m, cpu_state_t *s, int t, double *dH, int nth)
{
    const int N = m->N;
    const double Sc = m->Sc;
    const int F = m->F, R = m->R;
    double *E = s->E, *H = s->H;
    const double *J = m->J ? m->J + (size_t)t * (N + 1) : NULL;
    (void)nth;
    /* polarisation with E at the start of the step (BaseFDTD11.py:631) */
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int i = 0; i < R - F; ++i) {
        double Pn = m->A[i] * s->P[i] + m->B[i] * s->Pm[i] + m->C[i] * E[F + i];
        s->Pm[i] = s->P[i];
        s->P[i] = Pn;
    }
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int n = 1; n < N; ++n) {
        dH[n] = H[n] - H[n - 1];
        double j = J ? J[n] : 0.0;
        E[n] = E[n] + Sc * (dH[n] - j);
    }
    for (int n = m->e_l0; n < m->e_l1; ++n) {
        s->psiE[n] = m->bE[n] * s->psiE[n] + m->cE[n] * dH[n];
        E[n] += Sc * s->psiE[n];
    }
    for (int n = m->e_r0; n < m->e_r1; ++n) {
        s->psiE[n] = m->bE[n] * s->psiE[n] + m->cE[n] * dH[n];
        E[n] += Sc * s->psiE[n];
    }
    E[m->src_cell] += m->src[t];
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int i = 0; i < R - F; ++i) {
        int n = F + i;
        double j = J ? J[n] : 0.0;
        s->D[i] = s->D[i] + Sc * (dH[n] - j);
        E[n] = kerr_root(s->D[i] - s->P[i], m->eps[i], m->chi3);
    }
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int n = 0; n < N; ++n) {
        dH[n] = E[n + 1] - E[n];   /* reuse as dE */
        H[n] = H[n] + Sc * dH[n];
    }
    for (int n = m->h_l0; n < m->h_l1; ++n) {
        s->psiH[n] = m->bH[n] * s->psiH[n] + m->cH[n] * dH[n];
        H[n] += Sc * s->psiH[n];
    }
    for (int n = m->h_r0; n < m->h_r1; ++n) {
        s->psiH[n] = m->bH[n] * s->psiH[n] + m->cH[n] * dH[n];
        H[n] += Sc * s->psiH[n];
    }
}

/* Runs steps [t0, t1); returns wall seconds of the step loop. */
double cpu_phys_run(const cpu_phys_t *m, cpu_state_t *s, int t0, int t1, double *scratch, int nthreads)
{
    double start = now_s();
    for (int t = t0; t < t1; ++t) {
        one_step(m, s, t, scratch, nthreads);
        for (int k = 0; k < m->nprobe; ++k)
            m->probe_out[(size_t)k * m->steps + t] = s->E[m->probes[k]];
    }
    return now_s() - start;
}

/* Independent instances (batched sweep): OpenMP over instances. */
double cpu_phys_run_many(int count, const cpu_phys_t *const *ms, cpu_state_t *const *ss,
                         double *const *scratch, int nthreads)
{
    double start = now_s();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) if (nthreads > 1)
    for (int i = 0; i < count; ++i) {
        const cpu_phys_t *m = ms[i];
        for (int t = 0; t < m->steps; ++t) {
            one_step(m, ss[i], t, scratch[i], 1);
            for (int k = 0; k < m->nprobe; ++k)
                m->probe_out[(size_t)k * m->steps + t] = ss[i]->E[m->probes[k]];
        }
    }
    return now_s() - start;
}

int cpu_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
