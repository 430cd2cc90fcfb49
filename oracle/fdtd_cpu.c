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

static void one_step_j(const cpu_phys_t *m, cpu_state_t *s, int t, double *dH, int nth, const double *J);

static void one_step(const cpu_phys_t *m, cpu_state_t *s, int t, double *dH, int nth)
{
    one_step_j(m, s, t, dH, nth, m->J ? m->J + (size_t)t * (m->N + 1) : NULL);
}

static void one_step_j(const cpu_phys_t *m, cpu_state_t *s, int t, double *dH, int nth, const double *J)
{
    const int N = m->N;
    const double Sc = m->Sc;
    const int F = m->F, R = m->R;
    double *E = s->E, *H = s->H;
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

/*
 * Particles (oracle/pic.py restated): relativistic Boris push with linear
 * gather of E^n and the time-averaged H', CIC deposit of J' at x^{n+1/2},
 * absorption into cells 1 / N-2.  The push runs in parallel over particles.
 * With one thread the deposit sums each node in particle order, as numpy's
 * bincount does (two sums, J = (0 + B1) + B2), so the run is bitwise
 * oracle/pic.py's; with more threads each thread sums a contiguous range of
 * particles and the partials are added in thread order (deterministic for a
 * given thread count).
 */
typedef struct {
    int64_t n;
    double *x, *u;                           /* n, n x 3 */
    int8_t *alive;
    double a, jscale;
    double *Hprev;                           /* N+1: H' before the last field step */
    double *J;                               /* N+1 */
    double *B1, *B2;                         /* nthreads x (N+1) partial sums */
    int32_t *m0;                             /* n: node of the first CIC weight (-1: none) */
    double *w0, *w1;                         /* n: the two weights */
} cpu_beam_t;

static void push_particles(const cpu_phys_t *m, const cpu_state_t *s, cpu_beam_t *b, int nth)
{
    const int N = m->N;
    const double Sc = m->Sc, a = b->a;
    const double *E = s->E, *H = s->H, *Hp = b->Hprev;
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
    for (int64_t i = 0; i < b->n; ++i) {
        if (!b->alive[i]) { b->m0[i] = -1; continue; }
        const double x = b->x[i];
        const double fi = floor(x);
        const int64_t ci = (int64_t)fi;
        const double f = x - fi;
        const double Eg = (1.0 - f) * E[ci] + f * E[ci + 1];
        const double xh = x - 0.5;
        const double fj = floor(xh);
        const int64_t cj = (int64_t)fj;
        const double g = xh - fj;
        const double h0 = 0.5 * (Hp[cj] + H[cj]), h1 = 0.5 * (Hp[cj + 1] + H[cj + 1]);
        const double Hg = (1.0 - g) * h0 + g * h1;
        double *u = b->u + 3 * i;
        const double um0 = u[0] + a * Eg, um1 = u[1], um2 = u[2];
        const double gm = sqrt(1.0 + ((um0 * um0 + um1 * um1) + um2 * um2));
        const double ty = -a * Hg / gm;
        const double sy = 2.0 * ty / (1.0 + ty * ty);
        const double upx = um0 - um2 * ty;
        const double upz = um2 + um0 * ty;
        const double un0 = (um0 - upz * sy) + a * Eg;
        const double un2 = um2 + upx * sy;
        u[0] = un0;
        u[2] = un2;
        const double gam = sqrt(1.0 + ((un0 * un0 + um1 * um1) + un2 * un2));
        const double bx = un0 / gam, bz = un2 / gam;
        const double xd = x + 0.5 * Sc * bz;
        double xn = x + Sc * bz;
        const double fm = floor(xd);
        const double fh = xd - fm;
        const double q = b->jscale * bx;
        b->m0[i] = (int32_t)fm;
        b->w0[i] = q * (1.0 - fh);
        b->w1[i] = q * fh;
        const double cell = floor(xn);
        if (cell < 2.0) { xn = 1.5; b->alive[i] = 0; }
        else if (cell > N - 3) { xn = (double)(N - 2) + 0.5; b->alive[i] = 0; }
        b->x[i] = xn;
    }
}

static void deposit(const cpu_phys_t *m, cpu_beam_t *b, int nth)
{
    const int n1 = m->N + 1;
    if (nth <= 1) {
        memset(b->B1, 0, sizeof(double) * n1);
        memset(b->B2, 0, sizeof(double) * n1);
        for (int64_t i = 0; i < b->n; ++i) {
            if (b->m0[i] < 0) continue;
            b->B1[b->m0[i]] += b->w0[i];
            b->B2[b->m0[i] + 1] += b->w1[i];
        }
        for (int k = 0; k < n1; ++k) b->J[k] = (0.0 + b->B1[k]) + b->B2[k];
        return;
    }
#pragma omp parallel num_threads(nth)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        double *B1 = b->B1 + (size_t)t * n1, *B2 = b->B2 + (size_t)t * n1;
        memset(B1, 0, sizeof(double) * n1);
        memset(B2, 0, sizeof(double) * n1);
        const int64_t lo = b->n * t / nt, hi = b->n * (t + 1) / nt;
        for (int64_t i = lo; i < hi; ++i) {
            if (b->m0[i] < 0) continue;
            B1[b->m0[i]] += b->w0[i];
            B2[b->m0[i] + 1] += b->w1[i];
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int k = 0; k < n1; ++k) {
            double s1 = 0.0, s2 = 0.0;
            for (int q = 0; q < nt; ++q) { s1 += b->B1[(size_t)q * n1 + k]; s2 += b->B2[(size_t)q * n1 + k]; }
            b->J[k] = (0.0 + s1) + s2;
        }
    }
}

/* Coupled run over steps [t0, t1): push/deposit, then the field step with J'.
 * Returns wall seconds. */
double cpu_pic_run(const cpu_phys_t *m, cpu_state_t *s, cpu_beam_t *b, int t0, int t1, double *scratch,
                   int nthreads)
{
    const int n1 = m->N + 1;
    double start = now_s();
    for (int t = t0; t < t1; ++t) {
        push_particles(m, s, b, nthreads);
        deposit(m, b, nthreads);
        memcpy(b->Hprev, s->H, sizeof(double) * n1);
        one_step_j(m, s, t, scratch, nthreads, b->J);
        for (int k = 0; k < m->nprobe; ++k)
            m->probe_out[(size_t)k * m->steps + t] = s->E[m->probes[k]];
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
