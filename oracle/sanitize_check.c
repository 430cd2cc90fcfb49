/*
 * Sanitizer check of the compiled CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * tests/test_oracle_sanitizers.py builds this file with gcc
 * -fsanitize=address,undefined -fno-sanitize-recover=all -fopenmp and runs
 * it.  It includes oracle/fdtd_cpu.c itself (so it checks that source, with
 * its own structure definitions), runs a small model (CPML strips at both
 * ends, a Lorentz + Kerr slab, a soft source, two probes) for 300 steps on one
 * thread and on four, and a particle beam for 200 steps the same way, and
 * checks that the field runs agree bit for bit, the particle runs to 1e-12
 * (the four-thread deposit sums per-thread partials), and every value is
 * finite.  Prints "ok" and exits 0; any sanitizer report aborts.
 */
#include "fdtd_cpu.c"

#include <stdio.h>
#include <stdlib.h>

enum { N = 400, T = 300, TP = 200, F = 150, R = 250, W = 20, NP = 600 };

static double *zeros(size_t n) { return (double *)calloc(n, sizeof(double)); }

typedef struct {
    cpu_state_t s;
    double *E, *H, *psiE, *psiH, *D, *P, *Pm, *probe_out;
} run_t;

static void alloc_run(run_t *r)
{
    r->E = zeros(N + 1); r->H = zeros(N + 1); r->psiE = zeros(N + 1); r->psiH = zeros(N + 1);
    r->D = zeros(R - F); r->P = zeros(R - F); r->Pm = zeros(R - F);
    r->probe_out = zeros(2 * (size_t)(T > TP ? T : TP));
    cpu_state_t s = {r->E, r->H, r->psiE, r->psiH, r->D, r->P, r->Pm};
    r->s = s;
}

static void free_run(run_t *r)
{
    free(r->E); free(r->H); free(r->psiE); free(r->psiH); free(r->D); free(r->P); free(r->Pm);
    free(r->probe_out);
}

/* a equals b bitwise (tol = 0), or within tol of the largest |a| of the array */
static int same(const double *a, const double *b, size_t n, double tol)
{
    double scale = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (!isfinite(a[i]) || !isfinite(b[i])) return 0;
        scale = fmax(scale, fabs(a[i]));
    }
    for (size_t i = 0; i < n; ++i)
        if (tol == 0.0 ? a[i] != b[i] : fabs(a[i] - b[i]) > tol * scale) return 0;
    return 1;
}

int main(void)
{
    double *bE = zeros(N + 1), *cE = zeros(N + 1), *bH = zeros(N + 1), *cH = zeros(N + 1);
    for (int n = 0; n <= N; ++n) { bE[n] = bH[n] = 0.9; cE[n] = cH[n] = -0.05; }
    double *A = zeros(R - F), *B = zeros(R - F), *Cc = zeros(R - F), *eps = zeros(R - F);
    for (int i = 0; i < R - F; ++i) { A[i] = 1.9; B[i] = -0.95; Cc[i] = 0.01; eps[i] = 1.5 + 0.002 * i; }
    double *src = zeros(T);
    for (int t = 0; t < T; ++t) src[t] = 2.0 * exp(-0.5 * ((t - 60.0) / 15.0) * ((t - 60.0) / 15.0)) * sin(0.3 * t);
    int32_t probes[2] = {60, 340};
    run_t r1, r4;
    alloc_run(&r1);
    alloc_run(&r4);
    cpu_phys_t m = {N, T, 0.5, 1, W, N + 1 - W, N, 0, W, N - W, N, bE, cE, bH, cH, F, R, A, B, Cc, eps,
                    1e-3, 100, src, 2, probes, r1.probe_out, NULL};
    double *scratch = zeros(N + 1);
    cpu_phys_run(&m, &r1.s, 0, T, scratch, 1);
    m.probe_out = r4.probe_out;
    cpu_phys_run(&m, &r4.s, 0, T, scratch, 4);
    int ok = same(r1.E, r4.E, N + 1, 0.0) && same(r1.H, r4.H, N + 1, 0.0) && same(r1.psiE, r4.psiE, N + 1, 0.0) &&
             same(r1.psiH, r4.psiH, N + 1, 0.0) && same(r1.D, r4.D, R - F, 0.0) && same(r1.P, r4.P, R - F, 0.0) &&
             same(r1.Pm, r4.Pm, R - F, 0.0) && same(r1.probe_out, r4.probe_out, 2 * (size_t)T, 0.0);
    if (!ok) { fprintf(stderr, "field runs differ\n"); return 1; }

    /* a particle beam drifting through the grid */
    run_t q1, q4;
    alloc_run(&q1);
    alloc_run(&q4);
    double *x[2], *u[2], *Hprev[2], *J[2], *B1[2], *B2[2], *w0[2], *w1[2];
    int8_t *alive[2];
    int32_t *m0[2];
    cpu_beam_t b[2];
    for (int k = 0; k < 2; ++k) {
        const int nth = k == 0 ? 1 : 4;
        x[k] = zeros(NP); u[k] = zeros(3 * NP); Hprev[k] = zeros(N + 1); J[k] = zeros(N + 1);
        B1[k] = zeros((size_t)nth * (N + 1)); B2[k] = zeros((size_t)nth * (N + 1));
        w0[k] = zeros(NP); w1[k] = zeros(NP);
        alive[k] = (int8_t *)calloc(NP, 1);
        m0[k] = (int32_t *)calloc(NP, sizeof(int32_t));
        for (int i = 0; i < NP; ++i) {
            x[k][i] = 100.0 + 40.0 * (i + 0.5) / NP;
            u[k][3 * i] = 0.1 * ((i % 7) - 3) / 3.0;
            u[k][3 * i + 2] = 0.1 + 0.01 * (i % 5);
            alive[k][i] = 1;
        }
        cpu_beam_t bb = {NP, x[k], u[k], alive[k], 0.01, 1e-3, Hprev[k], J[k], B1[k], B2[k], m0[k], w0[k], w1[k]};
        b[k] = bb;
    }
    m.steps = TP;
    m.probe_out = q1.probe_out;
    cpu_pic_run(&m, &q1.s, &b[0], 0, TP, scratch, 1);
    m.probe_out = q4.probe_out;
    cpu_pic_run(&m, &q4.s, &b[1], 0, TP, scratch, 4);
    ok = same(q1.E, q4.E, N + 1, 1e-12) && same(q1.H, q4.H, N + 1, 1e-12) && same(x[0], x[1], NP, 1e-12) &&
         same(u[0], u[1], 3 * NP, 1e-12) && same(J[0], J[1], N + 1, 1e-12);
    /* the runs did something: the pulse reached the probe, the beam deposited */
    double emax = 0.0, jmax = 0.0;
    for (int n = 0; n <= N; ++n) { emax = fmax(emax, fabs(r1.E[n])); jmax = fmax(jmax, fabs(J[0][n])); }
    ok = ok && emax > 0.0 && jmax > 0.0;
    for (int k = 0; k < 2; ++k) {
        free(x[k]); free(u[k]); free(Hprev[k]); free(J[k]); free(B1[k]); free(B2[k]); free(w0[k]); free(w1[k]);
        free(alive[k]); free(m0[k]);
    }
    free_run(&r1); free_run(&r4); free_run(&q1); free_run(&q4);
    free(bE); free(cE); free(bH); free(cH); free(A); free(B); free(Cc); free(eps); free(src); free(scratch);
    if (!ok) { fprintf(stderr, "particle runs differ\n"); return 1; }
    printf("ok\n");
    return 0;
}
