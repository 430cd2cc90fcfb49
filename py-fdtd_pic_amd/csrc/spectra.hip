// FP64 DFT of probe series on the device: X(f) = sum_t x_t exp(-2 pi i f t dt) dt.
//
// The reference's spectral stage (TransformHandler.RefTester,
// TransformHandler.py:34-78, on x1ColBe / x1ColAf from the integrators) runs
// scipy FFTs on the host; reflection and transmission spectra at chosen
// frequencies (R = |X1 - X0| / |X0| at the reflection probe, T = |X1| / |X0|
// at the transmission probe, the two-pass protocol of
// Solver_Engine.py:275-371) need only these per-frequency sums.
//
// The series are FP64 in both precisions (the step kernel writes an FP32
// field as hi - c of its Kahan pair).  One block per (simulation, frequency
// chunk of 64): each of its kDftWaves waves takes every kDftWaves-th tile
// of 64 samples, stages the tile through LDS and, one thread per frequency,
// starts the twiddle from an exact sincos at the tile's first sample and
// advances it by a complex multiply per sample (the recurrence error stays
// ~1e-15 over 64 samples).  The waves' partial sums are added in wave order,
// so the result does not depend on the launch.  A thread holding a whole
// series would run T samples in one dependent chain; the block's waves cut it
// kDftWaves-fold.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

constexpr int kDftWaves = 16;

__global__ __launch_bounds__(64 * kDftWaves) void probe_dft(const double* const* __restrict__ series,
                                                            const int* __restrict__ lens,
                                                            const double* __restrict__ freqs, int nf, double dt,
                                                            double* __restrict__ out) {
    __shared__ double xs[kDftWaves][64];
    __shared__ double pre[kDftWaves][64], pim[kDftWaves][64];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int s = blockIdx.y;
    const int f = blockIdx.x * 64 + lane;
    const double* x = series[s];
    const int T = lens[s];
    const double w = f < nf ? -2.0 * M_PI * freqs[f] * dt : 0.0;
    double wr, wi;
    sincos(w, &wi, &wr);
    double re = 0.0, im = 0.0;
    const int rounds = ((T + 63) / 64 + kDftWaves - 1) / kDftWaves;   // the same for every wave
    for (int r = 0; r < rounds; ++r) {
        const int t0 = 64 * (r * kDftWaves + wv);
        const int n = max(0, min(64, T - t0));
        __syncthreads();
        if (lane < n) xs[wv][lane] = x[t0 + lane];
        __syncthreads();
        if (n > 0) {
            double cr, ci;
            sincos(w * (double)t0, &ci, &cr);
            for (int k = 0; k < n; ++k) {
                const double v = xs[wv][k];
                re = fma(v, cr, re);
                im = fma(v, ci, im);
                const double nr = cr * wr - ci * wi;
                ci = cr * wi + ci * wr;
                cr = nr;
            }
        }
    }
    pre[wv][lane] = re;
    pim[wv][lane] = im;
    __syncthreads();
    if (wv == 0 && f < nf) {
        double sr = pre[0][lane], si = pim[0][lane];
        for (int q = 1; q < kDftWaves; ++q) {
            sr += pre[q][lane];
            si += pim[q][lane];
        }
        out[((size_t)s * nf + f) * 2] = sr * dt;
        out[((size_t)s * nf + f) * 2 + 1] = si * dt;
    }
}

hipError_t launch_dft(const double* const* series, const int* lens, int nseries, const double* freqs, int nf,
                      double dt, double* out, hipStream_t st) {
    if (nseries <= 0 || nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(probe_dft, dim3((nf + 63) / 64, nseries), dim3(64 * kDftWaves), 0, st, series, lens, freqs,
                       nf, dt, out);
    return hipGetLastError();
}

}  // namespace fdtd
