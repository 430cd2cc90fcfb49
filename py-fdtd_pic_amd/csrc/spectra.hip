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
// chunk of 64): the block stages the probe series through LDS in tiles, each thread owns one frequency and
// advances its twiddle by a complex multiply per sample, re-anchored with an
// exact sincos every 64 samples so the recurrence error stays ~1e-15.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

constexpr int kDftTile = 256;

__global__ __launch_bounds__(64) void probe_dft(const double* const* __restrict__ series, const int* __restrict__ lens,
                                                const double* __restrict__ freqs, int nf, double dt,
                                                double* __restrict__ out) {
    __shared__ double xs[kDftTile];
    const int s = blockIdx.y;
    const int f = blockIdx.x * 64 + threadIdx.x;
    const double* x = series[s];
    const int T = lens[s];
    const double w = f < nf ? -2.0 * M_PI * freqs[f] * dt : 0.0;
    double re = 0.0, im = 0.0;
    double cr = 1.0, ci = 0.0, wr, wi;
    sincos(w, &wi, &wr);
    for (int t0 = 0; t0 < T; t0 += kDftTile) {
        const int n = min(kDftTile, T - t0);
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += 64) xs[k] = x[t0 + k];
        __syncthreads();
        for (int k = 0; k < n; ++k) {
            if ((k & 63) == 0) sincos(w * (double)(t0 + k), &ci, &cr);
            const double v = xs[k];
            re = fma(v, cr, re);
            im = fma(v, ci, im);
            const double nr = cr * wr - ci * wi;
            ci = cr * wi + ci * wr;
            cr = nr;
        }
    }
    if (f < nf) {
        out[((size_t)s * nf + f) * 2] = re * dt;
        out[((size_t)s * nf + f) * 2 + 1] = im * dt;
    }
}

hipError_t launch_dft(const double* const* series, const int* lens, int nseries, const double* freqs, int nf,
                      double dt, double* out, hipStream_t st) {
    if (nseries <= 0 || nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(probe_dft, dim3((nf + 63) / 64, nseries), dim3(64), 0, st, series, lens, freqs, nf, dt, out);
    return hipGetLastError();
}

}  // namespace fdtd
