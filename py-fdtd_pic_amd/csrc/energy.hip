// FP64 discrete field energy of whole simulations, on the device.
//
// The Yee leapfrog in normalised units (E, H' = Z0 H) conserves
//     W^{n+1} = sum_n (E^{n+1}_n)^2 + sum_n H'^{n+3/2}_n H'^{n+1/2}_n
// in a closed lossless cavity.  The reference takes its per-step
// observables at Solver_Engine.py:360-370 and has no such accumulator; the
// step kernel stores H' before a launch's last H update (SimDev::Hprev) and
// this kernel forms W from it and the launch's final E and H, one workgroup
// per simulation, every term in FP64 (FP32 fields through their Kahan
// compensations), summed in a fixed order: thread t takes cells t, t + 256,
// ... in order, then a fixed LDS tree, so the value depends only on the
// state.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

constexpr int kEnergyJobs = 16;       // simulations per launch (the job table goes by value)

template <typename R>
struct EnergyJobs {
    EnergyJob<R> j[kEnergyJobs];
};

template <typename R>
__global__ __launch_bounds__(256) void energy_sum(EnergyJobs<R> jobs) {
#pragma clang fp contract(off)
    __shared__ double part[256];
    const EnergyJob<R>& jb = jobs.j[blockIdx.x];
    const int t = (int)threadIdx.x;
    double acc = 0.0;
    for (int n = t; n <= jb.N; n += 256) {
        const double e = jb.cE ? (double)jb.E[n] - (double)jb.cE[n] : (double)jb.E[n];
        const double h = jb.cH ? (double)jb.H[n] - (double)jb.cH[n] : (double)jb.H[n];
        acc += e * e + h * jb.Hprev[n];
    }
    part[t] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) part[t] += part[t + w];
        __syncthreads();
    }
    if (t == 0) *jb.out = part[0];
}

template <typename R>
hipError_t launch_energy(const EnergyJob<R>* jobs, int n, hipStream_t st) {
    for (int k0 = 0; k0 < n; k0 += kEnergyJobs) {
        EnergyJobs<R> js;
        const int nk = n - k0 < kEnergyJobs ? n - k0 : kEnergyJobs;
        for (int k = 0; k < nk; ++k) js.j[k] = jobs[k0 + k];
        hipLaunchKernelGGL((energy_sum<R>), dim3(nk), dim3(256), 0, st, js);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template hipError_t launch_energy<double>(const EnergyJob<double>*, int, hipStream_t);
template hipError_t launch_energy<float>(const EnergyJob<float>*, int, hipStream_t);

}  // namespace fdtd
