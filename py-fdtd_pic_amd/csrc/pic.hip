// Launchers of the particle kernels (pic_kernels.hpp).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "pic_kernels.hpp"

namespace fdtd {

// With profiling events the kernels go through hipExtLaunchKernelGGL, which
// stamps the events from their own dispatch packets.
template <typename R>
hipError_t launch_push(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* push_items, int npush,
                       long long launch_idx, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    if (npush <= 0) return hipSuccess;
    const int nb = (npush + kPushWaves - 1) / kPushWaves;    // items are padded to whole blocks
    if (e0 || e1)
        hipExtLaunchKernelGGL((pic_push_count<R>), dim3(nb), dim3(64 * kPushWaves), 0, st, e0, e1, 0, sims, beams,
                              push_items, launch_idx);
    else
        hipLaunchKernelGGL((pic_push_count<R>), dim3(nb), dim3(64 * kPushWaves), 0, st, sims, beams, push_items,
                           launch_idx);
    return hipGetLastError();
}

// The scatter reads nothing the field step of the same launch writes
// (H^{n-1/2} comes from the copy in hold, and both form J' from the CIC
// sums), so it may follow that step or run beside it on another stream.
template <typename R>
hipError_t launch_scatter(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* scatter_blocks,
                          int nscatter, long long launch_idx, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    if (nscatter <= 0) return hipSuccess;
    if (e0 || e1)
        hipExtLaunchKernelGGL((pic_scatter<R>), dim3(nscatter), dim3(256), 0, st, e0, e1, 0, sims, beams,
                              scatter_blocks, launch_idx);
    else
        hipLaunchKernelGGL((pic_scatter<R>), dim3(nscatter), dim3(256), 0, st, sims, beams, scatter_blocks,
                           launch_idx);
    return hipGetLastError();
}

template hipError_t launch_push<double>(const SimDev<double>*, const BeamDev<double>*, const int2*, int, long long,
                                        hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_push<float>(const SimDev<float>*, const BeamDev<float>*, const int2*, int, long long,
                                       hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_scatter<double>(const SimDev<double>*, const BeamDev<double>*, const int2*, int,
                                           long long, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_scatter<float>(const SimDev<float>*, const BeamDev<float>*, const int2*, int, long long,
                                          hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace fdtd
