// Launchers of the particle kernels (pic_kernels.hpp).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>

#include "pic_kernels.hpp"

namespace fdtd {

// With profiling events the kernels go through hipExtLaunchKernelGGL, which
// stamps the events from their own dispatch packets.
//
// Push/count of one step: the one-cell beams' kernel (one item per
// workgroup) and the multi-cell beams' kernel (one item per wave, padded to
// whole workgroups), e0 at the start of the first launch, e1 at the end of
// the last.
// xr: LDS rows per cell of the one-cell kernel (1 .. kPushCellMaxRows);
// single: the one-cell items' beam when they all belong to one (items
// (single, 0 .. ncell-1)), else -1.
template <typename R>
hipError_t launch_push(const SimDev<R>* sims, const BeamDev<R>* beams, const int2* cell_items, int ncell,
                       const int2* group_items, int ngroup, long long launch_idx, int xr, int single, hipStream_t st,
                       hipEvent_t e0, hipEvent_t e1) {
    const bool ev = e0 || e1;
    if (ncell > 0) {
        if (xr < 1 || xr > kPushCellMaxRows<R>) return hipErrorInvalidValue;
        const size_t shm = push_cell_lds_bytes<R>(xr);
        // dynamic shared memory above 64 KiB needs the function attribute; it
        // is raised to the largest size asked for so far (host threads with
        // contexts of their own may launch concurrently, hence the atomic)
        static std::atomic<size_t> allowed{64 * 1024};
        size_t cur = allowed.load();
        while (shm > cur) {
            const hipError_t a = hipFuncSetAttribute((const void*)pic_push_cell<R>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
            if (a != hipSuccess) return a;
            if (allowed.compare_exchange_weak(cur, shm)) break;
        }
        hipEvent_t a0 = e0, a1 = ngroup > 0 ? nullptr : e1;
        if (ev)
            hipExtLaunchKernelGGL((pic_push_cell<R>), dim3(ncell), dim3(64 * kPushWaves), shm, st, a0, a1, 0, sims,
                                  beams, cell_items, launch_idx, xr, single);
        else
            hipLaunchKernelGGL((pic_push_cell<R>), dim3(ncell), dim3(64 * kPushWaves), shm, st, sims, beams,
                               cell_items, launch_idx, xr, single);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (ngroup > 0) {
        const int nb = (ngroup + kPushWaves - 1) / kPushWaves;    // items are padded to whole blocks
        hipEvent_t a0 = ncell > 0 ? nullptr : e0, a1 = e1;
        if (ev)
            hipExtLaunchKernelGGL((pic_push_count<R>), dim3(nb), dim3(64 * kPushWaves), 0, st, a0, a1, 0, sims,
                                  beams, group_items, launch_idx);
        else
            hipLaunchKernelGGL((pic_push_count<R>), dim3(nb), dim3(64 * kPushWaves), 0, st, sims, beams,
                               group_items, launch_idx);
        return hipGetLastError();
    }
    return hipSuccess;
}

// The scatter (node and move blocks) reads nothing the field step of the
// same launch writes (both form J' from the CIC sums; the movers come from
// the particle buffers), so it may follow that step or share its launch.
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

template hipError_t launch_push<double>(const SimDev<double>*, const BeamDev<double>*, const int2*, int,
                                        const int2*, int, long long, int, int, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_push<float>(const SimDev<float>*, const BeamDev<float>*, const int2*, int, const int2*,
                                       int, long long, int, int, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_scatter<double>(const SimDev<double>*, const BeamDev<double>*, const int2*, int,
                                           long long, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_scatter<float>(const SimDev<float>*, const BeamDev<float>*, const int2*, int, long long,
                                          hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace fdtd
