// One launch for the field step and the particle scatter of a PIC step.
//
// The scatter's node blocks read nothing the field step writes (the field
// step forms J' from the CIC sums itself), so both can share one grid of
// 256-thread workgroups: the node blocks first (few, with short dependent
// load chains that would otherwise trail the launch), then the field
// workgroups, four tiles each, one per wave.  One launch instead of two
// saves the second launch's ramp and drain.  The host fuses only small field
// launches (kFuseMaxTiles): the kernel has the field tile's register count.
#pragma once
#include "pic_kernels.hpp"
#include "stencil_kernels.hpp"

namespace fdtd {

constexpr int kFusedTileWaves = 4;    // a 256-thread workgroup: four field tiles or one scatter block

template <typename R, int KW, bool KAHAN, int KIND>
__global__ __launch_bounds__(256) void field_scatter(const SimDev<R>* __restrict__ sims,
                                                     const int2* __restrict__ tiles, int ntiles,
                                                     long long launch_idx, long long gstep0, int nsteps, int S,
                                                     const BeamDev<R>* __restrict__ beams,
                                                     const int2* __restrict__ sblocks, int nfirst) {
    const int nfb = (ntiles + kFusedTileWaves - 1) / kFusedTileWaves;
    const int b = (int)blockIdx.x;
    if (b < nfirst || b >= nfirst + nfb) {
        scatter_block<R>(sims, beams, sblocks[b < nfirst ? b : b - nfb], launch_idx, (int)threadIdx.x);
        return;
    }
    // wave-uniform tile index, so the descriptors come in through scalar loads
    const int tid = kFusedTileWaves * (b - nfirst) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (tid >= ntiles) return;
    const PersistArgs one{1, nullptr, nullptr, nullptr, 0};
    field_tile<R, KW, KAHAN, KIND>(sims, tiles, tid, launch_idx, gstep0, nsteps, S, one);
}

template <typename R, int K, bool KAHAN, int KIND>
static hipError_t launch_fused_k(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                                 long long gstep0, int nsteps, int S, const BeamDev<R>* beams,
                                 const int2* sblocks, int nsblocks, int nfirst, hipStream_t st, hipEvent_t ev0,
                                 hipEvent_t ev1) {
    const int grid = (ntiles + kFusedTileWaves - 1) / kFusedTileWaves + nsblocks;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL((field_scatter<R, K, KAHAN, KIND>), dim3(grid), dim3(256), 0, st, ev0, ev1, 0,
                              sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, beams, sblocks, nfirst);
    else
        hipLaunchKernelGGL((field_scatter<R, K, KAHAN, KIND>), dim3(grid), dim3(256), 0, st, sims, tiles,
                           ntiles, launch_idx, gstep0, nsteps, S, beams, sblocks, nfirst);
    return hipGetLastError();
}

// K: cells per lane of the field tiles (1 or 2; a PIC context steps once per
// launch, where wider tiles do not pay).  sblocks: nfirst node blocks, then
// the rest.
template <typename R, bool KAHAN, int KIND>
hipError_t launch_fused_kind(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                             long long gstep0, int nsteps, int K, int S, const BeamDev<R>* beams,
                             const int2* sblocks, int nsblocks, int nfirst, hipStream_t st, hipEvent_t ev0,
                             hipEvent_t ev1) {
    switch (K) {
        case 1: return launch_fused_k<R, 1, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, beams, sblocks, nsblocks, nfirst, st, ev0, ev1);
        case 2: return launch_fused_k<R, 2, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, S, beams, sblocks, nsblocks, nfirst, st, ev0, ev1);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fdtd
