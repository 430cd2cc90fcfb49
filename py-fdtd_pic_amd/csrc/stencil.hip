// Field-step launcher (called from host.hip): dispatch on the slab kind to
// the instances compiled in stencil_<precision>_<kind>.hip.
#include "stencil_kernels.hpp"

namespace fdtd {

extern template hipError_t launch_kind<float, true, SLAB_DIEL>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<float, true, SLAB_LINEAR>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<float, true, SLAB_KERR_E>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_DIEL>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_LINEAR>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_KERR_E>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_KERR_REF>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);

// kind: 0 (no D-form slab: NONE or DIEL share the DIEL kernel), SLAB_LINEAR,
// SLAB_KERR_REF (FP64 only), SLAB_KERR_E.
template <typename R, bool KAHAN>
hipError_t launch_stencil(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                          long long gstep0, int nsteps, int K, int S, int kind, hipStream_t st,
                          hipEvent_t ev0, hipEvent_t ev1) {
    switch (kind) {
        case SLAB_NONE:
        case SLAB_DIEL: return launch_kind<R, KAHAN, SLAB_DIEL>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st, ev0, ev1);
        case SLAB_LINEAR: return launch_kind<R, KAHAN, SLAB_LINEAR>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st, ev0, ev1);
        case SLAB_KERR_E: return launch_kind<R, KAHAN, SLAB_KERR_E>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st, ev0, ev1);
        case SLAB_KERR_REF:
            if constexpr (sizeof(R) == 8)
                return launch_kind<R, KAHAN, SLAB_KERR_REF>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, st, ev0, ev1);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_stencil<double, false>(const SimDev<double>*, const int2*, int, long long,
                                                  long long, int, int, int, int, hipStream_t, hipEvent_t,
                                                  hipEvent_t);
template hipError_t launch_stencil<float, true>(const SimDev<float>*, const int2*, int, long long,
                                                long long, int, int, int, int, hipStream_t, hipEvent_t,
                                                hipEvent_t);

}  // namespace fdtd
