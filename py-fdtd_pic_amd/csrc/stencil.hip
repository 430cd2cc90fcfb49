// Field-step launchers (called from host.hip): dispatch on the slab kind to
// the instances compiled in stencil_<precision>_<kind>.hip.
#include "field_scatter.hpp"

namespace fdtd {

extern template hipError_t launch_kind<float, true, SLAB_DIEL>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<float, true, SLAB_LINEAR>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<float, true, SLAB_KERR_E>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_DIEL>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_LINEAR>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_KERR_E>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_kind<double, false, SLAB_KERR_REF>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);

extern template int field_blocks_per_cu<float, true, SLAB_DIEL>(int);
extern template int field_blocks_per_cu<float, true, SLAB_LINEAR>(int);
extern template int field_blocks_per_cu<float, true, SLAB_KERR_E>(int);
extern template int field_blocks_per_cu<double, false, SLAB_DIEL>(int);
extern template int field_blocks_per_cu<double, false, SLAB_LINEAR>(int);
extern template int field_blocks_per_cu<double, false, SLAB_KERR_E>(int);
extern template int field_blocks_per_cu<double, false, SLAB_KERR_REF>(int);

extern template hipError_t launch_fused_kind<float, true, SLAB_DIEL>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const BeamDev<float>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<float, true, SLAB_LINEAR>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const BeamDev<float>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<float, true, SLAB_KERR_E>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const BeamDev<float>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<double, false, SLAB_DIEL>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const BeamDev<double>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<double, false, SLAB_LINEAR>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const BeamDev<double>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<double, false, SLAB_KERR_E>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const BeamDev<double>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template hipError_t launch_fused_kind<double, false, SLAB_KERR_REF>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const BeamDev<double>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);

// kind: 0 (no D-form slab: NONE or DIEL share the DIEL kernel), SLAB_LINEAR,
// SLAB_KERR_REF (FP64 only), SLAB_KERR_E.
template <typename R, bool KAHAN>
hipError_t launch_stencil(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                          long long gstep0, int nsteps, int K, int S, int kind, const PersistArgs& pa,
                          hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
#define FDTD_FIELD(KIND) \
    launch_kind<R, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, pa, st, ev0, ev1)
    switch (kind) {
        case SLAB_NONE:
        case SLAB_DIEL: return FDTD_FIELD(SLAB_DIEL);
        case SLAB_LINEAR: return FDTD_FIELD(SLAB_LINEAR);
        case SLAB_KERR_E: return FDTD_FIELD(SLAB_KERR_E);
        case SLAB_KERR_REF:
            if constexpr (sizeof(R) == 8) return FDTD_FIELD(SLAB_KERR_REF);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
#undef FDTD_FIELD
}

template hipError_t launch_stencil<double, false>(const SimDev<double>*, const int2*, int, long long,
                                                  long long, int, int, int, int, const PersistArgs&, hipStream_t,
                                                  hipEvent_t, hipEvent_t);
template hipError_t launch_stencil<float, true>(const SimDev<float>*, const int2*, int, long long,
                                                long long, int, int, int, int, const PersistArgs&, hipStream_t,
                                                hipEvent_t, hipEvent_t);

// Field workgroups one CU holds at once, for a slab kind and K.
template <typename R, bool KAHAN>
int field_capacity_per_cu(int K, int kind) {
    switch (kind) {
        case SLAB_NONE:
        case SLAB_DIEL: return field_blocks_per_cu<R, KAHAN, SLAB_DIEL>(K);
        case SLAB_LINEAR: return field_blocks_per_cu<R, KAHAN, SLAB_LINEAR>(K);
        case SLAB_KERR_E: return field_blocks_per_cu<R, KAHAN, SLAB_KERR_E>(K);
        case SLAB_KERR_REF:
            if constexpr (sizeof(R) == 8) return field_blocks_per_cu<R, KAHAN, SLAB_KERR_REF>(K);
            return 0;
        default: return 0;
    }
}

template int field_capacity_per_cu<double, false>(int, int);
template int field_capacity_per_cu<float, true>(int, int);

// The field step of one tile group and the particle scatter in one launch
// (field_scatter.hpp); same kinds as launch_stencil.
template <typename R, bool KAHAN>
hipError_t launch_field_scatter(const SimDev<R>* sims, const int2* tiles, int ntiles, long long launch_idx,
                                long long gstep0, int nsteps, int K, int S, int kind, const BeamDev<R>* beams,
                                const int2* sblocks, int nsblocks, int nfirst, hipStream_t st, hipEvent_t ev0,
                                hipEvent_t ev1) {
#define FDTD_FUSED(KIND) \
    launch_fused_kind<R, KAHAN, KIND>(sims, tiles, ntiles, launch_idx, gstep0, nsteps, K, S, beams, sblocks, nsblocks, \
                                      nfirst, st, ev0, ev1)
    switch (kind) {
        case SLAB_NONE:
        case SLAB_DIEL: return FDTD_FUSED(SLAB_DIEL);
        case SLAB_LINEAR: return FDTD_FUSED(SLAB_LINEAR);
        case SLAB_KERR_E: return FDTD_FUSED(SLAB_KERR_E);
        case SLAB_KERR_REF:
            if constexpr (sizeof(R) == 8) return FDTD_FUSED(SLAB_KERR_REF);
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
#undef FDTD_FUSED
}

template hipError_t launch_field_scatter<double, false>(const SimDev<double>*, const int2*, int, long long, long long,
                                                        int, int, int, int, const BeamDev<double>*, const int2*, int,
                                                        int, hipStream_t, hipEvent_t, hipEvent_t);
template hipError_t launch_field_scatter<float, true>(const SimDev<float>*, const int2*, int, long long, long long,
                                                      int, int, int, int, const BeamDev<float>*, const int2*, int,
                                                      int, hipStream_t, hipEvent_t, hipEvent_t);

}  // namespace fdtd
