// Field-step kernels, alone and fused with the particle scatter, of one
// precision and slab kind (one translation unit each, so the instances
// compile in parallel).
#include "field_scatter.hpp"

namespace fdtd {
template hipError_t launch_kind<double, false, SLAB_DIEL>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
template int field_blocks_per_cu<double, false, SLAB_DIEL>(int);
template hipError_t launch_fused_kind<double, false, SLAB_DIEL>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, const BeamDev<double>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
}  // namespace fdtd
