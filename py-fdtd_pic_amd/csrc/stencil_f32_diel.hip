// Field-step kernels, alone and fused with the particle scatter, of one
// precision and slab kind (one translation unit each, so the instances
// compile in parallel).
#include "field_scatter.hpp"

namespace fdtd {
template hipError_t launch_kind<float, true, SLAB_DIEL>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const PersistArgs&, hipStream_t, hipEvent_t, hipEvent_t);
template int field_blocks_per_cu<float, true, SLAB_DIEL>(int);
template hipError_t launch_fused_kind<float, true, SLAB_DIEL>(const SimDev<float>*, const int2*, int, long long, long long, int, int, int, const BeamDev<float>*, const int2*, int, int, hipStream_t, hipEvent_t, hipEvent_t);
}  // namespace fdtd
