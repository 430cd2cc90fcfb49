// Field-step kernels of one precision and slab kind (one translation unit
// each, so the instances compile in parallel).
#include "stencil_kernels.hpp"

namespace fdtd {
template hipError_t launch_kind<double, false, SLAB_KERR_E>(const SimDev<double>*, const int2*, int, long long, long long, int, int, int, hipStream_t, hipEvent_t, hipEvent_t);
}  // namespace fdtd
