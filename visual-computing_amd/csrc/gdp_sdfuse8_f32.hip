// Streamed down-strip kernels of the finest level carrying the PCG update (the fused
// update-down kernel), 8 coarse rows per strip, element type float.
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_fused_crs<float, 8>(int nu, int blocks, hipStream_t s, const DownArgs& a) {
  down_fused_nu<float, 8>(nu, blocks, s, a);
}
}  // namespace gdp
