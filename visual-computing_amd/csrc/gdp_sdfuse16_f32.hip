// Streamed down-strip kernels of the finest level carrying the PCG update (the fused
// update-down kernel), 16 coarse rows per strip, element type float.
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_fused_crs<float, 16>(int nu, int blocks, hipStream_t s, const DownArgs& a) {
  down_fused_nu<float, 16>(nu, blocks, s, a);
}
}  // namespace gdp
