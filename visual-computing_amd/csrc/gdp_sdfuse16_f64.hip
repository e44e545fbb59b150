// Streamed down-strip kernels of the finest level carrying the PCG update (the fused
// update-down kernel), 16 coarse rows per strip, element type double.
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_fused_crs<double, 16>(int nu, int blocks, hipStream_t s, const DownArgs& a) {
  down_fused_nu<double, 16>(nu, blocks, s, a);
}
}  // namespace gdp
