// Streamed down-strip kernels (wavefront-order pre-smooth + residual + restrict),
// element type double.
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_stream<double>(int crs, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (crs == 8) down_stream_nu<double, 8>(nu, blocks, s, a);
  else if (crs == 16) down_stream_nu<double, 16>(nu, blocks, s, a);
  else launch_down_stream32<double>(nu, blocks, s, a);
}
}  // namespace gdp
