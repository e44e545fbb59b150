// Streamed down-strip kernels (wavefront-order pre-smooth + residual + restrict),
// element type float.
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_stream<float>(int crs, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (crs == 8) down_stream_nu<float, 8>(nu, blocks, s, a);
  else if (crs == 16) down_stream_nu<float, 16>(nu, blocks, s, a);
  else launch_down_stream32<float>(nu, blocks, s, a);
}
}  // namespace gdp
