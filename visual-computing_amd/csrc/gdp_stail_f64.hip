// Register-strip coarse-tail kernels, element type double: the general kernel and the
// dispatch of every tail launch.
#include "gdp_stail_launch.h"

namespace gdp {
template <>
void launch_stail<double>(const STailSpec& sp, bool dots, int blocks, int smem, hipStream_t s,
                        const TailArgs& a) {
  stail_launch<double>(sp, dots, blocks, smem, s, a);
}
template <>
hipError_t stail_set_smem<double>(int smem) {
  return stail_smem<double>(smem);
}
}  // namespace gdp
