// Register-strip coarse-tail kernels, element type float: the general kernel and the
// dispatch of every tail launch.
#include "gdp_stail_launch.h"

namespace gdp {
template <>
void launch_stail<float>(const STailSpec& sp, bool dots, int blocks, int smem, hipStream_t s,
                        const TailArgs& a) {
  stail_launch<float>(sp, dots, blocks, smem, s, a);
}
template <>
hipError_t stail_set_smem<float>(int smem) {
  return stail_smem<float>(smem);
}
}  // namespace gdp

int gdp::stail_lds_bytes(int cm, int cn, bool exact) {
  // Cm, Cn, 1/lambda, Cn and Cm transposed, and the two work arrays
  const int tables = exact ? (2 * cm * cm + 2 * cn * cn + cm * cn) + 2 * cm * cn : 0;
  return kSTailXbBytes + kSTailXhBytes + 8 * tables;
}
int gdp::stail_max_rows(bool fp32) {
  return kSTailWaves * (fp32 ? STailRows<float>::kRows0 : STailRows<double>::kRows0);
}
