// Register-strip coarse-tail kernels fixed to Dirichlet tails of 2 to 4 levels (either
// coarsest method), element type float.
#include "gdp_stail_launch.h"

namespace gdp {
template <>
void launch_stail_spec<float, 0>(int nlev, int exact, bool dots, int blocks, int smem,
                                 hipStream_t s, const TailArgs& a) {
  stail_spec_launch<float, 0>(nlev, exact, dots, blocks, smem, s, a);
}
template <>
hipError_t stail_spec_set_smem<float, 0>(int smem) {
  return stail_spec_smem<float, 0>(smem);
}
}  // namespace gdp
