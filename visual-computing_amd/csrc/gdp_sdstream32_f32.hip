// Streamed down-strip kernels of 32 coarse rows, element type float, for nu = 1 and 2
// (the planner runs other sweep counts with 16-row strips).
#include "gdp_strip_launch.h"

namespace gdp {
template <>
void launch_down_stream32<float>(int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (nu == 1)
    klaunch(k_sdown_stream<float, 1, 32>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st,
            a.upart, a.check_conv, a.apart, a.dcur);
  else
    klaunch(k_sdown_stream<float, 2, 32>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st,
            a.upart, a.check_conv, a.apart, a.dcur);
}
}  // namespace gdp
