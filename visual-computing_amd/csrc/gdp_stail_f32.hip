// Register-strip coarse-tail kernels, element type float.
#include "gdp_stail_launch.h"

template void gdp::launch_stail<float>(bool, int, int, hipStream_t, const gdp::TailArgs&);
template hipError_t gdp::stail_set_smem<float>(int);

int gdp::stail_lds_bytes(int cm, int cn, bool exact) {
  // Cm, Cn, 1/lambda, Cn transposed, and the two work arrays
  const int tables = exact ? (cm * cm + 2 * cn * cn + cm * cn) + 2 * cm * cn : 0;
  return kSTailXbBytes + kSTailXhBytes + 8 * tables;
}
int gdp::stail_max_rows(bool fp32) {
  return kSTailWaves * (fp32 ? STailRows<float>::kRows0 : STailRows<double>::kRows0);
}
