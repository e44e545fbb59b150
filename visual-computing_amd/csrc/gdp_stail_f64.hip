// Register-strip coarse-tail kernels, element type double.
#include "gdp_stail_launch.h"

template void gdp::launch_stail<double>(bool, int, int, hipStream_t, const gdp::TailArgs&);
template hipError_t gdp::stail_set_smem<double>(int);
