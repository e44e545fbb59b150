// Down-strip kernels (pre-smooth + residual + restrict), element type float.
#include "gdp_strip_launch.h"

template void gdp::launch_down<float>(int, int, int, hipStream_t, const gdp::DownArgs&);
