// Down-strip kernels (pre-smooth + residual + restrict), element type double.
#include "gdp_strip_launch.h"

template void gdp::launch_down<double>(int, int, int, hipStream_t, const gdp::DownArgs&);
