// Up-strip kernels (prolong + post-smooth [+ PCG dot products]), element type double.
#include "gdp_strip_launch.h"

template void gdp::launch_up<double>(int, int, bool, int, hipStream_t, const gdp::UpArgs&);
