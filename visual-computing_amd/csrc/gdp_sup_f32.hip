// Up-strip kernels (prolong + post-smooth [+ PCG dot products]), element type float.
#include "gdp_strip_launch.h"

template void gdp::launch_up<float>(int, int, bool, int, hipStream_t, const gdp::UpArgs&);
