// Block-tiled down kernels of the finest level carrying the PCG update (the fused
// update-down kernel), element type double.
#include "gdp_sblock_launch.h"

template void gdp::launch_down_blk_fused<double>(int, int, int, hipStream_t, const gdp::DownArgs&);
