// Block-tiled level kernels (k_sdown_blk, k_sup_blk, gdp_sblock.h), element type double.
#include "gdp_sblock_launch.h"

template void gdp::launch_down_blk<double>(int, int, int, hipStream_t, const gdp::DownArgs&);
template void gdp::launch_up_blk<double>(int, int, bool, int, hipStream_t, const gdp::UpArgs&);
