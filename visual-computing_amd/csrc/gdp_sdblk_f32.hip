// Block-tiled level kernels (k_sdown_blk, k_sup_blk, gdp_sblock.h), element type float.
#include "gdp_sblock_launch.h"

template void gdp::launch_down_blk<float>(int, int, int, hipStream_t, const gdp::DownArgs&);
template void gdp::launch_up_blk<float>(int, int, bool, int, hipStream_t, const gdp::UpArgs&);
