// Host dispatch of the block-tiled level kernels: (rows per block, nu) -> template
// instance.  Included by one translation unit per element type.
#pragma once
#include "gdp_internal.h"
#include "gdp_sblock.h"

namespace gdp {

template <typename T, int CB>
void down_blk_nu(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWNB(NU) \
  klaunch(k_sdown_blk<T, NU, CB>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv, a.apart, a.dcur)
  switch (nu) {
    case 0: GDP_DOWNB(0); break;
    case 1: GDP_DOWNB(1); break;
    case 2: GDP_DOWNB(2); break;
    default: GDP_DOWNB(3); break;
  }
#undef GDP_DOWNB
}

template <typename T>
void launch_down_blk(int cb, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (cb == 8) down_blk_nu<T, 8>(nu, blocks, s, a);
  else if (cb == 16) down_blk_nu<T, 16>(nu, blocks, s, a);
  else down_blk_nu<T, 32>(nu, blocks, s, a);
}

// The same blocks carrying the PCG update of the finest level (gdp_sdblkfuse*.hip).
template <typename T, int CB>
void down_blk_fused_nu(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWNBF(NU) \
  klaunch(k_sdown_blk<T, NU, CB, true>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv, a.apart, a.dcur)
  switch (nu) {
    case 0: GDP_DOWNBF(0); break;
    case 1: GDP_DOWNBF(1); break;
    case 2: GDP_DOWNBF(2); break;
    default: GDP_DOWNBF(3); break;
  }
#undef GDP_DOWNBF
}

template <typename T>
void launch_down_blk_fused(int cb, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (cb == 8) down_blk_fused_nu<T, 8>(nu, blocks, s, a);
  else if (cb == 16) down_blk_fused_nu<T, 16>(nu, blocks, s, a);
  else down_blk_fused_nu<T, 32>(nu, blocks, s, a);
}

template <typename T, int TB, bool DOTS>
void up_blk_nu(int nu, int blocks, hipStream_t s, const UpArgs& a) {
#define GDP_UPB(NU) \
  klaunch(k_sup_blk<T, NU, TB, DOTS>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.zpart, a.order, a.rupd)
  switch (nu) {
    case 0: GDP_UPB(0); break;
    case 1: GDP_UPB(1); break;
    case 2: GDP_UPB(2); break;
    default: GDP_UPB(3); break;
  }
#undef GDP_UPB
}

template <typename T>
void launch_up_blk(int tb, int nu, bool dots, int blocks, hipStream_t s, const UpArgs& a) {
  if (dots) {
    if (tb == 16) up_blk_nu<T, 16, true>(nu, blocks, s, a);
    else up_blk_nu<T, 32, true>(nu, blocks, s, a);
  } else {
    if (tb == 16) up_blk_nu<T, 16, false>(nu, blocks, s, a);
    else up_blk_nu<T, 32, false>(nu, blocks, s, a);
  }
}

}  // namespace gdp
