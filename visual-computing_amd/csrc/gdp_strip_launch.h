// Host dispatch of the strip kernels: (menu, nu) -> template instance.  Included by
// one small translation unit per kernel and element type so they compile in parallel.
#pragma once
#include "gdp_internal.h"
#include "gdp_strip.h"

namespace gdp {

template <typename T, int CR>
void down_nu(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWN(NU) \
  klaunch(k_sdown<T, NU, CR>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv)
  switch (nu) {
    case 0: GDP_DOWN(0); break;
    case 1: GDP_DOWN(1); break;
    case 2: GDP_DOWN(2); break;
    default: GDP_DOWN(3); break;
  }
#undef GDP_DOWN
}

// Masked levels: kernels of their own, every strip height of the menu.
template <typename T, int CR>
void down_masked(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWN(NU) \
  klaunch(k_sdown<T, NU, CR, true>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv)
  switch (nu) {
    case 0: GDP_DOWN(0); break;
    case 1: GDP_DOWN(1); break;
    case 2: GDP_DOWN(2); break;
    default: GDP_DOWN(3); break;
  }
#undef GDP_DOWN
}

// Streamed strips (menu entries 4, 5, 6): the wavefront-order kernel with CRS = 8, 16 or
// 32 coarse rows, in translation units of their own (gdp_sdstream*.hip; 32 rows for
// nu = 1 and 2 only: the planner gives other sweep counts the 16-row entry).
template <typename T, int CRS>
void down_stream_nu(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWNS(NU) \
  klaunch(k_sdown_stream<T, NU, CRS>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv, a.apart, a.dcur)
  switch (nu) {
    case 0: GDP_DOWNS(0); break;
    case 1: GDP_DOWNS(1); break;
    case 2: GDP_DOWNS(2); break;
    default: GDP_DOWNS(3); break;
  }
#undef GDP_DOWNS
}

// The same strips carrying the PCG update of the finest level (gdp_sdfuse*.hip).
template <typename T, int CRS>
void down_fused_nu(int nu, int blocks, hipStream_t s, const DownArgs& a) {
#define GDP_DOWNF(NU) \
  klaunch(k_sdown_stream<T, NU, CRS, true>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.upart, a.check_conv, a.apart, a.dcur)
  switch (nu) {
    case 0: GDP_DOWNF(0); break;
    case 1: GDP_DOWNF(1); break;
    case 2: GDP_DOWNF(2); break;
    default: GDP_DOWNF(3); break;
  }
#undef GDP_DOWNF
}

template <typename T>
void launch_down(int menu, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (a.masked) {
    // (the planner gives masked levels menu entries 0 .. 3 only)
    switch (menu) {
      case 0: down_masked<T, 1>(nu, blocks, s, a); break;
      case 1: down_masked<T, 2>(nu, blocks, s, a); break;
      case 2: down_masked<T, 4>(nu, blocks, s, a); break;
      default: down_masked<T, 8>(nu, blocks, s, a); break;
    }
    return;
  }
  switch (menu) {
    case 0: down_nu<T, 1>(nu, blocks, s, a); break;
    case 1: down_nu<T, 2>(nu, blocks, s, a); break;
    case 2: down_nu<T, 4>(nu, blocks, s, a); break;
    case 3: down_nu<T, 8>(nu, blocks, s, a); break;
    case 4: case 5: case 6:
      launch_down_stream<T>(menu == 4 ? 8 : (menu == 5 ? 16 : 32), nu, blocks, s, a);
      break;
    case 7: case 8: case 9:
      launch_down_blk<T>(menu == 7 ? 8 : (menu == 8 ? 16 : 32), nu, blocks, s, a);
      break;
    // entries 10 .. 15: the streamed strips and the blocks, which on the finest level of a
    // PCG iteration also carry its update (launch_down_fused, launch_down_blk_fused); here,
    // where no update is pending, plain
    case 10: case 11: case 12:
      launch_down_stream<T>(menu == 10 ? 8 : (menu == 11 ? 16 : 32), nu, blocks, s, a);
      break;
    default: launch_down_blk<T>(menu == 13 ? 8 : (menu == 14 ? 16 : 32), nu, blocks, s, a); break;
  }
}

template <typename T, int TR, bool DOTS>
void up_nu(int nu, int blocks, hipStream_t s, const UpArgs& a) {
#define GDP_UP(NU) \
  klaunch(k_sup<T, NU, TR, DOTS>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.zpart, a.order, a.rupd)
  switch (nu) {
    case 0: GDP_UP(0); break;
    case 1: GDP_UP(1); break;
    case 2: GDP_UP(2); break;
    default: GDP_UP(3); break;
  }
#undef GDP_UP
}

template <typename T, bool DOTS>
void launch_up_d(int menu, int nu, int blocks, hipStream_t s, const UpArgs& a) {
  switch (menu) {
    case 0: up_nu<T, 2, DOTS>(nu, blocks, s, a); break;
    case 1: up_nu<T, 4, DOTS>(nu, blocks, s, a); break;
    case 2: up_nu<T, 8, DOTS>(nu, blocks, s, a); break;
    case 3: up_nu<T, 16, DOTS>(nu, blocks, s, a); break;
    default: launch_up_blk<T>(menu == 4 ? 16 : 32, nu, DOTS, blocks, s, a); break;
  }
}

template <typename T, int TR, bool DOTS>
void up_masked(int nu, int blocks, hipStream_t s, const UpArgs& a) {
#define GDP_UP(NU) \
  klaunch(k_sup<T, NU, TR, DOTS, true>, dim3(blocks), dim3(kThreads), 0, s, a.tiles, a.st, a.zpart, a.order, a.rupd)
  switch (nu) {
    case 0: GDP_UP(0); break;
    case 1: GDP_UP(1); break;
    case 2: GDP_UP(2); break;
    default: GDP_UP(3); break;
  }
#undef GDP_UP
}

template <typename T, bool DOTS>
void launch_up_masked(int menu, int nu, int blocks, hipStream_t s, const UpArgs& a) {
  switch (menu) {
    case 0: up_masked<T, 2, DOTS>(nu, blocks, s, a); break;
    case 1: up_masked<T, 4, DOTS>(nu, blocks, s, a); break;
    case 2: up_masked<T, 8, DOTS>(nu, blocks, s, a); break;
    default: up_masked<T, 16, DOTS>(nu, blocks, s, a); break;
  }
}

template <typename T>
void launch_up(int menu, int nu, bool dots, int blocks, hipStream_t s, const UpArgs& a) {
  if (a.masked) {
    if (dots) launch_up_masked<T, true>(menu, nu, blocks, s, a);
    else launch_up_masked<T, false>(menu, nu, blocks, s, a);
    return;
  }
  if (dots) launch_up_d<T, true>(menu, nu, blocks, s, a);
  else launch_up_d<T, false>(menu, nu, blocks, s, a);
}

}  // namespace gdp
