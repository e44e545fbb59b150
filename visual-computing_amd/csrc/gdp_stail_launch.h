// Host dispatch of the register-strip coarse tail.
#pragma once
#include "gdp_internal.h"
#include "gdp_stail.h"

namespace gdp {

static_assert(kStailBlockThreads == kSTailThreads && kStailXItems == kXItems,
              "host and device strip-tail constants");

template <typename T>
void launch_stail(bool dots, int blocks, int smem, hipStream_t s, const TailArgs& a) {
  if (dots)
    klaunch(k_stail<T, true>, dim3(blocks + a.xblocks), dim3(kSTailThreads), smem, s, a.tails,
            a.levels, a.st, a.kds, a.upart, a.rk, a.zpart, a.rz, a.check_conv, a.pre, a.post,
            a.coarsest_sweeps, a.order, blocks, a.xt, a.nxt, a.xth, a.xtw);
  else
    klaunch(k_stail<T, false>, dim3(blocks + a.xblocks), dim3(kSTailThreads), smem, s, a.tails,
            a.levels, a.st, a.kds, a.upart, a.rk, a.zpart, a.rz, a.check_conv, a.pre, a.post,
            a.coarsest_sweeps, a.order, blocks, a.xt, a.nxt, a.xth, a.xtw);
}

template <typename T>
hipError_t stail_set_smem(int smem) {
  hipError_t e = hipFuncSetAttribute((const void*)k_stail<T, true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, smem);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_stail<T, false>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, smem);
}

}  // namespace gdp
