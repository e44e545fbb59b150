// Host dispatch of the register-strip coarse tail.
#pragma once
#include "gdp_internal.h"
#include "gdp_stail.h"

namespace gdp {

static_assert(kStailBlockThreads == kSTailThreads && kStailXItems == kXItems,
              "host and device strip-tail constants");

template <typename T, int NLEV, int NEU, int EXACT>
void launch_stail_k(bool dots, int blocks, int smem, hipStream_t s, const TailArgs& a) {
  if (dots)
    klaunch(k_stail<T, true, NLEV, NEU, EXACT>, dim3(blocks + a.xblocks), dim3(kSTailThreads),
            smem, s, a.tails, a.levels, a.st, a.kds, a.upart, a.rk, a.zpart, a.rz,
            a.check_conv, a.pre, a.post, a.coarsest_sweeps, a.order, blocks, a.xt, a.nxt,
            a.xth, a.xtw);
  else
    klaunch(k_stail<T, false, NLEV, NEU, EXACT>, dim3(blocks + a.xblocks), dim3(kSTailThreads),
            smem, s, a.tails, a.levels, a.st, a.kds, a.upart, a.rk, a.zpart, a.rz,
            a.check_conv, a.pre, a.post, a.coarsest_sweeps, a.order, blocks, a.xt, a.nxt,
            a.xth, a.xtw);
}

template <typename T, int NLEV, int NEU, int EXACT>
hipError_t stail_set_smem_k(int smem) {
  hipError_t e = hipFuncSetAttribute((const void*)k_stail<T, true, NLEV, NEU, EXACT>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, smem);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)k_stail<T, false, NLEV, NEU, EXACT>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, smem);
}

// The fixed-kind kernels of boundary NEU: tail level counts kStailSpecMin ..
// kStailSpecMax, both coarsest methods.
template <typename T, int NEU>
void stail_spec_launch(int nlev, int exact, bool dots, int blocks, int smem, hipStream_t s,
                       const TailArgs& a) {
  static_assert(kStailSpecMin == 2 && kStailSpecMax == 4, "fixed-kind tail kernels");
  switch (nlev * 2 + (exact ? 1 : 0)) {
    case 4: launch_stail_k<T, 2, NEU, 0>(dots, blocks, smem, s, a); break;
    case 5: launch_stail_k<T, 2, NEU, 1>(dots, blocks, smem, s, a); break;
    case 6: launch_stail_k<T, 3, NEU, 0>(dots, blocks, smem, s, a); break;
    case 7: launch_stail_k<T, 3, NEU, 1>(dots, blocks, smem, s, a); break;
    case 8: launch_stail_k<T, 4, NEU, 0>(dots, blocks, smem, s, a); break;
    default: launch_stail_k<T, 4, NEU, 1>(dots, blocks, smem, s, a); break;
  }
}

template <typename T, int NEU>
hipError_t stail_spec_smem(int smem) {
  hipError_t e;
  if ((e = stail_set_smem_k<T, 2, NEU, 0>(smem)) != hipSuccess) return e;
  if ((e = stail_set_smem_k<T, 2, NEU, 1>(smem)) != hipSuccess) return e;
  if ((e = stail_set_smem_k<T, 3, NEU, 0>(smem)) != hipSuccess) return e;
  if ((e = stail_set_smem_k<T, 3, NEU, 1>(smem)) != hipSuccess) return e;
  if ((e = stail_set_smem_k<T, 4, NEU, 0>(smem)) != hipSuccess) return e;
  return stail_set_smem_k<T, 4, NEU, 1>(smem);
}

// The dispatch of a tail launch (instantiated with the general kernel).
template <typename T>
void stail_launch(const STailSpec& sp, bool dots, int blocks, int smem, hipStream_t s,
                  const TailArgs& a) {
  if (sp.nlev >= kStailSpecMin && sp.nlev <= kStailSpecMax && sp.neu >= 0 && sp.exact >= 0) {
    if (sp.neu) launch_stail_spec<T, 1>(sp.nlev, sp.exact, dots, blocks, smem, s, a);
    else launch_stail_spec<T, 0>(sp.nlev, sp.exact, dots, blocks, smem, s, a);
    return;
  }
  launch_stail_k<T, 0, -1, -1>(dots, blocks, smem, s, a);
}

template <typename T>
hipError_t stail_smem(int smem) {
  hipError_t e;
  if ((e = stail_set_smem_k<T, 0, -1, -1>(smem)) != hipSuccess) return e;
  if ((e = stail_spec_set_smem<T, 0>(smem)) != hipSuccess) return e;
  return stail_spec_set_smem<T, 1>(smem);
}

}  // namespace gdp
