// Host-side declarations shared by the translation units of libgdpoisson.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "gdp_device.h"

namespace gdp {

// Launch timing of the plan profiler.  While a profiling scope is open on this thread it
// points at its events and the scope's one launch carries them: hipExtLaunchKernelGGL
// stamps them from the kernel's own dispatch, so no event packet runs between kernels
// and the times match a kernel trace.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool used = false;
};
extern thread_local LaunchEvents* t_launch_events;

template <typename... KArgs, typename... Args>
inline void klaunch(void (*k)(KArgs...), dim3 grid, dim3 block, unsigned smem, hipStream_t s,
                    Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  if (LaunchEvents* e = t_launch_events) {
    // a later launch of the same scope moves only the stop: the scope times them all
    hipExtLaunchKernelGGL(k, grid, block, smem, s, e->used ? nullptr : e->start, e->stop, 0u,
                          static_cast<KArgs>(args)...);
    e->used = true;
  } else {
    k<<<grid, block, smem, s>>>(static_cast<KArgs>(args)...);
  }
}

// Launches of the strip kernels of the tiled multigrid levels (gdp_sdown_*.hip,
// gdp_sup_*.hip).  menu: strip rows (kDownRows / kUpRows index), nu: sweeps.
struct DownArgs {
  const VTile* tiles;
  ProblemState* st;
  // ||r||^2 partials: read by the convergence test, written by the fused update
  double2* upart;
  // the level whose kernel tests convergence, plus one (0: no test)
  int check_conv;
  int masked;  // finest level of a plan with masked domains (strip height menu 0)
  // the fused update-down kernel: <d, q> partials (for alpha) and the direction buffer
  // whose x half the update leaves to a later kernel
  const double2* apart = nullptr;
  int dcur = 0;
};
struct UpArgs {
  const VTile* tiles;
  const ProblemState* st;
  double2* zpart;
  int order;
  int masked;
  // finest level after the fused update-down kernel: r -= alpha q of the owned points
  int rupd = 0;
};
template <typename T>
void launch_down(int menu, int nu, int blocks, hipStream_t s, const DownArgs& a);
template <typename T>
void launch_up(int menu, int nu, bool dots, int blocks, hipStream_t s, const UpArgs& a);
// The streamed down kernel (k_sdown_stream), crs = 8, 16 or 32 coarse rows per strip
// (32 for nu = 1 and 2 only, in launch_down_stream32).
template <typename T>
void launch_down_stream(int crs, int nu, int blocks, hipStream_t s, const DownArgs& a);
template <typename T>
void launch_down_stream32(int nu, int blocks, hipStream_t s, const DownArgs& a);
// The streamed down kernel of the finest level that also carries the PCG update
// (k_sdown_stream with UPD): CRS = 8, 16 or 32 coarse rows (32 for nu = 1 and 2 only),
// one translation unit each (gdp_sdfuse*.hip).
template <typename T, int CRS>
void launch_down_fused_crs(int nu, int blocks, hipStream_t s, const DownArgs& a);
template <typename T>
inline void launch_down_fused(int crs, int nu, int blocks, hipStream_t s, const DownArgs& a) {
  if (crs == 8) launch_down_fused_crs<T, 8>(nu, blocks, s, a);
  else if (crs == 16 || (nu != 1 && nu != 2)) launch_down_fused_crs<T, 16>(nu, blocks, s, a);
  else launch_down_fused_crs<T, 32>(nu, blocks, s, a);
}
// The block-tiled down kernel (k_sdown_blk), cb = 8, 16 or 32 coarse rows per block, and
// up kernel (k_sup_blk), tb = 16 or 32 fine rows per block.
template <typename T>
void launch_down_blk(int cb, int nu, int blocks, hipStream_t s, const DownArgs& a);
// The block-tiled down kernel of the finest level carrying the PCG update (k_sdown_blk
// with UPD), cb = 8, 16 or 32.
template <typename T>
void launch_down_blk_fused(int cb, int nu, int blocks, hipStream_t s, const DownArgs& a);
template <typename T>
void launch_up_blk(int tb, int nu, bool dots, int blocks, hipStream_t s, const UpArgs& a);

// The register-strip coarse tail (gdp_stail_*.hip): one 512-thread block per problem.
constexpr int kStailBlockThreads = 512;  // threads of a strip-tail block
constexpr int kStailXItems = 4;          // deferred-x items of a thread per round
struct TailArgs {
  const TailDesc* tails;
  const LevelDesc* levels;
  ProblemState* st;
  const KrylovDesc* kds;
  const double2* upart;
  const SlotRange* rk;
  double2* zpart;
  const SlotRange* rz;
  int check_conv, pre, post, coarsest_sweeps, order;
  // deferred x updates (st_xupdate): extra blocks past the tail blocks, over the nxt
  // Krylov tiles of xth x xtw points at xt (xblocks = 0: none)
  int xblocks = 0;
  const KTile* xt = nullptr;
  int nxt = 0, xth = 0, xtw = 0;
};
// Which strip-tail kernel a launch runs: one whose code is fixed to tail problems of
// nlev levels with the given boundary (neu 1 / 0) and coarsest method (exact 1 / 0), when
// every strip-tail problem of the plan is of that kind and kStailSpecMin <= nlev <=
// kStailSpecMax, else the general kernel (nlev 0), which reads them per problem.
struct STailSpec {
  int nlev = 0, neu = -1, exact = -1;
};
constexpr int kStailSpecMin = 2, kStailSpecMax = 4;
template <typename T>
void launch_stail(const STailSpec& sp, bool dots, int blocks, int smem, hipStream_t s,
                  const TailArgs& a);
// Allow smem bytes of dynamic LDS for every strip-tail kernel of element type T.
template <typename T>
hipError_t stail_set_smem(int smem);
// The fixed-kind kernels of one boundary (gdp_stail_spec_*.hip).
template <typename T, int NEU>
void launch_stail_spec(int nlev, int exact, bool dots, int blocks, int smem, hipStream_t s,
                       const TailArgs& a);
template <typename T, int NEU>
hipError_t stail_spec_set_smem(int smem);
// Dynamic LDS bytes of a strip tail whose coarsest level is cm x cn.
int stail_lds_bytes(int cm, int cn, bool exact);
// Largest first level a strip tail takes: rows / columns.
int stail_max_rows(bool fp32);
constexpr int kStailMaxCols = 128;
constexpr int kStailMaxLevels = 5;

hipError_t launch_laplacian(const double* u, double* out, int m, int n, int planes,
                            double spacing, int bc, hipStream_t s);
hipError_t launch_gradient(const double* img, double* gx, double* gy, int h, int w, int planes,
                           hipStream_t s);
hipError_t launch_divergence(const double* gx, const double* gy, double* out, int h, int w,
                             int planes, hipStream_t s);
hipError_t launch_fold(const double* rhs, const double* bd, double* out, int h, int w,
                       int planes, double spacing, hipStream_t s);
hipError_t launch_smooth(double* u, const double* b, int m, int n, double spacing, int sweeps,
                         int bc, int order, hipStream_t s);
hipError_t launch_restrict(const double* f, double* c, int m, int n, int mc, int nc, int bc,
                           hipStream_t s);
hipError_t launch_prolong(const double* c, double* f, int m, int n, int mc, int nc, int bc,
                          hipStream_t s);
hipError_t launch_resnorm_partial(const double* u, const double* b, int m, int n, int planes,
                                  double spacing, int bc, double2* partial, int blocks,
                                  hipStream_t s);

hipError_t launch_clone_rhs(const double* src, int sh, int sw, int row, int col,
                            const uint8_t* mask, const double* tgt, int h, int w, int planes,
                            int top, int left, int hh, int ww, int mode,
                            const uint8_t* interior, double* out, hipStream_t s);
hipError_t launch_masked_rhs(const double* gx, const double* gy, const double* img,
                             const uint8_t* interior, int h, int w, int planes, int top,
                             int left, int hh, int ww, double* out, hipStream_t s);
hipError_t launch_guidance(const double* ax, const double* ay, const double* bx,
                           const double* by, const uint8_t* mask, const double* factor,
                           double fscale, int mode, size_t hw, int planes, double* ox,
                           double* oy, hipStream_t s);
hipError_t launch_illumination(const double* gx, const double* gy, size_t n, double scale,
                               double beta, double epsilon, double* ox, double* oy,
                               hipStream_t s);

void set_error(const char* fmt, ...);

}  // namespace gdp
