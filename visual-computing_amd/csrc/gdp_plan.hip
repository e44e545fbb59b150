// Host side of libgdpoisson: level hierarchy, workspace layout, tile tables, launch
// sequences, hipGraph capture of PCG iteration pairs, and the C ABI of gdpoisson.h.
//
// Level hierarchy (per problem):
//   Dirichlet  m -> (m - 1) // 2 until min side <= coarsest_size or a coarse side < 2
//              (multigrid.py:143-144, 206)
//   Neumann    m -> (m + 1) // 2 until max side <= coarsest_size or min side < 3
// The coarse tail of a problem (a level and every coarser one) runs in one block per
// problem: k_stail (registers, gdp_stail.h) from the first level at most 128 columns
// wide that it can hold, else k_tail (LDS) from the first level whose tail fits in LDS.
// Finer levels use the strip kernels k_sdown / k_sup (gdp_strip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "gdp_device.h"
#include "gdp_internal.h"
#include "gdp_kernels.h"
#include "gdp_apply.h"
#include "gdpoisson.h"

namespace gdp {

thread_local LaunchEvents* t_launch_events = nullptr;

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

}  // namespace gdp

using namespace gdp;

#define GDP_HIP(x)                                                              \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      set_error("%s failed: %s", #x, hipGetErrorString(e_));                    \
      return GDP_EHIP;                                                          \
    }                                                                           \
  } while (0)

#define GDP_CHECK_LAUNCH()                                                      \
  do {                                                                          \
    hipError_t e_ = hipGetLastError();                                          \
    if (e_ != hipSuccess) {                                                     \
      set_error("kernel launch failed: %s", hipGetErrorString(e_));             \
      return GDP_EHIP;                                                          \
    }                                                                           \
  } while (0)

namespace {

struct Shape2 {
  int h, w;
};
// Tile menus.  Kind 0 = down: coarse rows per strip (a block stacks four strips; a
// strip is StripGeo<nu>::kDownCols coarse columns wide); entries 4, 5 and 6 are the
// streamed (wavefront-order) down kernel, k_sdown_stream, the last one compiled for
// nu = 1 and 2 only; entries 7, 8 and 9 the block-tiled kernel, k_sdown_blk, whose block
// owns 8, 16 or 32 coarse rows (four times the listed rows, like the others).  Kind 1 =
// up: fine rows per strip (kUpCols fine columns wide); entries 4 and 5 the block-tiled
// k_sup_blk, whose block owns 16 or 32 fine rows.  Kind 2 = Krylov: fine points per
// block.  Down entries 10 .. 15 (finest level only) also carry the PCG update of an
// iteration, in place of k_update (the finest up kernel then stores r): 10, 11, 12 the
// streamed strips of 8, 16 and 32 coarse rows (k_sdown_stream with UPD), 13, 14, 15 the
// block-tiled kernel's blocks of 8, 16 and 32 coarse rows (k_sdown_blk with UPD).
const int kDownRows[16] = {1, 2, 4, 8, 8, 16, 32, 2, 4, 8, 8, 16, 32, 2, 4, 8};
constexpr int kFusedDown = 10;     // first fused update-down entry
constexpr int kFusedDownBlk = 13;  // first block-tiled one
const int kUpRows[6] = {2, 4, 8, 16, 4, 8};
const Shape2 kKryMenu[6] = {{8, 64}, {16, 64}, {32, 64}, {8, 128}, {16, 128}, {32, 128}};
constexpr int kMenu = 6;       // up-kernel menu (rows per strip)
constexpr int kDownMenuN = 16;  // down-kernel menu
constexpr int kKryMenuN = 6;   // Krylov tile menu
int menu_size(int kind) { return kind == 2 ? kKryMenuN : (kind == 0 ? kDownMenuN : kMenu); }
// The 32-row streamed down entry exists for nu = 1 and 2; other sweep counts run the
// 16-row one.
int down_entry(int idx, int nu) {
  return ((idx == 6 || idx == 12) && nu != 1 && nu != 2) ? idx - 1 : idx;
}
constexpr int kStrips = kThreads / 64;  // strips per block
int down_cols(int nu) { return (125 - 4 * nu) / 2; }
int up_cols(int nu) { return 124 - 4 * nu; }
constexpr int kMaxDepth = 32;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
// Bytes of the problem states rounded up to 64: the offset of the check's flag behind them,
// on the device and in the pinned host block alike.
size_t state_span(int count) { return (sizeof(ProblemState) * size_t(count) + 63) & ~size_t(63); }
int round_up(int x, int a) { return (x + a - 1) / a * a; }
int cdiv(int a, int b) { return (a + b - 1) / b; }

// ----------------------------------------------------------------- launch dispatch
#define GDP_KRY_DISPATCH(menu, CALL)                 \
  switch (menu) {                                    \
    case 0: { constexpr int TH = 8, TW = 64; CALL; } break;   \
    case 1: { constexpr int TH = 16, TW = 64; CALL; } break;  \
    case 2: { constexpr int TH = 32, TW = 64; CALL; } break;  \
    case 3: { constexpr int TH = 8, TW = 128; CALL; } break;  \
    case 4: { constexpr int TH = 16, TW = 128; CALL; } break; \
    default: { constexpr int TH = 32, TW = 128; CALL; } break; \
  }

// ----------------------------------------------------------------- plan
struct LevelInfo {
  int m, n, pitch;
  double h;
  size_t off_u, off_v, off_b;
  size_t off_mask = 0;  // coarse levels of a Dirichlet problem: mask by injection (m x n)
};

struct ProbInfo {
  gdp_problem p;
  bool exact = false;  // exact transform solve on the coarsest level
  std::vector<LevelInfo> lv;
  int tail_first = 0;
  int tail_bytes = 0;
  bool stail = false;  // register-strip tail (k_stail) rather than the LDS tail (k_tail)
  int level_base = 0;
  size_t off_x = 0, off_r = 0, off_d0 = 0, off_d1 = 0, off_q = 0, off_dct = 0;
  int tail_off[kMaxTailLevels][2];
  int scratch_off = 0, coarse_off = 0, tab_off = 0;
  const uint8_t* mask = nullptr;  // masked Dirichlet domain (device, m x n), or null
};

struct Table {
  size_t off = 0;
  int count = 0;
};

}  // namespace

struct gdp_plan {
  gdp_options opt;
  bool fp32 = false;
  size_t tsize = 8;
  std::vector<ProbInfo> probs;
  int depth = 0;
  int down_menu[kMaxDepth];
  int up_menu[kMaxDepth];
  int kry_menu = 1;
  // coarsest sweeps of the problems without an exact coarsest solve (the option's count,
  // or the reference's 60 where the option asked for exact solves and a problem's
  // tables did not fit)
  int csweeps = 0;
  int total_levels = 0;
  int tail_smem = 0;   // dynamic LDS of the LDS tail (k_tail)
  int stail_smem = 0;  // dynamic LDS of the register-strip tail (k_stail)
  int n_stail = 0;     // strip-tail problems: the first n_stail entries of the tail table
  STailSpec stail_spec;  // the strip-tail kernel fixed to its problems' kind, when they agree
  int n_cu = 0;        // compute units of the bound device (blocks for deferred x updates)
  int max_slots = 1;
  size_t ws_bytes = 0;
  size_t off_levels = 0, off_kry = 0, off_tails = 0, off_state = 0;
  size_t off_apart = 0, off_upart = 0, off_zpart = 0, off_spart = 0;
  size_t off_ranges_kry = 0, off_ranges_up0 = 0, off_ranges_down0 = 0, off_io = 0,
         off_init = 0;
  size_t off_hist = 0;
  size_t off_tiles = 0;     // start of the tile tables region
  size_t tiles_bytes = 0;   // reserved bytes for the tile tables
  Table down[kMaxDepth], up[kMaxDepth], kry, kapply;
  // runtime
  char* ws = nullptr;
  IoDesc* h_io = nullptr;
  const double** h_init = nullptr;
  ProblemState* h_state = nullptr;
  void* h_block = nullptr;     // pinned block holding h_io, h_init, h_state
  size_t h_block_bytes = 0;
  hipStream_t cap = nullptr;
  hipEvent_t io_ev = nullptr;  // the last upload of h_io / h_init has been consumed
  bool io_valid = false;       // the device copy holds the pointers now in h_io
  hipGraphExec_t pair_exec[2] = {nullptr, nullptr};   // [0] mgcg, [1] cg
  // a whole first batch (set-up, first preconditioning, N pairs, device-side check) per N
  std::vector<std::pair<int, hipGraphExec_t>> head_exec[2];
  hipStream_t graph_stream = nullptr;  // stream of the last graph launch
  int last_iters = 0;
  bool bound = false;
  // built-in kernel timing (HIP events around every launch on the launch stream)
  bool prof = false;
  struct ProfRec {
    int id;
    hipEvent_t a, b;
    double bytes;
  };
  struct ProfAgg {
    std::string name;
    double ms = 0.0, bytes = 0.0;
    long launches = 0;
  };
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  std::vector<ProfAgg> prof_agg;

  template <typename X>
  X* at(size_t off) const { return reinterpret_cast<X*>(ws + off); }
  LevelDesc* d_levels() const { return at<LevelDesc>(off_levels); }
  KrylovDesc* d_kry() const { return at<KrylovDesc>(off_kry); }
  TailDesc* d_tails() const { return at<TailDesc>(off_tails); }
  ProblemState* d_state() const { return at<ProblemState>(off_state); }
  double2* d_apart() const { return at<double2>(off_apart); }  // <d, q> partials
  double2* d_upart() const { return at<double2>(off_upart); }  // ||r||^2 partials
  double2* d_zpart() const { return at<double2>(off_zpart); }  // <r, z>, <q, z> partials
  double2* d_spart() const { return at<double2>(off_spart); }  // set-up / residual partials
  SlotRange* d_rkry() const { return at<SlotRange>(off_ranges_kry); }
  SlotRange* d_rup0() const { return at<SlotRange>(off_ranges_up0); }
  // ||r||^2 partials of the fused update-down kernel, per problem
  SlotRange* d_rdown0() const { return at<SlotRange>(off_ranges_down0); }
  IoDesc* d_io() const { return at<IoDesc>(off_io); }
  const double** d_init() const { return at<const double*>(off_init); }
  double* d_hist() const { return at<double>(off_hist); }
  int count() const { return int(probs.size()); }
};

namespace {

std::vector<std::pair<int, int>> level_shapes(const gdp_problem& p, const gdp_options& o) {
  std::vector<std::pair<int, int>> s;
  int m = p.rows, n = p.cols;
  s.push_back({m, n});
  for (;;) {
    if (p.boundary == GDP_NEUMANN) {
      if (std::max(m, n) <= o.coarsest_size || std::min(m, n) < 3) break;
      m = (m + 1) / 2;
      n = (n + 1) / 2;
    } else {
      const int mc = (m - 1) / 2, nc = (n - 1) / 2;
      if (std::min(m, n) <= o.coarsest_size || std::min(mc, nc) < 2) break;
      m = mc;
      n = nc;
    }
    s.push_back({m, n});
  }
  return s;
}

// Pinned host staging blocks of plans are kept for reuse rather than freed: freeing
// pinned memory may wait for the whole device, which would stall the work other
// streams (other plans) have in flight.  Blocks come in power-of-two size classes.
struct PinnedPool {
  std::mutex mu;
  std::unordered_map<size_t, std::vector<void*>> free;
};
PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool;  // lives until the process ends
  return *pool;
}
size_t pinned_class(size_t bytes) {
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  return c;
}
hipError_t pinned_get(size_t bytes, void** out) {
  const size_t c = pinned_class(bytes);
  {
    PinnedPool& pool = pinned_pool();
    std::lock_guard<std::mutex> lock(pool.mu);
    std::vector<void*>& v = pool.free[c];
    if (!v.empty()) {
      *out = v.back();
      v.pop_back();
      return hipSuccess;
    }
  }
  return hipHostMalloc(out, c, hipHostMallocDefault);
}
void pinned_put(void* p, size_t bytes) {
  PinnedPool& pool = pinned_pool();
  std::lock_guard<std::mutex> lock(pool.mu);
  pool.free[pinned_class(bytes)].push_back(p);
}

// coarse mask (mc x nc) from the next finer one (row length nf) by injection: the coarse
// point (I, J) of a Dirichlet hierarchy coincides with the fine point (2I + 1, 2J + 1)
__global__ __launch_bounds__(256) void k_mask_inject(const uint8_t* __restrict__ fine, int nf,
                                                     uint8_t* __restrict__ coarse, int mc, int nc) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= mc * nc) return;
  const int I = t / nc, J = t - I * nc;
  coarse[t] = fine[(size_t)(2 * I + 1) * nf + 2 * J + 1] ? 1 : 0;
}

bool has_mask(const gdp_plan* P) {
  for (const ProbInfo& pi : P->probs)
    if (pi.mask) return true;
  return false;
}

// The down-kernel menu entry of level l: the masked kernels come in the sweep-order form
// only, so a plan with masked domains runs a streamed or block-tiled entry as the
// tallest one of that form.
int down_menu_of(const gdp_plan* P, int l) {
  const int m = down_entry(P->down_menu[l], P->opt.pre_smoothing);
  return (m >= 4 && has_mask(P)) ? 3 : m;
}
// The finest level's down kernel carries the PCG update (fused update-down entries)
// when its entry asks for it, no domain is masked and every problem's finest level is a
// tiled one (a problem whose finest level is in the tail needs k_update).
bool fused_update(const gdp_plan* P) {
  if (P->depth < 1 || down_menu_of(P, 0) < kFusedDown) return false;
  for (const ProbInfo& pi : P->probs)
    if (pi.tail_first < 1) return false;
  return true;
}
// The up-kernel menu entry of level l: likewise the masked up kernels are strips only.
int up_menu_of(const gdp_plan* P, int l) {
  const int m = P->up_menu[l];
  return (m >= 4 && has_mask(P)) ? 3 : m;
}

// Raises the dynamic-LDS limit of a tail kernel family (0 / 1: strip tail float / double,
// 2 / 3: LDS tail float / double) to at least smem.  The limit is process-wide and only
// ever rises: a plan bound later, possibly in another thread, must not lower a limit
// that launches of an earlier plan rely on.
hipError_t raise_smem(int family, int smem) {
  static std::mutex mu;
  static int cur[4] = {0, 0, 0, 0};
  std::lock_guard<std::mutex> lock(mu);
  if (smem <= cur[family]) return hipSuccess;
  hipError_t e = hipSuccess;
  switch (family) {
    case 0: e = stail_set_smem<float>(smem); break;
    case 1: e = stail_set_smem<double>(smem); break;
    case 2:
      e = hipFuncSetAttribute((const void*)k_tail<float, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, smem);
      if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_tail<float, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, smem);
      break;
    default:
      e = hipFuncSetAttribute((const void*)k_tail<double, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, smem);
      if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)k_tail<double, false>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, smem);
  }
  if (e == hipSuccess) cur[family] = smem;
  return e;
}

// Instantiated hipGraphs of a plan are tied to its tables: drop them on re-bind /
// destroy, once the stream they last ran on has finished them (other streams, and so
// other plans' work, are not waited for).
void destroy_graphs(gdp_plan* P) {
  bool any = !P->head_exec[0].empty() || !P->head_exec[1].empty();
  any = any || P->pair_exec[0] || P->pair_exec[1];
  if (any && P->graph_stream) (void)hipStreamSynchronize(P->graph_stream);
  P->graph_stream = nullptr;
  for (auto& v : P->head_exec) {
    for (auto& e : v)
      if (e.second) (void)hipGraphExecDestroy(e.second);
    v.clear();
  }
  for (int g = 0; g < 2; ++g)
    if (P->pair_exec[g]) {
      (void)hipGraphExecDestroy(P->pair_exec[g]);
      P->pair_exec[g] = nullptr;
    }
}

// LDS bytes of a tail holding levels [first, L) (u and b per level, one residual
// scratch of the first level, two FP64 arrays of the coarsest).  Fills the offsets.
size_t tail_layout(ProbInfo& pi, int first, size_t tsize) {
  const int L = int(pi.lv.size());
  if (L - first > kMaxTailLevels) return size_t(1) << 40;
  size_t off = 0;
  auto a16 = [](size_t x) { return (x + 15) & ~size_t(15); };
  for (int l = first; l < L; ++l) {
    const size_t pts = size_t(pi.lv[l].m) * pi.lv[l].n;
    if (pts > (size_t(1) << 24)) return size_t(1) << 40;  // far beyond any LDS
    pi.tail_off[l - first][0] = int(off);
    off += a16(pts * tsize);
    pi.tail_off[l - first][1] = int(off);
    off += a16(pts * tsize);
  }
  pi.scratch_off = int(off);
  off += a16(size_t(pi.lv[first].m) * pi.lv[first].n * tsize);
  pi.coarse_off = int(off);
  const size_t cm = pi.lv[L - 1].m, cn = pi.lv[L - 1].n;
  off += 2 * cm * cn * 8;
  pi.tab_off = int(off);
  if (pi.exact) off += (cm * cm + cn * cn + cm * cn) * 8;
  return off;
}

// Block tile of a menu entry: down in coarse points, up and Krylov in fine points; the
// strip widths depend on the number of sweeps nu of the kernel.
Shape2 menu_shape(int kind, int idx, int nu) {
  if (kind == 0) return Shape2{kStrips * kDownRows[down_entry(idx, nu)], down_cols(nu)};
  if (kind == 1) return Shape2{kStrips * kUpRows[idx], up_cols(nu)};
  return kKryMenu[idx];
}

// Strip tiles of the PCG apply (k_apply_s) over an m x n grid.
int apply_tiles(int m, int n) {
  return cdiv(m, kApplyBlockRows) * std::max(1, cdiv(n - 1, kApplyCols));
}

// Worst case tile count of a grid over every entry of a menu.
int max_tiles(int kind, int m, int n, int nu) {
  int best = 0;
  for (int i = 0; i < menu_size(kind); ++i) {
    const Shape2 s = menu_shape(kind, i, nu);
    best = std::max(best, cdiv(m, s.h) * cdiv(n, s.w));
  }
  return best;
}

// Block b of a launch runs on XCD b % 8; give each XCD a contiguous run of tiles so
// neighbouring tiles (which share halo rows) hit the same L2.
template <typename TT>
std::vector<TT> xcd_swizzle(const std::vector<TT>& in) {
  const int N = int(in.size()), X = 8;
  const int base = N / X, rem = N % X;
  std::vector<TT> out(N);
  for (int b = 0; b < N; ++b) {
    const int x = b % X, k = b / X;
    const int start = x * base + std::min(x, rem);
    out[b] = in[start + k];
  }
  return out;
}

}  // namespace

// ----------------------------------------------------------------- C ABI
extern "C" {

const char* gdp_last_error(void) { return g_err.c_str(); }
int gdp_version(void) { return 10000; }

int gdp_tile_menu(int kind, int index, int nu, int* th, int* tw) {
  if (kind < 0 || kind > 2 || index < 0 || index >= menu_size(kind) || nu < 0 || nu > 3 ||
      !th || !tw) {
    set_error("gdp_tile_menu: bad kind/index/nu");
    return GDP_EINVAL;
  }
  const Shape2 s = menu_shape(kind, index, nu);
  *th = s.h;
  *tw = s.w;
  return GDP_OK;
}

int gdp_plan_create(const gdp_problem* problems, int count, const gdp_options* options,
                    gdp_plan** out) {
  if (!problems || count <= 0 || !options || !out) {
    set_error("gdp_plan_create: null argument or empty batch");
    return GDP_EINVAL;
  }
  const gdp_options& o = *options;
  if (o.precision != GDP_FP64 && o.precision != GDP_MIXED) {
    set_error("unknown precision %d", o.precision);
    return GDP_EINVAL;
  }
  if (o.pre_smoothing < 0 || o.pre_smoothing > 3 || o.post_smoothing < 0 ||
      o.post_smoothing > 3) {
    set_error("pre/post smoothing must be in [0, 3] (got %d/%d)", o.pre_smoothing,
              o.post_smoothing);
    return GDP_ENOTSUP;
  }
  if (o.coarsest_size < 1 || o.coarsest_sweeps < 0 || o.max_iterations < 0) {
    set_error("bad coarsest_size / coarsest_sweeps / max_iterations");
    return GDP_EINVAL;
  }
  gdp_plan* P = new gdp_plan();
  P->opt = o;
  P->fp32 = o.precision == GDP_MIXED;
  P->tsize = P->fp32 ? 4 : 8;
  for (int i = 0; i < kMaxDepth; ++i) {
    P->down_menu[i] = i == 0 ? 2 : 1;
    P->up_menu[i] = i == 0 ? 2 : 1;
  }
  P->kry_menu = 1;
  P->csweeps = o.coarsest_sweeps > 0 ? o.coarsest_sweeps : 60;
  int lvl_total = 0;
  for (int i = 0; i < count; ++i) {
    const gdp_problem& pr = problems[i];
    if (pr.rows < 1 || pr.cols < 1 || !(pr.spacing > 0.0) ||
        (pr.boundary != GDP_DIRICHLET && pr.boundary != GDP_NEUMANN) ||
        (pr.flags & ~GDP_PROBLEM_MASKED) ||
        ((pr.flags & GDP_PROBLEM_MASKED) && pr.boundary != GDP_DIRICHLET)) {
      set_error("problem %d: bad shape (%d, %d), spacing, boundary or flags", i, pr.rows,
                pr.cols);
      delete P;
      return GDP_EINVAL;
    }
    ProbInfo pi;
    pi.p = pr;
    pi.exact = pr.boundary == GDP_NEUMANN || o.coarsest_sweeps == 0;
    auto shapes = level_shapes(pr, o);
    double h = pr.spacing;
    for (auto& s : shapes) {
      LevelInfo li;
      li.m = s.first;
      li.n = s.second;
      li.pitch = round_up(s.second, 32);
      li.h = h;
      h *= 2.0;
      pi.lv.push_back(li);
    }
    const int L = int(pi.lv.size());
    int first = -1;
    // a masked problem's tail holds its coarsest level only: the strip kernels smooth
    // every other level with the mask
    const int f0 = (pr.flags & GDP_PROBLEM_MASKED) ? std::min(L - 1, kMaxDepth - 1) : 0;
    auto place_tail = [&]() {
      // the register-strip tail takes the first level at most 128 columns wide and
      // stail_max_rows rows tall whose chain to the coarsest fits its level count and LDS
      for (int f = f0; f < L; ++f) {
        const auto& c = pi.lv[L - 1];
        const int bytes = stail_lds_bytes(c.m, c.n, pi.exact);
        if (pi.lv[f].m <= stail_max_rows(P->fp32) && pi.lv[f].n <= kStailMaxCols &&
            L - f <= kStailMaxLevels && bytes <= kTailLdsBytes) {
          first = f;
          pi.stail = true;
          pi.tail_bytes = bytes;
          return;
        }
      }
      for (int f = f0; f < L; ++f) {
        const size_t bytes = tail_layout(pi, f, P->tsize);
        if (bytes <= size_t(kTailLdsBytes)) {
          first = f;
          pi.tail_bytes = int(bytes);
          return;
        }
      }
    };
    place_tail();
    if (first < 0 && pi.exact) {
      // a coarsest level whose transform tables overflow LDS (a thin grid: its short
      // side stops the coarsening while its long side is still long) is solved by
      // red-black sweeps instead, the reference's coarsest rule (multigrid.py:206-207)
      pi.exact = false;
      place_tail();
    }
    if (first < 0) {
      set_error("problem %d (%d x %d): coarsest level %d x %d does not fit in LDS", i, pr.rows,
                pr.cols, pi.lv[L - 1].m, pi.lv[L - 1].n);
      delete P;
      return GDP_ENOTSUP;
    }
    if (first >= kMaxDepth) {
      set_error("problem %d: too many levels", i);
      delete P;
      return GDP_ENOTSUP;
    }
    pi.tail_first = first;
    if (!pi.stail) tail_layout(pi, first, P->tsize);
    pi.level_base = lvl_total;
    lvl_total += L;
    P->depth = std::max(P->depth, first);
    if (pi.stail) {
      P->stail_smem = std::max(P->stail_smem, pi.tail_bytes);
      P->n_stail++;
    } else {
      P->tail_smem = std::max(P->tail_smem, pi.tail_bytes);
    }
    P->probs.push_back(pi);
  }
  P->total_levels = lvl_total;
  {
    // the strip-tail problems share a level count, boundary and coarsest method: their
    // launches run the kernel fixed to that kind (same arithmetic, less code to fetch)
    STailSpec sp;
    bool first_st = true, same = true;
    for (const ProbInfo& pi : P->probs) {
      if (!pi.stail) continue;
      const int nl = int(pi.lv.size()) - pi.tail_first;
      const int ne = pi.p.boundary == GDP_NEUMANN ? 1 : 0, ex = pi.exact ? 1 : 0;
      if (first_st) {
        sp = STailSpec{nl, ne, ex};
        first_st = false;
      } else if (sp.nlev != nl || sp.neu != ne || sp.exact != ex) {
        same = false;
      }
    }
    if (!first_st && same && sp.nlev >= kStailSpecMin && sp.nlev <= kStailSpecMax)
      P->stail_spec = sp;
  }

  // ---- workspace layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o2 = off;
    off = align256(off + bytes);
    return o2;
  };
  for (auto& pi : P->probs) {
    for (size_t l = 0; l < pi.lv.size(); ++l) {
      LevelInfo& li = pi.lv[l];
      const size_t a = size_t(li.m) * li.pitch * P->tsize;
      li.off_u = take(a);
      li.off_v = take(a);
      li.off_b = take(a);
      // room for the level's mask should the problem be given a masked domain
      if (l > 0 && pi.p.boundary == GDP_DIRICHLET) li.off_mask = take(size_t(li.m) * li.n);
    }
    const size_t k = size_t(pi.lv[0].m) * pi.lv[0].pitch * 8;
    pi.off_x = take(k);
    pi.off_r = take(k);
    pi.off_d0 = take(k);
    pi.off_d1 = take(k);
    pi.off_q = take(k);
    const auto& c = pi.lv.back();
    if (pi.exact) pi.off_dct = take(size_t(c.m * c.m + c.n * c.n + c.m * c.n) * 8);
  }
  const int P_ = count;
  P->off_levels = take(sizeof(LevelDesc) * lvl_total);
  P->off_kry = take(sizeof(KrylovDesc) * P_);
  P->off_tails = take(sizeof(TailDesc) * P_);
  P->off_state = take(state_span(P_));
  P->off_ranges_kry = take(sizeof(SlotRange) * P_);
  P->off_ranges_up0 = take(sizeof(SlotRange) * P_);
  P->off_ranges_down0 = take(sizeof(SlotRange) * P_);
  P->off_io = take(sizeof(IoDesc) * P_);
  P->off_init = take(sizeof(double*) * P_);
  P->off_hist = take(sizeof(double) * P_ * (size_t(o.max_iterations) + 2));
  // tile tables (worst case over the menus)
  int slots = 1;
  size_t tiles = 0;
  int kry_slots = 0, up0_slots = 0, apply_slots = 0, down0_slots = 0;
  for (auto& pi : P->probs) {
    kry_slots += max_tiles(2, pi.lv[0].m, pi.lv[0].n, 0);
    apply_slots += apply_tiles(pi.lv[0].m, pi.lv[0].n);
    if (pi.tail_first > 0) {
      up0_slots += max_tiles(1, pi.lv[0].m, pi.lv[0].n, o.post_smoothing);
      // the fused update-down kernel's ||r||^2 partials: one slot per finest down tile
      down0_slots += max_tiles(0, pi.lv[1].m, pi.lv[1].n, o.pre_smoothing);
    }
    for (int l = 0; l < pi.tail_first; ++l) {
      tiles += max_tiles(0, pi.lv[l + 1].m, pi.lv[l + 1].n, o.pre_smoothing);
      tiles += max_tiles(1, pi.lv[l].m, pi.lv[l].n, o.post_smoothing);
    }
  }
  tiles += kry_slots;
  slots = std::max({slots, kry_slots, apply_slots, up0_slots + P_, down0_slots});
  P->max_slots = slots;
  P->off_apart = take(sizeof(double2) * slots);
  P->off_upart = take(sizeof(double2) * slots);
  P->off_zpart = take(sizeof(double2) * slots);
  P->off_spart = take(sizeof(double2) * slots);
  P->tiles_bytes = sizeof(VTile) * (tiles + 8 * (2 * P->depth + 1)) + 256 * (2 * P->depth + 2) +
                   sizeof(KTile) * (apply_slots + 8) + 256;
  P->off_tiles = take(P->tiles_bytes);
  P->ws_bytes = off;
  *out = P;
  return GDP_OK;
}

int gdp_plan_workspace_bytes(const gdp_plan* P, size_t* bytes) {
  if (!P || !bytes) {
    set_error("null argument");
    return GDP_EINVAL;
  }
  *bytes = P->ws_bytes;
  return GDP_OK;
}

int gdp_plan_levels(const gdp_plan* P, int problem, int* levels) {
  if (!P || !levels || problem < 0 || problem >= P->count()) {
    set_error("bad problem index");
    return GDP_EINVAL;
  }
  *levels = int(P->probs[problem].lv.size());
  return GDP_OK;
}

int gdp_plan_level_shape(const gdp_plan* P, int problem, int level, int* m, int* n) {
  if (!P || !m || !n || problem < 0 || problem >= P->count() || level < 0 ||
      level >= int(P->probs[problem].lv.size())) {
    set_error("bad problem / level index");
    return GDP_EINVAL;
  }
  *m = P->probs[problem].lv[level].m;
  *n = P->probs[problem].lv[level].n;
  return GDP_OK;
}

int gdp_plan_set_mask(gdp_plan* P, int problem, const uint8_t* mask) {
  if (!P || problem < 0 || problem >= P->count()) {
    set_error("gdp_plan_set_mask: bad plan or problem");
    return GDP_EINVAL;
  }
  if (mask && P->probs[problem].p.boundary != GDP_DIRICHLET) {
    set_error("gdp_plan_set_mask: a masked domain needs the Dirichlet boundary");
    return GDP_EINVAL;
  }
  P->probs[problem].mask = mask;
  if (P->bound) P->bound = false;  // tables must be rebuilt by gdp_plan_bind
  destroy_graphs(P);
  return GDP_OK;
}

int gdp_plan_set_tile(gdp_plan* P, int kind, int level, int menu_index) {
  if (!P || kind < 0 || kind > 2 || menu_index < 0 || menu_index >= menu_size(kind) ||
      (kind < 2 && (level < 0 || level >= kMaxDepth))) {
    set_error("gdp_plan_set_tile: bad argument");
    return GDP_EINVAL;
  }
  if (kind == 0 && level > 0 && menu_index >= kFusedDown) {
    set_error("gdp_plan_set_tile: down entries %d and up carry the PCG update: finest level only",
              kFusedDown);
    return GDP_EINVAL;
  }
  if (kind == 0) P->down_menu[level] = menu_index;
  else if (kind == 1) P->up_menu[level] = menu_index;
  else P->kry_menu = menu_index;
  if (P->bound) P->bound = false;  // tables must be rebuilt by gdp_plan_bind
  destroy_graphs(P);
  return GDP_OK;
}

int gdp_plan_destroy(gdp_plan* P) {
  if (!P) return GDP_OK;
  destroy_graphs(P);
  if (P->cap) (void)hipStreamDestroy(P->cap);
  if (P->io_ev) (void)hipEventDestroy(P->io_ev);
  for (auto& r : P->prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : P->prof_pool) (void)hipEventDestroy(e);
  if (P->h_block) pinned_put(P->h_block, P->h_block_bytes);
  delete P;
  return GDP_OK;
}

int gdp_plan_bind(gdp_plan* P, void* workspace, size_t bytes) {
  if (!P || !workspace) {
    set_error("null argument");
    return GDP_EINVAL;
  }
  if (bytes < P->ws_bytes) {
    set_error("workspace of %zu bytes is smaller than the %zu the plan needs", bytes, P->ws_bytes);
    return GDP_ENOMEM;
  }
  if ((reinterpret_cast<uintptr_t>(workspace) & 255) != 0) {
    set_error("workspace must be 256-byte aligned");
    return GDP_EINVAL;
  }
  P->ws = static_cast<char*>(workspace);
  P->io_valid = false;  // the pointer tables live in the workspace
  {
    int dev = 0;
    GDP_HIP(hipGetDevice(&dev));
    GDP_HIP(hipDeviceGetAttribute(&P->n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  const int NP = P->count();
  if (!P->h_block) {
    auto a64 = [](size_t x) { return (x + 63) & ~size_t(63); };
    const size_t b_io = a64(sizeof(IoDesc) * NP), b_init = a64(sizeof(double*) * NP);
    P->h_block_bytes = b_io + b_init + state_span(NP);
    GDP_HIP(pinned_get(P->h_block_bytes, &P->h_block));
    char* hb = static_cast<char*>(P->h_block);
    P->h_io = reinterpret_cast<IoDesc*>(hb);
    P->h_init = reinterpret_cast<const double**>(hb + b_io);
    P->h_state = reinterpret_cast<ProblemState*>(hb + b_io + b_init);
  }
  if (!P->cap) GDP_HIP(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
  if (!P->io_ev) GDP_HIP(hipEventCreateWithFlags(&P->io_ev, hipEventDisableTiming));
  // Host-to-device uploads of the tables, on the plan's own non-blocking stream (never
  // the null stream, which other threads' graph captures forbid), each finished before
  // its host buffer can go away.
  auto h2d = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, P->cap);
    return e == hipSuccess ? hipStreamSynchronize(P->cap) : e;
  };

  // ---- descriptors
  std::vector<LevelDesc> levels(P->total_levels);
  std::vector<KrylovDesc> kry(NP);
  std::vector<TailDesc> tails(NP);
  std::vector<double> dct_host;
  for (int p = 0; p < NP; ++p) {
    ProbInfo& pi = P->probs[p];
    const int L = int(pi.lv.size());
    for (int l = 0; l < L; ++l) {
      const LevelInfo& li = pi.lv[l];
      LevelDesc& d = levels[pi.level_base + l];
      memset(&d, 0, sizeof(d));
      d.u = P->ws + li.off_u;
      d.v = P->ws + li.off_v;
      d.b = P->ws + li.off_b;
      d.m = li.m;
      d.n = li.n;
      d.pitch = li.pitch;
      d.bc = pi.p.boundary;
      d.h2 = li.h * li.h;
      d.inv_h2 = 1.0 / d.h2;
      d.coarse = l + 1 < L ? pi.level_base + l + 1 : -1;
      d.problem = p;
      d.level = l;
    }
    KrylovDesc& k = kry[p];
    memset(&k, 0, sizeof(k));
    k.x = P->at<double>(pi.off_x);
    k.r = P->at<double>(pi.off_r);
    k.d[0] = P->at<double>(pi.off_d0);
    k.d[1] = P->at<double>(pi.off_d1);
    k.q = P->at<double>(pi.off_q);
    k.rhs = nullptr;
    k.m = pi.lv[0].m;
    k.n = pi.lv[0].n;
    k.pitch = pi.lv[0].pitch;
    k.bc = pi.p.boundary;
    k.h2 = pi.p.spacing * pi.p.spacing;
    k.inv_h2 = 1.0 / k.h2;
    k.level0 = pi.level_base;
    TailDesc& t = tails[p];
    memset(&t, 0, sizeof(t));
    t.first = pi.level_base + pi.tail_first;
    t.nlev = L - pi.tail_first;
    t.problem = p;
    t.krylov = pi.tail_first == 0 ? p : -1;
    for (int l = 0; l < t.nlev; ++l) {
      t.lds_off[l][0] = pi.tail_off[l][0];
      t.lds_off[l][1] = pi.tail_off[l][1];
    }
    t.scratch_off = pi.scratch_off;
    t.coarse_off = pi.coarse_off;
    t.tab_off = pi.tab_off;
    t.dct = nullptr;
    if (pi.exact) {
      const int m = pi.lv.back().m, n = pi.lv.back().n;
      std::vector<double> tab(size_t(m * m + n * n + m * n));
      const bool neu = pi.p.boundary == GDP_NEUMANN;
      // orthonormal DCT-II (Neumann) or DST-I (Dirichlet) matrices, row k = mode k
      auto fill = [neu](double* C, int s) {
        for (int k = 0; k < s; ++k)
          for (int i = 0; i < s; ++i) {
            if (neu) {
              const double sc = k == 0 ? std::sqrt(1.0 / s) : std::sqrt(2.0 / s);
              C[k * s + i] = sc * std::cos(M_PI * (2.0 * i + 1.0) * k / (2.0 * s));
            } else {
              C[k * s + i] = std::sqrt(2.0 / (s + 1)) *
                             std::sin(M_PI * (i + 1.0) * (k + 1.0) / (s + 1.0));
            }
          }
      };
      fill(tab.data(), m);
      fill(tab.data() + m * m, n);
      double* il = tab.data() + m * m + n * n;
      for (int k = 0; k < m; ++k)
        for (int l = 0; l < n; ++l) {
          double lam;
          if (neu)
            lam = 2.0 * std::cos(M_PI * k / m) + 2.0 * std::cos(M_PI * l / n) - 4.0;
          else
            lam = 2.0 * std::cos(M_PI * (k + 1.0) / (m + 1.0)) +
                  2.0 * std::cos(M_PI * (l + 1.0) / (n + 1.0)) - 4.0;
          il[k * n + l] = (neu && k == 0 && l == 0) ? 0.0 : 1.0 / lam;
        }
      t.dct = P->at<double>(pi.off_dct);
      GDP_HIP(h2d(P->ws + pi.off_dct, tab.data(), tab.size() * 8));
    }
  }
  GDP_HIP(h2d(P->d_levels(), levels.data(), sizeof(LevelDesc) * levels.size()));
  GDP_HIP(h2d(P->d_kry(), kry.data(), sizeof(KrylovDesc) * NP));
  {
    // strip-tail problems first (k_stail takes entries [0, n_stail), k_tail the rest)
    std::vector<TailDesc> ordered;
    for (int p = 0; p < NP; ++p)
      if (P->probs[p].stail) ordered.push_back(tails[p]);
    for (int p = 0; p < NP; ++p)
      if (!P->probs[p].stail) ordered.push_back(tails[p]);
    GDP_HIP(h2d(P->d_tails(), ordered.data(), sizeof(TailDesc) * NP));
  }

  // ---- tile tables (self-contained records; see VTile / KTile)
  size_t toff = P->off_tiles;
  const size_t tend = P->off_tiles + P->tiles_bytes;
  auto upload = [&](Table& tb, const void* data, size_t rec, size_t count) -> int {
    tb.off = toff;
    tb.count = int(count);
    const size_t b = align256(rec * std::max<size_t>(count, 1));
    if (toff + rec * count > tend) {
      set_error("internal: tile table overflow");
      return GDP_EINVAL;
    }
    if (count && h2d(P->ws + toff, data, rec * count) != hipSuccess) {
      set_error("hipMemcpy of a tile table failed");
      return GDP_EHIP;
    }
    toff += std::min(b, tend - toff);
    return GDP_OK;
  };
  // Krylov slot ranges (fixed by the Krylov tile shape)
  std::vector<SlotRange> rk(NP), ru(NP), ra(NP), rd(NP);
  {
    const Shape2 s = kKryMenu[P->kry_menu];
    int c = 0, ca = 0;
    for (int p = 0; p < NP; ++p) {
      const auto& l0 = P->probs[p].lv[0];
      rk[p].begin = c;
      c += cdiv(l0.m, s.h) * cdiv(l0.n, s.w);
      rk[p].count = c - rk[p].begin;
      ra[p].begin = ca;  // the strip apply's tiles, in the order the table lists them
      ca += apply_tiles(l0.m, l0.n);
      ra[p].count = ca - ra[p].begin;
    }
  }
  auto vtile = [&](int p, int l, int i0, int j0, int slot, int lead) {
    const ProbInfo& pi = P->probs[p];
    const LevelInfo& f = pi.lv[l];
    VTile t;
    memset(&t, 0, sizeof(t));
    t.u = P->ws + f.off_u;
    t.v = P->ws + f.off_v;
    t.b = P->ws + f.off_b;
    if (l + 1 < int(pi.lv.size())) {
      const LevelInfo& c = pi.lv[l + 1];
      t.cu = P->ws + c.off_u;
      t.cb = P->ws + c.off_b;
      t.cm = c.m;
      t.cn = c.n;
      t.cpitch = c.pitch;
    }
    t.kr = P->at<double>(pi.off_r);
    t.kq = P->at<double>(pi.off_q);
    t.h2 = f.h * f.h;
    t.inv_h2 = 1.0 / t.h2;
    t.m = f.m;
    t.n = f.n;
    t.pitch = f.pitch;
    t.bc = pi.p.boundary;
    t.level = l;
    t.problem = p;
    t.i0 = i0;
    t.j0 = j0;
    t.slot = slot;
    t.lead = lead;
    t.rbeg = rk[p].begin;
    t.rcnt = rk[p].count;
    t.abeg = ra[p].begin;
    t.acnt = ra[p].count;
    if (pi.mask) t.mask = l == 0 ? pi.mask : reinterpret_cast<const uint8_t*>(P->ws + f.off_mask);
    return t;
  };
  if (has_mask(P)) {
    // coarse masks by injection: coarse point (I, J) is an unknown when its coincident
    // fine point (2I + 1, 2J + 1) is
    for (const ProbInfo& pi : P->probs) {
      if (!pi.mask) continue;
      for (size_t l = 1; l < pi.lv.size(); ++l) {
        const LevelInfo& f = pi.lv[l - 1];
        const LevelInfo& c = pi.lv[l];
        const uint8_t* src = l == 1 ? pi.mask : reinterpret_cast<const uint8_t*>(P->ws + f.off_mask);
        klaunch(k_mask_inject, dim3((c.m * c.n + 255) / 256), dim3(256), 0, P->cap,
            src, f.n, reinterpret_cast<uint8_t*>(P->ws + c.off_mask), c.m, c.n);
        GDP_HIP(hipGetLastError());
      }
    }
  }
  for (int l = 0; l < P->depth; ++l) {
    std::vector<VTile> vd, vu;
    const Shape2 sd = menu_shape(0, down_menu_of(P, l), P->opt.pre_smoothing);
    const Shape2 su = menu_shape(1, up_menu_of(P, l), P->opt.post_smoothing);
    for (int p = 0; p < NP; ++p) {
      const ProbInfo& pi = P->probs[p];
      if (pi.tail_first <= l) continue;
      const auto& f = pi.lv[l];
      const auto& c = pi.lv[l + 1];
      const size_t d0 = vd.size();
      for (int i = 0; i < c.m; i += sd.h)
        for (int j = 0; j < c.n; j += sd.w)
          vd.push_back(vtile(p, l, i, j, int(vd.size()), vd.size() == d0));
      if (l == 0) rd[p] = SlotRange{int(d0), int(vd.size() - d0)};
      if (l == 0) ru[p].begin = int(vu.size());
      for (int i = 0; i < f.m; i += su.h)
        for (int j = 0; j < f.n; j += su.w) vu.push_back(vtile(p, l, i, j, int(vu.size()), 0));
      if (l == 0) ru[p].count = int(vu.size()) - ru[p].begin;
    }
    const auto swd = xcd_swizzle(vd);
    const auto swu = xcd_swizzle(vu);
    if (int rc = upload(P->down[l], swd.data(), sizeof(VTile), swd.size())) return rc;
    if (int rc = upload(P->up[l], swu.data(), sizeof(VTile), swu.size())) return rc;
  }
  {
    // problems whose finest level is in the tail get one direct <r,z> slot each,
    // after the slots of the finest up-kernel table
    int zb = P->depth > 0 ? P->up[0].count : 0;
    for (int p = 0; p < NP; ++p)
      if (P->probs[p].tail_first == 0) ru[p] = SlotRange{zb++, 1};
  }
  {
    std::vector<KTile> v;
    const Shape2 s = kKryMenu[P->kry_menu];
    for (int p = 0; p < NP; ++p) {
      const ProbInfo& pi = P->probs[p];
      const auto& l0 = pi.lv[0];
      for (int i = 0; i < l0.m; i += s.h)
        for (int j = 0; j < l0.n; j += s.w) {
          KTile t;
          memset(&t, 0, sizeof(t));
          t.x = P->at<double>(pi.off_x);
          t.r = P->at<double>(pi.off_r);
          t.d0 = P->at<double>(pi.off_d0);
          t.d1 = P->at<double>(pi.off_d1);
          t.q = P->at<double>(pi.off_q);
          t.z = P->ws + l0.off_u;
          t.b0 = P->ws + l0.off_b;
          t.inv_h2 = 1.0 / (pi.p.spacing * pi.p.spacing);
          t.m = l0.m;
          t.n = l0.n;
          t.pitch = l0.pitch;
          t.bc = pi.p.boundary;
          t.problem = p;
          t.i0 = i;
          t.j0 = j;
          t.slot = int(v.size());
          t.lead = int(v.size()) == rk[p].begin;
          t.kbeg = rk[p].begin;
          t.kcnt = rk[p].count;
          t.zbeg = ru[p].begin;
          t.zcnt = ru[p].count;
          t.abeg = ra[p].begin;
          t.acnt = ra[p].count;
          t.mask = pi.mask;
          v.push_back(t);
        }
    }
    // the strip apply's tiles: their own <d, q> slots, every other field as above
    std::vector<KTile> va;
    for (int p = 0; p < NP; ++p) {
      const ProbInfo& pi = P->probs[p];
      const auto& l0 = pi.lv[0];
      for (int i = 0; i < l0.m; i += kApplyBlockRows)
        for (int j = 1; j == 1 || j <= l0.n - 1; j += kApplyCols) {
          KTile t;
          memset(&t, 0, sizeof(t));
          t.x = P->at<double>(pi.off_x);
          t.r = P->at<double>(pi.off_r);
          t.d0 = P->at<double>(pi.off_d0);
          t.d1 = P->at<double>(pi.off_d1);
          t.q = P->at<double>(pi.off_q);
          t.z = P->ws + l0.off_u;
          t.b0 = P->ws + l0.off_b;
          t.inv_h2 = 1.0 / (pi.p.spacing * pi.p.spacing);
          t.m = l0.m;
          t.n = l0.n;
          t.pitch = l0.pitch;
          t.bc = pi.p.boundary;
          t.problem = p;
          t.i0 = i;
          t.j0 = j;
          t.slot = int(va.size());
          t.lead = int(va.size()) == ra[p].begin;
          t.kbeg = rk[p].begin;
          t.kcnt = rk[p].count;
          t.zbeg = ru[p].begin;
          t.zcnt = ru[p].count;
          t.abeg = ra[p].begin;
          t.acnt = ra[p].count;
          t.mask = pi.mask;
          va.push_back(t);
        }
    }
    const auto sw = xcd_swizzle(v);
    if (int rc = upload(P->kry, sw.data(), sizeof(KTile), sw.size())) return rc;
    const auto swa = xcd_swizzle(va);
    if (int rc = upload(P->kapply, swa.data(), sizeof(KTile), swa.size())) return rc;
  }
  GDP_HIP(h2d(P->d_rkry(), rk.data(), sizeof(SlotRange) * NP));
  GDP_HIP(h2d(P->d_rup0(), ru.data(), sizeof(SlotRange) * NP));
  GDP_HIP(h2d(P->d_rdown0(), rd.data(), sizeof(SlotRange) * NP));
  GDP_HIP(hipMemsetAsync(P->d_state(), 0, sizeof(ProblemState) * NP, P->cap));
  for (size_t o : {P->off_apart, P->off_upart, P->off_zpart, P->off_spart})
    GDP_HIP(hipMemsetAsync(P->ws + o, 0, sizeof(double2) * P->max_slots, P->cap));

  // the tail kernels may use more than the default dynamic LDS
  if (P->n_stail) GDP_HIP(raise_smem(P->fp32 ? 0 : 1, P->stail_smem));
  if (P->tail_smem) GDP_HIP(raise_smem(P->fp32 ? 2 : 3, P->tail_smem));
  // the uploads and clears above ran on the plan's own stream: finish them before any
  // caller stream uses the tables
  GDP_HIP(hipStreamSynchronize(P->cap));
  destroy_graphs(P);
  P->bound = true;
  return GDP_OK;
}

}  // extern "C"

// ----------------------------------------------------------------- launch sequences
namespace {

// Algorithmic bytes per fine point of the Krylov grid (p = V-cycle element size).
double krylov_points(const gdp_plan* P) {
  double n = 0.0;
  for (const auto& pi : P->probs) n += double(pi.lv[0].m) * pi.lv[0].n;
  return n;
}

double level_points(const gdp_plan* P, int l) {
  double n = 0.0;
  for (const auto& pi : P->probs)
    if (pi.tail_first > l) n += double(pi.lv[l].m) * pi.lv[l].n;
  return n;
}

int prof_id(gdp_plan* P, const std::string& name) {
  for (size_t i = 0; i < P->prof_agg.size(); ++i)
    if (P->prof_agg[i].name == name) return int(i);
  P->prof_agg.push_back(gdp_plan::ProfAgg{name});
  return int(P->prof_agg.size()) - 1;
}

hipEvent_t prof_event(gdp_plan* P) {
  if (!P->prof_pool.empty()) {
    hipEvent_t e = P->prof_pool.back();
    P->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// A profiled launch (nothing when not profiling): the scope's kernel launches (klaunch)
// carry its start and stop events, stamped from the dispatches themselves.
struct ProfScope {
  gdp_plan* P;
  int id = -1;
  double bytes;
  LaunchEvents ev;
  ProfScope(gdp_plan* P_, hipStream_t, const char* name, double bytes_) : P(P_), bytes(bytes_) {
    if (!P->prof) return;
    id = prof_id(P, name);
    ev.start = prof_event(P);
    ev.stop = prof_event(P);
    t_launch_events = &ev;
  }
  ~ProfScope() {
    if (id < 0) return;
    t_launch_events = nullptr;
    if (ev.used) {
      P->prof_pending.push_back(gdp_plan::ProfRec{id, ev.start, ev.stop, bytes});
    } else {  // nothing was launched (an empty level)
      P->prof_pool.push_back(ev.start);
      P->prof_pool.push_back(ev.stop);
    }
  }
};

void prof_collect(gdp_plan* P) {
  for (auto& r : P->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      auto& g = P->prof_agg[r.id];
      g.ms += ms;
      g.bytes += r.bytes;
      g.launches += 1;
    }
    P->prof_pool.push_back(r.a);
    P->prof_pool.push_back(r.b);
  }
  P->prof_pending.clear();
}

// Whether k_update leaves its x half (read x and d, write x: 24 B per Krylov point) to
// extra blocks of the strip-tail launch.  Those blocks run on the CUs the one block per
// tail problem leaves free, while the latency-bound tail runs (about 20 us); one CU
// streams about 40 GB/s, so 1 MiB per free CU is what they finish in that time.  With
// more, the x update would outlast the tail on too few CUs (a batch of about as many
// problems as CUs), and the whole-chip k_update does it instead.
constexpr double kXBytesPerFreeCu = 1048576.0;
// The x-update blocks of an FP32 tail launch each take about this much of the x stream:
// fewer blocks than free CUs where the stream is small.  The stream's rate saturates
// well below one block per free CU, and its traffic slows the tail's own latency-bound
// chain: at 1024x1024 RGB mixed (75 MB) 128-152 blocks made the tail launch 1.1-1.8 us
// shorter than 253, and 108 or 64 blocks made it 2 or 13 us longer.  An FP64 plan's
// tail starts one level coarser and is done sooner, so the stream has to be: there 144
// blocks made the launch 3 us longer than one per free CU, which it keeps.
constexpr double kXBytesPerBlock = 524288.0;
bool defer_x_update(const gdp_plan* P) {
  const int free_cu = P->n_cu - P->n_stail;
  if (P->n_stail <= 0 || free_cu <= 0) return false;
  return 24.0 * krylov_points(P) <= kXBytesPerFreeCu * free_cu;
}

// xupd: the previous k_update deferred its x half (defer_x); the strip-tail launch adds
// it in blocks of its own.  fused (fused_update plans, PCG iterations that defer x): the
// finest down kernel carries the iteration's update (in place of k_update, dcur its
// direction buffer), the tail kernels test its convergence at their entry (from that
// kernel's partials, with their own loads in flight) and the finest up kernel stores its
// r.
template <typename T>
int enqueue_vcycle_t(gdp_plan* P, hipStream_t s, bool dots, bool check_conv,
                     bool xupd = false, bool fused = false, int dcur = 0) {
  const gdp_options& o = P->opt;
  for (int l = 0; l < P->depth; ++l) {
    if (!P->down[l].count) continue;
    char nm[32];
    const bool fl = fused && l == 0;
    snprintf(nm, sizeof nm, fl ? "down L%d+update" : "down L%d", l);
    ProfScope ps(P, s, nm, ((fl ? 16.0 : 0.0) + 2.25 * sizeof(T)) * level_points(P, l));
    const int cc = check_conv && !fused ? 1 : 0;
    DownArgs a{P->at<VTile>(P->down[l].off), P->d_state(), P->d_upart(), cc,
               has_mask(P) ? 1 : 0};
    if (fl) {
      a.check_conv = 0;
      a.apart = P->d_apart();
      a.dcur = dcur;
      const int e = down_menu_of(P, 0);
      if (e >= kFusedDownBlk)
        launch_down_blk_fused<T>(8 << (e - kFusedDownBlk), o.pre_smoothing, P->down[l].count, s,
                                 a);
      else
        launch_down_fused<T>(8 << (e - kFusedDown), o.pre_smoothing, P->down[l].count, s, a);
    } else {
      launch_down<T>(down_menu_of(P, l), o.pre_smoothing, P->down[l].count, s, a);
    }
    GDP_CHECK_LAUNCH();
  }
  // Neumann problems use the symmetric cycle (post-smoothing colour 1 then 0); a plan
  // mixing boundary types uses the order of its first problem for the whole batch.
  const int order = P->probs[0].p.boundary == GDP_NEUMANN ? 1 : 0;
  double tail_pts = 0.0;
  for (const auto& pi : P->probs)
    tail_pts += double(pi.lv[pi.tail_first].m) * pi.lv[pi.tail_first].n;
  {
    ProfScope ps(P, s, "tail",
                 2.0 * sizeof(T) * tail_pts + (xupd ? 24.0 * krylov_points(P) : 0.0));
    // 1: the problems whose finest level is in the tail test the update's convergence
    // (k_update's partials); 2: every problem does (the fused update-down kernel's)
    const int cc = !check_conv ? 0 : (fused ? 2 : 1);
    const SlotRange* rk = fused ? P->d_rdown0() : P->d_rkry();
    const int ns = P->n_stail, no = P->count() - P->n_stail;
    if (ns) {
      TailArgs a{P->d_tails(), P->d_levels(), P->d_state(), P->d_kry(), P->d_upart(),
                 rk, P->d_zpart(), P->d_rup0(), cc, o.pre_smoothing,
                 o.post_smoothing, P->csweeps, order};
      if (xupd) {
        // one block per kXBytesPerBlock of the x stream (FP32 tails), on CUs the tail
        // blocks leave free, at most one per kXItems items
        const Shape2 ks = kKryMenu[P->kry_menu];
        const int items =
            P->kry.count * ((ks.h * ks.w / 2 + kStailBlockThreads - 1) / kStailBlockThreads);
        const int by_bytes =
            P->fp32 ? int(24.0 * krylov_points(P) / kXBytesPerBlock) + 1 : P->n_cu;
        a.xblocks = std::max(1, std::min(std::min(P->n_cu - ns, by_bytes),
                                         (items + kStailXItems - 1) / kStailXItems));
        a.xt = P->at<KTile>(P->kry.off);
        a.nxt = P->kry.count;
        a.xth = ks.h;
        a.xtw = ks.w;
      }
      launch_stail<T>(P->stail_spec, dots, ns, P->stail_smem, s, a);
      GDP_CHECK_LAUNCH();
    }
    if (no) {
      if (dots)
        klaunch(k_tail<T, true>, dim3(no), dim3(kTailThreads), P->tail_smem, s,
            P->d_tails() + ns, P->d_levels(), P->d_state(), P->d_kry(), P->d_upart(),
            rk, P->d_zpart(), P->d_rup0(), cc, o.pre_smoothing, o.post_smoothing,
            P->csweeps, order);
      else
        klaunch(k_tail<T, false>, dim3(no), dim3(kTailThreads), P->tail_smem, s,
            P->d_tails() + ns, P->d_levels(), P->d_state(), P->d_kry(), P->d_upart(),
            rk, P->d_zpart(), P->d_rup0(), cc, o.pre_smoothing, o.post_smoothing,
            P->csweeps, order);
      GDP_CHECK_LAUNCH();
    }
  }
  for (int l = P->depth - 1; l >= 0; --l) {
    if (!P->up[l].count) continue;
    char nm[32];
    const bool dl = dots && l == 0;
    const bool ru = dl && fused;  // + the r store the fused update left (8 B a point)
    snprintf(nm, sizeof nm, dl ? "up L%d+dots" : "up L%d", l);
    ProfScope ps(P, s, nm,
                 (3.25 * sizeof(T) + (dl ? 16.0 : 0.0) + (ru ? 8.0 : 0.0)) * level_points(P, l));
    UpArgs a{P->at<VTile>(P->up[l].off), P->d_state(), P->d_zpart(), order,
             has_mask(P) ? 1 : 0};
    a.rupd = ru ? 1 : 0;
    launch_up<T>(up_menu_of(P, l), o.post_smoothing, dl, P->up[l].count, s, a);
    GDP_CHECK_LAUNCH();
  }
  return GDP_OK;
}

int enqueue_vcycle(gdp_plan* P, hipStream_t s, bool dots, bool check_conv) {
  return P->fp32 ? enqueue_vcycle_t<float>(P, s, dots, check_conv)
                 : enqueue_vcycle_t<double>(P, s, dots, check_conv);
}

template <typename T>
int enqueue_jacobi_t(gdp_plan* P, hipStream_t s, bool check_conv) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  ProfScope ps(P, s, "jacobi", (8.0 + sizeof(T)) * krylov_points(P));
  GDP_KRY_DISPATCH(P->kry_menu,
                   (klaunch(k_jacobi<T, TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                       t, P->d_state(), P->d_upart(), P->d_zpart(), check_conv ? 1 : 0)));
  GDP_CHECK_LAUNCH();
  return GDP_OK;
}

// One PCG iteration: direction and operator, update, preconditioner.  last: the
// preconditioner is left out, and the update's convergence test runs in the check that
// follows (the host runs the preconditioner before continuing if a problem is still
// open).  next_apply: the next
// kernel sequence on the stream starts with another iteration's apply (no true-residual
// check in between).  x_pending: the x updates the preceding update left to this apply,
// 0, 1 (which the apply holds back for the next update) or 2 (which it adds to x); only
// the profiler's byte counts use it, the kernels read the state.  Returns through
// *deferred the same for the next apply.
template <typename T>
int enqueue_pcg_iter_t(gdp_plan* P, hipStream_t s, int din, bool jacobi, bool last,
                       bool next_apply, int x_pending, int* deferred) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  const double N = krylov_points(P);
  // Where the x half of this iteration's update goes: to the strip-tail launch of its
  // V-cycle when the tails leave enough CUs idle (defer_x_update), else to the next
  // iteration's apply when one follows before any true-residual check; k_update does it
  // itself after the last update of a batch, for Jacobi and when no apply follows.
  const bool tail_mode = defer_x_update(P);
  const bool apply_mode = !jacobi && !tail_mode;  // this apply adds a pending x update
  const bool tail_x = !jacobi && !last && tail_mode;
  const bool apply_x = apply_mode && !last && next_apply;
  // the Jacobi preconditioner publishes its <r, z> partials with the Krylov tiles
  {
    // with a double x update: x read and written, the held direction read
    const double xb = apply_mode && x_pending == 2 ? 24.0 : 0.0;
    ProfScope ps(P, s, "apply", (24.0 + sizeof(T) + xb) * N);
    const KTile* at = P->at<KTile>(P->kapply.off);
    const int nt = P->kapply.count, jc = jacobi ? 1 : 0;
    if (has_mask(P)) {
      if (apply_mode)
        klaunch(k_apply_s<T, true, true>, dim3(nt), dim3(kThreads), 0, s, at, P->d_state(),
                P->d_zpart(), P->d_apart(), din, jc);
      else
        klaunch(k_apply_s<T, true, false>, dim3(nt), dim3(kThreads), 0, s, at, P->d_state(),
                P->d_zpart(), P->d_apart(), din, jc);
    } else {
      if (apply_mode)
        klaunch(k_apply_s<T, false, true>, dim3(nt), dim3(kThreads), 0, s, at, P->d_state(),
                P->d_zpart(), P->d_apart(), din, jc);
      else
        klaunch(k_apply_s<T, false, false>, dim3(nt), dim3(kThreads), 0, s, at, P->d_state(),
                P->d_zpart(), P->d_apart(), din, jc);
    }
    GDP_CHECK_LAUNCH();
  }
  const bool defer_x = tail_x || apply_x;
  // an update after an apply that held one back: both to the next apply, or both to x
  const bool held = apply_mode && x_pending == 1;
  if (deferred) *deferred = apply_x ? (held ? 2 : 1) : 0;
  // the update in the finest down kernel of the V-cycle that follows it
  if (defer_x && fused_update(P)) return enqueue_vcycle_t<T>(P, s, true, true, tail_x, true, din ^ 1);
  {
    ProfScope ps(P, s, defer_x ? "update (x deferred)" : "update",
                 ((defer_x ? 24.0 : held ? 56.0 : 48.0) + sizeof(T)) * N);
    if (defer_x) {
      GDP_KRY_DISPATCH(P->kry_menu,
                       (klaunch(k_update<T, TH, TW, true>, dim3(nb), dim3(kThreads), 0, s, t,
                                P->d_state(), P->d_apart(), P->d_upart(), din ^ 1)));
    } else {
      GDP_KRY_DISPATCH(P->kry_menu,
                       (klaunch(k_update<T, TH, TW, false>, dim3(nb), dim3(kThreads), 0, s, t,
                                P->d_state(), P->d_apart(), P->d_upart(), din ^ 1)));
    }
    GDP_CHECK_LAUNCH();
  }
  if (last) return GDP_OK;  // the check that follows tests its convergence
  if (jacobi) return enqueue_jacobi_t<T>(P, s, true);
  return enqueue_vcycle_t<T>(P, s, true, true, tail_x);
}

// Two PCG iterations.  more: another pair follows in the same launch sequence, so the
// second iteration's update may leave its x half to that pair's first apply; a pair
// followed by the true-residual check leaves x complete.  *pending: in, the x updates
// the preceding update left to this pair's first apply (enqueue_pcg_iter_t); out, the
// same for the next pair.
int enqueue_pcg_pair(gdp_plan* P, hipStream_t s, bool jacobi, bool last = false,
                     bool more = false, int* pending = nullptr) {
  int rc;
  const bool more2 = more && !last;
  int in = pending ? *pending : 0, mid = 0, out = 0;
  if (P->fp32) {
    if ((rc = enqueue_pcg_iter_t<float>(P, s, 0, jacobi, false, true, in, &mid))) return rc;
    rc = enqueue_pcg_iter_t<float>(P, s, 1, jacobi, last, more2, mid, &out);
  } else {
    if ((rc = enqueue_pcg_iter_t<double>(P, s, 0, jacobi, false, true, in, &mid))) return rc;
    rc = enqueue_pcg_iter_t<double>(P, s, 1, jacobi, last, more2, mid, &out);
  }
  if (pending) *pending = out;
  return rc;
}

struct ScalarArgs {
  int kind = 1, mode = 0;
  double tol = 0.0;
  int max_iter = 0, zero_done = 0, respect_done = 0;
  bool hist = false;
  int reopen = 0;  // kind 2 of a PCG check: the restart decision
};

int enqueue_scalar(gdp_plan* P, hipStream_t s, const ScalarArgs& a) {
  ProfScope ps(P, s, a.kind == 1 ? "scalar init" : "scalar residual", 16.0 * P->kry.count);
  klaunch(k_scalar, dim3(P->count()), dim3(kThreads), 0, s,
      P->d_kry(), P->d_state(), P->d_spart(), P->d_rkry(), a.kind, a.mode, a.tol,
      a.max_iter, P->opt.flexible, a.zero_done, a.respect_done, a.hist ? P->d_hist() : nullptr,
      P->opt.max_iterations + 2, a.reopen);
  GDP_CHECK_LAUNCH();
  return GDP_OK;
}

template <typename T>
int enqueue_init_t(gdp_plan* P, hipStream_t s, int mode, int project, double tol, int max_it,
                   bool with_init, int zero_done) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  const double N = krylov_points(P);
  {
    // the partial sums of b go to the apply's partial-sum slots, which no kernel reads
    // before the first apply writes them, so that k_init can reduce them while it writes
    // its own to spart
    ProfScope ps(P, s, "init_stats", 8.0 * N);
    GDP_KRY_DISPATCH(P->kry_menu, (klaunch(k_init_stats<TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                                      t, P->d_io(), P->d_apart())));
    GDP_CHECK_LAUNCH();
  }
  {
    ProfScope ps(P, s, "init", (mode == 0 ? 24.0 + sizeof(T) : 16.0) * N);
    GDP_KRY_DISPATCH(P->kry_menu,
                     (klaunch(k_init<T, TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                         t, P->d_io(), with_init ? P->d_init() : nullptr, P->d_state(),
                         P->d_spart(), mode, P->d_apart(), P->d_rkry(), project)));
    GDP_CHECK_LAUNCH();
  }
  ScalarArgs a;
  a.kind = 1;
  a.tol = tol;
  a.max_iter = max_it;
  a.zero_done = zero_done;
  return enqueue_scalar(P, s, a);
}

template <typename T>
int enqueue_residual_t(gdp_plan* P, hipStream_t s, int mode, int respect_done, bool hist,
                       bool reopen, bool check_conv) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  {
    ProfScope ps(P, s, "residual", (24.0 + sizeof(T)) * krylov_points(P));
    GDP_KRY_DISPATCH(P->kry_menu,
                     (klaunch(k_residual<T, TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                         t, P->d_io(), P->d_state(), P->d_spart(), mode, respect_done,
                         P->d_upart(), P->d_rkry(), check_conv ? 1 : 0)));
    GDP_CHECK_LAUNCH();
  }
  ScalarArgs a;
  a.kind = 2;
  a.mode = mode;
  a.tol = P->opt.tolerance;
  a.respect_done = respect_done;
  a.hist = hist;
  a.reopen = reopen ? 1 : 0;
  return enqueue_scalar(P, s, a);
}

template <typename T>
int enqueue_correct_t(gdp_plan* P, hipStream_t s) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  ProfScope ps(P, s, "correct", (16.0 + sizeof(T)) * krylov_points(P));
  GDP_KRY_DISPATCH(P->kry_menu, (klaunch(k_correct<T, TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                                    t, P->d_state())));
  GDP_CHECK_LAUNCH();
  return GDP_OK;
}

int enqueue_finalize(gdp_plan* P, hipStream_t s, int mean_fix, int only_checked = 0) {
  const int nb = P->kry.count;
  const KTile* t = P->at<KTile>(P->kry.off);
  ProfScope ps(P, s, "finalize", 16.0 * krylov_points(P));
  GDP_KRY_DISPATCH(P->kry_menu, (klaunch(k_finalize<TH, TW>, dim3(nb), dim3(kThreads), 0, s,
                                    t, P->d_io(), P->d_state(), mean_fix, only_checked)));
  GDP_CHECK_LAUNCH();
  return GDP_OK;
}

#define GDP_T(call) (P->fp32 ? call<float> : call<double>)

int stage_io(gdp_plan* P, const double* const* b, double* const* u, const double* const* init,
             hipStream_t s) {
  const int NP = P->count();
  for (int p = 0; p < NP; ++p)
    if (!b[p] || !u[p]) {
      set_error("problem %d: null b or u pointer", p);
      return GDP_EINVAL;
    }
  // The same buffers as last time (a solve loop reusing its arrays): the device copy is
  // current, so only order this stream after the upload that wrote it.
  bool same = P->io_valid && !init;
  for (int p = 0; same && p < NP; ++p) same = P->h_io[p].b == b[p] && P->h_io[p].u == u[p];
  if (same) {
    GDP_HIP(hipStreamWaitEvent(s, P->io_ev, 0));
    return GDP_OK;
  }
  // the pinned staging buffers are reused: wait until the previous upload has run
  GDP_HIP(hipEventSynchronize(P->io_ev));
  for (int p = 0; p < NP; ++p) {
    P->h_io[p].b = b[p];
    P->h_io[p].u = u[p];
    if (init) P->h_init[p] = init[p];
  }
  GDP_HIP(hipMemcpyAsync(P->d_io(), P->h_io, sizeof(IoDesc) * NP, hipMemcpyHostToDevice, s));
  if (init)
    GDP_HIP(hipMemcpyAsync(P->d_init(), P->h_init, sizeof(double*) * NP, hipMemcpyHostToDevice,
                           s));
  GDP_HIP(hipEventRecord(P->io_ev, s));
  P->io_valid = true;
  return GDP_OK;
}

int read_state(gdp_plan* P, hipStream_t s) {
  GDP_HIP(hipMemcpyAsync(P->h_state, P->d_state(), sizeof(ProblemState) * P->count(),
                         hipMemcpyDeviceToHost, s));
  GDP_HIP(hipStreamSynchronize(s));
  if (P->prof) prof_collect(P);
  return GDP_OK;
}

bool all_checked(const gdp_plan* P) {
  for (int p = 0; p < P->count(); ++p)
    if (!P->h_state[p].checked) return false;
  return true;
}

// Problems the last check reopened: a reopened problem is open with restart set, which
// every update clears, and a check follows a pair of updates.
int reopened(const gdp_plan* P) {
  int k = 0;
  for (int p = 0; p < P->count(); ++p)
    k += !P->h_state[p].done && P->h_state[p].restart;
  return k;
}

bool all_done(const gdp_plan* P) {
  for (int p = 0; p < P->count(); ++p)
    if (!P->h_state[p].done) return false;
  return true;
}

void fill_reports(const gdp_plan* P, gdp_report* rep) {
  if (!rep) return;
  for (int p = 0; p < P->count(); ++p) {
    const ProblemState& S = P->h_state[p];
    rep[p].iterations = S.iter;
    rep[p].converged = S.converged;
    rep[p].relative_residual = S.bnorm2 > 0.0 ? std::sqrt(S.true_rr / S.bnorm2) : 0.0;
    rep[p].recursive_residual = S.bnorm2 > 0.0 ? std::sqrt(S.rr / S.bnorm2) : 0.0;
  }
}

// Replays exec on s, capturing enqueue(stream) into it on first use.  Without graphs
// (or while profiling, which brackets launches with events) it enqueues directly.
template <typename F>
int launch_captured(gdp_plan* P, hipStream_t s, hipGraphExec_t& exec, F enqueue) {
  if (!P->opt.use_graph || P->prof) return enqueue(s);
  if (!exec) {
    hipGraph_t graph = nullptr;
    GDP_HIP(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(P->cap);
    const hipError_t ee = hipStreamEndCapture(P->cap, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    GDP_HIP(ee);
    const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GDP_HIP(ie);
  }
  GDP_HIP(hipGraphLaunch(exec, s));
  P->graph_stream = s;
  return GDP_OK;
}

int launch_pair(gdp_plan* P, hipStream_t s, bool jacobi) {
  return launch_captured(P, s, P->pair_exec[jacobi ? 1 : 0],
                         [&](hipStream_t cs) { return enqueue_pcg_pair(P, cs, jacobi); });
}

// PCG set-up (b statistics, x = 0, r = -b, norms) and the first preconditioned residual.
int enqueue_start(gdp_plan* P, hipStream_t s, bool jacobi) {
  const gdp_options& o = P->opt;
  int rc = GDP_T(enqueue_init_t)(P, s, 0, 1, o.tolerance, o.max_iterations, false, 1);
  if (rc) return rc;
  return jacobi ? GDP_T(enqueue_jacobi_t)(P, s, false) : enqueue_vcycle(P, s, true, false);
}

// The device-side end of a batch of pairs: FP64 true residual of the problems that
// stopped, reopen (residual replacement) or accept, and the output of the accepted ones.
int enqueue_check(gdp_plan* P, hipStream_t s, bool check_conv = false) {
  int rc;
  if ((rc = GDP_T(enqueue_residual_t)(P, s, 0, 2, false, true, check_conv))) return rc;
  return enqueue_finalize(P, s, 1, 1);
}

// The first batch of a solve as one graph: set-up, first preconditioning, `pairs`
// iteration pairs and the check (one graph launch instead of pairs + 1, each of which
// leaves the device idle for several microseconds).  Its last iteration ends with the
// convergence test instead of a preconditioner: the batch is sized from the last solve,
// so that test usually ends the solve, and a V-cycle after it would only be skipped
// launch by launch.  Kept per pair count (at most 4).
int launch_head(gdp_plan* P, hipStream_t s, bool jacobi, int pairs) {
  auto& v = P->head_exec[jacobi ? 1 : 0];
  size_t at = v.size();
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i].first == pairs) at = i;
  if (at == v.size()) {
    if (v.size() >= 4) {
      if (P->graph_stream) (void)hipStreamSynchronize(P->graph_stream);
      if (v.front().second) (void)hipGraphExecDestroy(v.front().second);
      v.erase(v.begin());
    }
    v.emplace_back(pairs, nullptr);
    at = v.size() - 1;
  }
  return launch_captured(P, s, v[at].second, [&](hipStream_t cs) {
    int rc;
    if ((rc = enqueue_start(P, cs, jacobi))) return rc;
    int pending = 0;
    for (int k = 0; k < pairs; ++k)
      if ((rc = enqueue_pcg_pair(P, cs, jacobi, k == pairs - 1, k < pairs - 1, &pending)))
        return rc;
    return enqueue_check(P, cs, true);
  });
}

int run_pcg(gdp_plan* P, const double* const* b, double* const* u, gdp_report* rep,
            hipStream_t s, bool jacobi) {
  if (!P || !P->bound) {
    set_error("plan is not bound to a workspace");
    return GDP_EINVAL;
  }
  int rc;
  if ((rc = stage_io(P, b, u, nullptr, s))) return rc;
  // First batch, one graph: set-up, z0 = M r0 with its <r0, z0>, as many iteration pairs
  // as the last solve of this plan needed, and the device-side check of the problems that
  // finished (FP64 true residual, reopen for residual replacement or accept, output of
  // the accepted ones).  The host then reads the states once; usually that ends the
  // solve.  Otherwise it continues one pair at a time with the same check.
  const int first_pairs = std::max(1, (P->last_iters + 1) / 2);
  if ((rc = launch_head(P, s, jacobi, first_pairs))) return rc;
  if ((rc = read_state(P, s))) return rc;
  bool precondition = true;  // the first batch left its last preconditioner out
  while (!all_checked(P)) {
    if (reopened(P) > 0 || precondition) {
      // z = M r of the open problems: those the first batch left before their
      // preconditioner, and reopened ones, which restart from their true residual (no
      // convergence test: the recursive residual they stopped on is stale)
      if (jacobi) rc = GDP_T(enqueue_jacobi_t)(P, s, false);
      else rc = enqueue_vcycle(P, s, true, false);
      if (rc) return rc;
    }
    precondition = false;
    if ((rc = launch_pair(P, s, jacobi))) return rc;
    if ((rc = enqueue_check(P, s))) return rc;
    if ((rc = read_state(P, s))) return rc;
  }
  int mx = 0;
  for (int p = 0; p < P->count(); ++p) mx = std::max(mx, P->h_state[p].iter);
  P->last_iters = mx;
  fill_reports(P, rep);
  return GDP_OK;
}

}  // namespace

extern "C" {

int gdp_mgcg_solve(gdp_plan* P, const double* const* b, double* const* u, gdp_report* rep,
                   void* stream) {
  return run_pcg(P, b, u, rep, static_cast<hipStream_t>(stream), false);
}

int gdp_plan_profile(gdp_plan* P, int enable) {
  if (!P) return (set_error("null plan"), GDP_EINVAL);
  P->prof = enable != 0;
  P->prof_agg.clear();
  return GDP_OK;
}

int gdp_plan_profile_count(gdp_plan* P, int* count) {
  if (!P || !count) return (set_error("null argument"), GDP_EINVAL);
  GDP_HIP(hipDeviceSynchronize());
  prof_collect(P);
  *count = int(P->prof_agg.size());
  return GDP_OK;
}

int gdp_plan_profile_entry(gdp_plan* P, int index, char* name, int name_len, double* ms,
                           long long* launches, double* bytes) {
  if (!P || index < 0 || index >= int(P->prof_agg.size()) || !name || name_len < 1 || !ms ||
      !launches || !bytes)
    return (set_error("bad profile entry request"), GDP_EINVAL);
  const auto& g = P->prof_agg[index];
  snprintf(name, size_t(name_len), "%s", g.name.c_str());
  *ms = g.ms;
  *launches = g.launches;
  *bytes = g.bytes;
  return GDP_OK;
}

int gdp_cg_solve(gdp_plan* P, const double* const* b, double* const* u, gdp_report* rep,
                 void* stream) {
  return run_pcg(P, b, u, rep, static_cast<hipStream_t>(stream), true);
}

int gdp_vcycle(gdp_plan* P, const double* const* u_in, const double* const* b,
               double* const* u_out, void* stream) {
  if (!P || !P->bound || !u_in || !b || !u_out) {
    set_error("gdp_vcycle: unbound plan or null argument");
    return GDP_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc;
  if ((rc = stage_io(P, b, u_out, u_in, s))) return rc;
  if ((rc = GDP_T(enqueue_init_t)(P, s, 1, 0, 0.0, 1, true, 0))) return rc;
  if ((rc = GDP_T(enqueue_residual_t)(P, s, 1, 0, false, false, false))) return rc;
  if ((rc = enqueue_vcycle(P, s, false, false))) return rc;
  if ((rc = GDP_T(enqueue_correct_t)(P, s))) return rc;
  return enqueue_finalize(P, s, 0);
}

int gdp_mg_solve(gdp_plan* P, const double* const* b, const double* const* initial,
                 double* const* u, gdp_report* rep, double* history, void* stream) {
  if (!P || !P->bound || !b || !u) {
    set_error("gdp_mg_solve: unbound plan or null argument");
    return GDP_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const gdp_options& o = P->opt;
  int rc;
  if ((rc = stage_io(P, b, u, initial, s))) return rc;
  if ((rc = GDP_T(enqueue_init_t)(P, s, 1, 1, 0.0, o.max_iterations, initial != nullptr, 1)))
    return rc;
  if ((rc = GDP_T(enqueue_residual_t)(P, s, 1, 1, true, false, false))) return rc;
  int cycles = 0;
  int batch = std::max(1, P->last_iters);
  for (;;) {
    for (int k = 0; k < batch; ++k) {
      if ((rc = enqueue_vcycle(P, s, false, false))) return rc;
      if ((rc = GDP_T(enqueue_correct_t)(P, s))) return rc;
      if ((rc = GDP_T(enqueue_residual_t)(P, s, 2, 1, true, false, false))) return rc;
    }
    cycles += batch;
    batch = 1;
    if ((rc = read_state(P, s))) return rc;
    if (all_done(P) || cycles > o.max_iterations) break;
  }
  if ((rc = enqueue_finalize(P, s, 1))) return rc;
  if ((rc = read_state(P, s))) return rc;
  int mx = 0;
  for (int p = 0; p < P->count(); ++p) mx = std::max(mx, P->h_state[p].iter);
  P->last_iters = mx;
  if (rep) {
    fill_reports(P, rep);
    for (int p = 0; p < P->count(); ++p)
      rep[p].relative_residual = rep[p].recursive_residual;  // MG's residual is the true one
  }
  if (history) {
    const size_t stride = size_t(o.max_iterations) + 2;
    std::vector<double> h(stride * P->count());
    GDP_HIP(hipMemcpyAsync(h.data(), P->d_hist(), h.size() * 8, hipMemcpyDeviceToHost, s));
    GDP_HIP(hipStreamSynchronize(s));
    for (int p = 0; p < P->count(); ++p)
      for (int k = 0; k <= o.max_iterations; ++k)
        history[size_t(p) * (o.max_iterations + 1) + k] = h[p * stride + k];
  }
  return GDP_OK;
}

// ----------------------------------------------------------------- single grids
static int check_grid(const void* a, const void* b, int m, int n, int planes) {
  if (!a || !b || m < 1 || n < 1 || planes < 1) {
    set_error("null pointer or empty grid");
    return GDP_EINVAL;
  }
  return GDP_OK;
}

int gdp_laplacian(const double* u, double* out, int m, int n, int planes, double spacing,
                  int boundary, void* stream) {
  if (int rc = check_grid(u, out, m, n, planes)) return rc;
  GDP_HIP(launch_laplacian(u, out, m, n, planes, spacing, boundary,
                           static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_gradient(const double* img, double* gx, double* gy, int h, int w, int planes,
                 void* stream) {
  if (int rc = check_grid(img, gx, h, w, planes)) return rc;
  if (!gy) return (set_error("null gy"), GDP_EINVAL);
  GDP_HIP(launch_gradient(img, gx, gy, h, w, planes, static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_divergence(const double* gx, const double* gy, double* out, int h, int w, int planes,
                   void* stream) {
  if (int rc = check_grid(gx, gy, h, w, planes)) return rc;
  if (!out) return (set_error("null out"), GDP_EINVAL);
  GDP_HIP(launch_divergence(gx, gy, out, h, w, planes, static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_fold_boundary(const double* rhs, const double* boundary, double* out, int h, int w,
                      int planes, double spacing, void* stream) {
  if (int rc = check_grid(rhs, boundary, h - 2, w - 2, planes)) return rc;
  if (!out) return (set_error("null out"), GDP_EINVAL);
  GDP_HIP(launch_fold(rhs, boundary, out, h, w, planes, spacing,
                      static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

static int clone_rhs_checked(const double* source, int sh, int sw, int row, int col,
                      const uint8_t* mask, const double* target, int h, int w, int planes,
                      int top, int left, int hh, int ww, int mode, const uint8_t* interior,
                      double* out, void* stream) {
  if (!source || !mask || !target || !out) return (set_error("null argument"), GDP_EINVAL);
  if (planes < 1 || sh < 1 || sw < 1 || h < 3 || w < 3)
    return (set_error("bad sizes"), GDP_EINVAL);
  if (row < 0 || col < 0 || row + sh > h || col + sw > w)
    return (set_error("the source does not fit in the target at this offset"), GDP_EINVAL);
  if (top < 0 || left < 0 || hh < 3 || ww < 3 || top + hh > h || left + ww > w)
    return (set_error("the window must be at least 3 x 3 and inside the target"), GDP_EINVAL);
  if (mode < 0 || mode > 2) return (set_error("unknown clone mode %d", mode), GDP_EINVAL);
  GDP_HIP(launch_clone_rhs(source, sh, sw, row, col, mask, target, h, w, planes, top, left, hh,
                           ww, mode, interior, out, static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_clone_rhs(const double* source, int sh, int sw, int row, int col,
                  const uint8_t* mask, const double* target, int h, int w, int planes, int top,
                  int left, int hh, int ww, int mode, double* out, void* stream) {
  return clone_rhs_checked(source, sh, sw, row, col, mask, target, h, w, planes, top, left, hh,
                           ww, mode, nullptr, out, stream);
}

int gdp_clone_rhs_masked(const double* source, int sh, int sw, int row, int col,
                         const uint8_t* mask, const double* target, int h, int w, int planes,
                         int top, int left, int hh, int ww, int mode, const uint8_t* interior,
                         double* out, void* stream) {
  if (!interior) return (set_error("null interior mask"), GDP_EINVAL);
  return clone_rhs_checked(source, sh, sw, row, col, mask, target, h, w, planes, top, left, hh,
                           ww, mode, interior, out, stream);
}

int gdp_masked_rhs(const double* gx, const double* gy, const double* image,
                   const uint8_t* interior, int h, int w, int planes, int top, int left, int hh,
                   int ww, double* out, void* stream) {
  if (!gx || !gy || !image || !interior || !out) return (set_error("null argument"), GDP_EINVAL);
  if (planes < 1 || h < 3 || w < 3 || top < 0 || left < 0 || hh < 3 || ww < 3 ||
      top + hh > h || left + ww > w)
    return (set_error("the window must be at least 3 x 3 and inside the image"), GDP_EINVAL);
  GDP_HIP(launch_masked_rhs(gx, gy, image, interior, h, w, planes, top, left, hh, ww, out,
                            static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_guidance(const double* ax, const double* ay, const double* bx, const double* by,
                 const uint8_t* mask, const double* factor, double scale, int mode, int h, int w,
                 int planes, double* out_x, double* out_y, void* stream) {
  if (!ax || !ay || !out_x || !out_y || h < 1 || w < 1 || planes < 1 || mode < 0 || mode > 3)
    return (set_error("bad guidance arguments"), GDP_EINVAL);
  if (mode <= 2 && (!bx || !by)) return (set_error("null second field"), GDP_EINVAL);
  if (mode == 0 && !mask) return (set_error("null mask"), GDP_EINVAL);
  GDP_HIP(launch_guidance(ax, ay, bx, by, mask, factor, scale, mode, size_t(h) * size_t(w),
                          planes, out_x, out_y, static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_illumination_field(const double* gx, const double* gy, size_t n, double alpha,
                           double beta, double epsilon, double* out_x, double* out_y,
                           void* stream) {
  if (!gx || !gy || !out_x || !out_y) return (set_error("bad illumination arguments"), GDP_EINVAL);
  // alpha ** beta once on the host, as the reference forms it (a scalar power)
  GDP_HIP(launch_illumination(gx, gy, n, std::pow(alpha, beta), beta, epsilon, out_x, out_y,
                              static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_smooth(double* u, const double* b, int m, int n, double spacing, int sweeps,
               int boundary, int order, void* stream) {
  if (int rc = check_grid(u, b, m, n, 1)) return rc;
  if (sweeps < 0) return (set_error("negative sweeps"), GDP_EINVAL);
  if (sweeps == 0) return GDP_OK;
  GDP_HIP(launch_smooth(u, b, m, n, spacing, sweeps, boundary, order,
                        static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_restrict(const double* fine, double* coarse, int m, int n, int mc, int nc, int boundary,
                 void* stream) {
  if (int rc = check_grid(fine, coarse, m, n, 1)) return rc;
  if (mc < 1 || nc < 1) return (set_error("empty coarse grid"), GDP_EINVAL);
  GDP_HIP(launch_restrict(fine, coarse, m, n, mc, nc, boundary,
                          static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_prolong(const double* coarse, double* fine, int m, int n, int mc, int nc, int boundary,
                void* stream) {
  if (int rc = check_grid(fine, coarse, m, n, 1)) return rc;
  if (mc < 1 || nc < 1) return (set_error("empty coarse grid"), GDP_EINVAL);
  GDP_HIP(launch_prolong(coarse, fine, m, n, mc, nc, boundary,
                         static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_relative_residual(const double* u, const double* b, int m, int n, int planes,
                          double spacing, int boundary, double* result, void* stream) {
  if (int rc = check_grid(u, b, m, n, planes)) return rc;
  if (!result) return (set_error("null result"), GDP_EINVAL);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int blocks = 64;
  double2* part = nullptr;
  GDP_HIP(hipMallocAsync((void**)&part, sizeof(double2) * blocks * planes, s));
  GDP_HIP(launch_resnorm_partial(u, b, m, n, planes, spacing, boundary, part, blocks, s));
  std::vector<double2> h(size_t(blocks) * planes);
  GDP_HIP(hipMemcpyAsync(h.data(), part, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, s));
  GDP_HIP(hipFreeAsync(part, s));
  GDP_HIP(hipStreamSynchronize(s));
  for (int p = 0; p < planes; ++p) {
    double rr = 0.0, bb = 0.0;
    for (int k = 0; k < blocks; ++k) {
      rr += h[size_t(p) * blocks + k].x;
      bb += h[size_t(p) * blocks + k].y;
    }
    result[p] = bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : 0.0;
  }
  return GDP_OK;
}

}  // extern "C"
