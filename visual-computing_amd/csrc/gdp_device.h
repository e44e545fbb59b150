// Device-side data structures shared by the kernels and the host planner.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Global address space: pointer accesses through it compile to global (not flat) memory
// instructions, whose completions are counted apart from LDS and scalar loads.
#define GDP_GLOBAL __attribute__((address_space(1)))

namespace gdp {

constexpr int kDirichlet = 0;
constexpr int kNeumann = 1;

// One level of one problem.  Arrays are row-major with `pitch` elements per row and
// hold the V-cycle working precision T (float or double).
struct LevelDesc {
  void* u;        // final iterate after post-smoothing (level 0: z = M r)
  void* v;        // iterate after pre-smoothing
  void* b;        // right-hand side
  int32_t m, n, pitch, bc;
  double h2;      // spacing^2 at this level
  double inv_h2;  // 1 / spacing^2
  int32_t coarse; // index of the next coarser LevelDesc, -1 at the coarsest
  int32_t problem;
  int32_t level;
  int32_t pad;
};

// FP64 Krylov vectors of one problem on its fine grid (pitched like level 0).
struct KrylovDesc {
  double* x;
  double* r;
  double* d[2];   // ping-pong search directions
  double* q;
  double* rhs;    // FP64 copy of the (mean-projected) right-hand side
  int32_t m, n, pitch, bc;
  double h2, inv_h2;
  int32_t level0; // LevelDesc index of level 0
  int32_t pad;
};

// Caller-owned dense (pitch = n) input / output of one solve.
struct IoDesc {
  const double* b;
  double* u;
};

// Per-problem scalar state of the PCG / MG iteration (FP64).
struct ProblemState {
  double delta;      // <r, z> of the current direction
  double delta_old;  // <r, z> of the previous direction (for beta)
  double dq;         // <d, q>
  double alpha;
  double beta;
  double rr;         // ||r||^2 (recursive)
  double target2;    // (tol * ||rhs||)^2
  double bnorm2;     // ||rhs||^2
  double bmean;      // mean of b (Neumann projection)
  double true_rr;    // ||rhs - (-A) x||^2 recomputed
  double sumx;       // sum of x (Neumann zero-mean output)
  double xalpha;     // alpha of the direction whose x update an apply held back (xhold)
  int32_t iter;
  int32_t converged; // recursive residual met the target
  int32_t done;      // converged or out of iterations: every kernel skips this problem
  int32_t restart;   // next preconditioned direction starts a new Krylov sequence
  int32_t max_iter;
  int32_t flexible;
  int32_t checked;   // done, true residual checked, output written (PCG)
  int32_t xpend;     // 1 + the direction buffer whose alpha d the last k_update left out
                     // of x for the tail launch of the same iteration or the next apply
                     // to add; 0: none
  int32_t xhold;     // the last apply left a pending alpha d in x's debt (xalpha, and the
                     // direction buffer the update after it does not write)
  int32_t xdouble;   // the pending x update covers two directions: xalpha times the
                     // buffer other than xpend's, then alpha times xpend's
  int32_t restarts;  // residual replacements so far (at most kMaxRestarts)
};
constexpr int kMaxRestarts = 4;

// One workgroup's work item of a V-cycle kernel: everything the block needs about its
// level and the next coarser one, so a block starts with a single descriptor load.
struct VTile {
  void* u;        // level arrays (T)
  void* v;
  void* b;
  void* cu;       // next coarser level: iterate and right-hand side (T)
  void* cb;
  double* kr;     // finest level only: the FP64 residual and q = -A d of the problem
  double* kq;
  double h2, inv_h2;
  int32_t m, n, pitch, bc;
  int32_t cm, cn, cpitch, level;
  int32_t problem, i0, j0, slot;  // slot: this tile's partial-sum slot (finest up kernel)
  int32_t lead;                   // the one tile per problem that writes its state
  int32_t rbeg, rcnt;             // ||r||^2 partials to test convergence from (level 0)
  int32_t pad;
  const uint8_t* mask;            // finest level of a masked domain: m x n bytes, nonzero =
                                  // unknown (the rest held at zero); else null
  int32_t abeg, acnt;             // level 0: slots of the problem's <d, q> partials (the
                                  // down kernel that also carries the PCG update)
};
static_assert(sizeof(VTile) == 152, "VTile layout");

// One workgroup's work item of a Krylov (finest-grid FP64) kernel.
struct KTile {
  double* x;
  double* r;
  double* d0;
  double* d1;
  double* q;
  void* z;        // finest V-cycle level: iterate (the preconditioned residual), T
  void* b0;       // finest V-cycle level: right-hand side, T
  double inv_h2;
  int32_t m, n, pitch, bc;
  int32_t problem, i0, j0, slot;
  int32_t lead;
  int32_t kbeg, kcnt;  // slots of this problem's Krylov-tile partials
  int32_t zbeg, zcnt;  // slots of this problem's <r, z> partials (finest up kernel / tail)
  int32_t abeg, acnt;  // slots of this problem's <d, q> partials (the strip apply's tiles)
  int32_t pad;
  // masked domain (Dirichlet): m x n bytes, nonzero = unknown, the rest held at zero; null
  // = every point is an unknown
  const uint8_t* mask;
};
static_assert(sizeof(KTile) == 136, "KTile layout");

// Slot range of one problem in a partial-sum buffer.
struct SlotRange {
  int32_t begin, count;
};

// Coarse-tail layout of one problem: levels [first, last] live in LDS.
struct TailDesc {
  int32_t first;          // LevelDesc index of the first tail level
  int32_t nlev;           // number of tail levels (the last one is the coarsest)
  int32_t problem;
  int32_t krylov;         // KrylovDesc index if first is level 0, else -1
  int32_t lds_off[16][2]; // byte offsets of (u, b) of each tail level in LDS
  int32_t scratch_off;    // residual scratch
  int32_t coarse_off;     // FP64 coarsest working area (two arrays of the coarsest level)
  int32_t tab_off;        // FP64 DCT tables in LDS (Neumann)
  int32_t pad;
  const double* dct;      // exact coarsest solve: Cm (m*m), Cn (n*n), 1/lambda (m*n)
};

constexpr int kTailThreads = 1024;
constexpr int kTailLdsBytes = 155648;  // of the 160 KB of LDS a gfx950 workgroup may use
constexpr int kMaxTailLevels = 16;

}  // namespace gdp
