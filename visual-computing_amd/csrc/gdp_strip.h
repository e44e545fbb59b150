// Register-resident wave strips: the smoothing kernels of the tiled multigrid levels.
//
// A wave owns a strip 128 fine columns wide and FH rows tall.  Lane l keeps the two
// adjacent columns 2l (x) and 2l+1 (y) of every row of the strip in registers, so in a
// red-black half-sweep a point finds its vertical neighbours and one horizontal
// neighbour in its own lane and the other horizontal neighbour one lane away, fetched
// with a DPP wave shift.  A half-sweep of a row costs one update per lane (x or y,
// fixed by the row's colour) of five VALU operations; there is no LDS traffic and no
// barrier.  Each block stacks four independent strips (its tile record covers four
// strips' worth of rows) so the block-level reductions of the PCG scalars keep their
// one-slot-per-block layout.
//
// Halo: a strip is loaded with a margin around the points it owns.  After k
// half-sweeps only the points at least k rows/columns inside the strip agree with a
// whole-grid sweep (the strip's outer ring is never updated, or is updated from the
// zeros a wave shift brings in), so the margins below are sized for the number of
// sweeps; everything the kernels write lies inside the exact region.
//
// Boundaries: points outside the grid stay zero (inverse diagonal 0, right-hand side
// 0), so the sum of the four neighbours is the sum over in-grid neighbours for both
// boundary conditions; only the inverse diagonal differs (1/4 Dirichlet, 1/#in-grid
// neighbours Neumann).  It is kept per lane for the three possible vertical counts of
// a row, and the row picks one.  A strip whose rows are all strictly inside the grid
// runs a variant with no per-row tests.
//
// Arithmetic follows gradient-domain/src/gradient_domain/multigrid.py: smoother
// :147-166 (colour 0 then 1, or the reverse for the symmetric Neumann post-smoother),
// residual :210, restriction 0.25 P^T r P :169-174, prolongation P e P^T :177-181.
#pragma once
#include <type_traits>

#include "gdp_reduce.h"

namespace gdp {

template <typename T>
struct Pair;
template <>
struct Pair<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct Pair<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// Wave shifts: lane l receives lane l-1's value (shr) or lane l+1's (shl); the lane
// with no source receives 0.  Every lane of the wave must be active.
__device__ inline int dpp_shr1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ inline int dpp_shl1_i(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ inline float dpp_shr1(float v) { return __int_as_float(dpp_shr1_i(__float_as_int(v))); }
__device__ inline float dpp_shl1(float v) { return __int_as_float(dpp_shl1_i(__float_as_int(v))); }
__device__ inline double dpp_shr1(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_shr1_i(int(unsigned(b)));
  const unsigned hi = (unsigned)dpp_shr1_i(int(unsigned(b >> 32)));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double dpp_shl1(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_shl1_i(int(unsigned(b)));
  const unsigned hi = (unsigned)dpp_shl1_i(int(unsigned(b >> 32)));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// a * b + c rounded once.  The relaxation, residual and restriction sums of the level
// kernels are written with it, so their bits do not depend on which products the compiler
// chooses to contract (the sweep-order and streamed down kernels give the same bits).
__device__ inline float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// 1 / k for the neighbour counts k = 1..4 (0 for k <= 0).
template <typename T>
__device__ inline T inv_count(int k) {
  return k == 4 ? T(0.25) : (k == 3 ? T(1.0 / 3.0) : (k == 2 ? T(0.5) : (k == 1 ? T(1) : T(0))));
}

// Geometry of the two strip kernels for NU pre/post sweeps.
template <int NU>
struct StripGeo {
  // down: coarse tile columns per strip and the fine offset of the tile's first point
  static constexpr int kOff = 2 * NU + 2;
  static constexpr int kDownCols = (125 - 4 * NU) / 2;
  // up: margins (rows / columns) and owned fine columns per strip.  The columns get
  // one more pair: the interpolation of a strip's first and last column needs coarse
  // values from outside the wave, so those two columns start out inexact.
  static constexpr int kHalo = 2 * NU;
  static constexpr int kColHalo = 2 * NU + 2;
  static constexpr int kUpCols = 128 - 2 * kColHalo;
};

// Per-lane column data of a strip whose first fine column is gc0 (even).
template <typename T>
struct LaneCols {
  int gjx;       // global column of x (y is gjx + 1)
  int inx, iny;  // columns inside the grid (0 / 1)
  int cx, cy;    // in-grid horizontal neighbours (Neumann) or 2 (Dirichlet)
  T hx, hy;      // the same as T
  T ix2, iy2;    // inverse diagonal of a row with two in-grid vertical neighbours

  __device__ __attribute__((always_inline)) void init(int gc0, int lane, int n, bool neu) {
    gjx = gc0 + 2 * lane;
    inx = gjx >= 0 && gjx < n;
    iny = gjx + 1 >= 0 && gjx + 1 < n;
    cx = neu ? (gjx > 0) + (gjx < n - 1) : 2;
    cy = neu ? (gjx + 1 > 0) + (gjx + 1 < n - 1) : 2;
    hx = T(cx);
    hy = T(cy);
    ix2 = inx ? inv_count<T>(cx + 2) : T(0);
    iy2 = iny ? inv_count<T>(cy + 2) : T(0);
  }
  // Inverse diagonal for a row with vc in-grid vertical neighbours.  Rows with fewer
  // than two (the grid's first and last) compute it; a select between two stored
  // fields would be merged into one load from a computed address and force the
  // record into memory.
  __device__ __attribute__((always_inline)) T wx(int vc) const {
    return vc == 2 ? ix2 : (inx ? inv_count<T>(cx + vc) : T(0));
  }
  __device__ __attribute__((always_inline)) T wy(int vc) const {
    return vc == 2 ? iy2 : (iny ? inv_count<T>(cy + vc) : T(0));
  }
};

// Vertical in-grid neighbour count of global row gi (2 for Dirichlet: diag is 4).
__device__ inline int row_count(int gi, int m, bool neu) {
  return neu ? (gi > 0) + (gi < m - 1) : 2;
}

// Unknown bits of a masked finest level: bit 2a is the lane's x column of strip row a,
// bit 2a + 1 its y column.
__device__ inline bool mask_bit(uint64_t mb, int bit) { return (mb >> bit) & 1ull; }

// One red-black half-sweep over rows 1 .. FH-2 of a strip.  Row a updates x when
// (a + S + FLIP) is even, else y.  EDGE: rows may lie outside the grid or on its
// first/last row.  MASK: the inverse diagonals come from wt (mask_weights), zero at fixed
// points and outside the grid, where hb' is zero too, so those points stay at zero; the
// sweep then needs no per-row branch.
template <typename T, int FH, int S, int FLIP, bool EDGE, bool MASK = false>
__device__ __attribute__((always_inline)) void strip_halfsweep(typename Pair<T>::type* u, const typename Pair<T>::type* hb,
                                       const LaneCols<T>& lc, int gr0, int m, bool neu,
                                       const typename Pair<T>::type* wt = nullptr) {
#pragma unroll
  for (int a = 1; a < FH - 1; ++a) {
    if constexpr (MASK) {
      if (((a + S + FLIP) & 1) == 0) {
        const T nb = ((u[a - 1].x + u[a + 1].x) + dpp_shr1(u[a].y)) + u[a].y;
        u[a].x = fmad(wt[a].x, nb, -hb[a].x);
      } else {
        const T nb = ((u[a - 1].y + u[a + 1].y) + u[a].x) + dpp_shl1(u[a].x);
        u[a].y = fmad(wt[a].y, nb, -hb[a].y);
      }
      continue;
    }
    int vc = 2;
    if (EDGE) {
      const int gi = gr0 + a;
      if (gi < 0 || gi >= m) continue;
      vc = row_count(gi, m, neu);
    }
    if (((a + S + FLIP) & 1) == 0) {
      const T nb = ((u[a - 1].x + u[a + 1].x) + dpp_shr1(u[a].y)) + u[a].y;
      const T w = EDGE ? lc.wx(vc) : lc.ix2;
      u[a].x = fmad(w, nb, -hb[a].x);
    } else {
      const T nb = ((u[a - 1].y + u[a + 1].y) + u[a].x) + dpp_shl1(u[a].x);
      const T w = EDGE ? lc.wy(vc) : lc.iy2;
      u[a].y = fmad(w, nb, -hb[a].y);
    }
  }
}

template <typename T, int FH, int S0, int NS, int FLIP, bool EDGE, bool MASK = false>
__device__ __attribute__((always_inline)) void strip_sweeps(typename Pair<T>::type* u, const typename Pair<T>::type* hb,
                                    const LaneCols<T>& lc, int gr0, int m, bool neu,
                                    const typename Pair<T>::type* wt = nullptr) {
  if constexpr (NS > 0) {
    strip_halfsweep<T, FH, S0, FLIP, EDGE, MASK>(u, hb, lc, gr0, m, neu, wt);
    strip_sweeps<T, FH, S0 + 1, NS - 1, FLIP, EDGE, MASK>(u, hb, lc, gr0, m, neu, wt);
  }
}

// Inverse diagonals of the strip rows of a masked level: the grid's own (as LaneCols
// gives them) at unknowns, zero at fixed points and outside the grid (whose mask bits
// load_mask_bits leaves clear).
template <typename T, int FH>
__device__ __attribute__((always_inline)) void mask_weights(typename Pair<T>::type* wt, uint64_t mb,
                                                            const LaneCols<T>& lc, int gr0, int m,
                                                            bool neu) {
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    const int vc = row_count(gr0 + a, m, neu);
    wt[a].x = mask_bit(mb, 2 * a) ? lc.wx(vc) : T(0);
    wt[a].y = mask_bit(mb, 2 * a + 1) ? lc.wy(vc) : T(0);
  }
}

// Unknown bits of the strip rows gr0 .. gr0 + FH - 1 at the lane's column pair (gjx,
// gjx + 1) of a masked level (mask: m x n bytes, m, n >= 1; null: every grid point is an
// unknown).  Every byte load is issued (at a clamped, in-bounds address) before any is
// waited on; bits of rows or columns outside the grid are clear.
template <int FH>
__device__ inline uint64_t load_mask_bits(const uint8_t* __restrict__ mask, int gr0, int gjx,
                                          int m, int n, bool work) {
  static_assert(2 * FH <= 64, "mask bits of a strip must fit 64 bits");
  const GDP_GLOBAL uint8_t* mk = (const GDP_GLOBAL uint8_t*)mask;
  const int jx = min(max(gjx, 0), n - 1), jy = min(max(gjx + 1, 0), n - 1);
  const bool xin = work && gjx >= 0 && gjx < n, yin = work && gjx + 1 >= 0 && gjx + 1 < n;
  uint32_t bx[FH], by[FH];
  if (mk) {
#pragma unroll
    for (int a = 0; a < FH; ++a) {
      const size_t row = (size_t)min(max(gr0 + a, 0), m - 1) * n;
      bx[a] = mk[row + jx];
      by[a] = mk[row + jy];
    }
  } else {
#pragma unroll
    for (int a = 0; a < FH; ++a) bx[a] = by[a] = 1;
  }
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    const bool rin = gr0 + a >= 0 && gr0 + a < m;
    const uint32_t bits = ((rin && xin && bx[a]) ? 1u : 0u) | ((rin && yin && by[a]) ? 2u : 0u);
    if (2 * a < 32) lo |= bits << (2 * a);
    else hi |= bits << (2 * a - 32);
  }
  return ((uint64_t)hi << 32) | lo;
}

// Full-weighting restriction of one coarse point per lane from the scaled residual rows
// (multigrid.py:169-174): Neumann taps 1/4, 3/4, 3/4, 1/4 over fine rows 2I-1 .. 2I+2 and
// columns 2J-1 .. 2J+2 (the lane's pair and its neighbours' by DPP), Dirichlet 1/2, 1, 1/2
// over 2I .. 2I+2 and 2J .. 2J+2; s = 1 / (4 h^2).  The sums run in this order with one
// rounding per product-add.
template <typename T>
__device__ __attribute__((always_inline)) T restrict_neumann(
    const typename Pair<T>::type& r0, const typename Pair<T>::type& r1,
    const typename Pair<T>::type& r2, const typename Pair<T>::type& r3, T s) {
  const T q0 = T(0.25) * s, q1 = T(0.75) * s;
  T vx = q0 * r0.x;
  vx = fmad(q1, r1.x, vx);
  vx = fmad(q1, r2.x, vx);
  vx = fmad(q0, r3.x, vx);
  T vy = q0 * r0.y;
  vy = fmad(q1, r1.y, vy);
  vy = fmad(q1, r2.y, vy);
  vy = fmad(q0, r3.y, vy);
  T c = T(0.25) * dpp_shr1(vy);
  c = fmad(T(0.75), vx, c);
  c = fmad(T(0.75), vy, c);
  return fmad(T(0.25), dpp_shl1(vx), c);
}

template <typename T>
__device__ __attribute__((always_inline)) T restrict_dirichlet(
    const typename Pair<T>::type& r0, const typename Pair<T>::type& r1,
    const typename Pair<T>::type& r2, T s) {
  const T h = T(0.5) * s;
  T vx = h * r0.x;
  vx = fmad(s, r1.x, vx);
  vx = fmad(h, r2.x, vx);
  T vy = h * r0.y;
  vy = fmad(s, r1.y, vy);
  vy = fmad(h, r2.y, vy);
  const T c = fmad(T(0.5), vx, vy);
  return fmad(T(0.5), dpp_shl1(vx), c);
}

// ------------------------------------------------------------------ down
// Pre-smoothing (NU sweeps from a zero guess) + residual + restriction.  A strip
// produces CR coarse rows x kDownCols coarse columns of the next level's right-hand
// side and the pre-smoothed iterate v of the fine points those coarse points own.
//
// Fine rows of the strip: gr0 = 2 I0 - kOff .. gr0 + FH - 1 (likewise columns, 128 of
// them).  The residual is kept scaled by h^2: r' = h^2 b - (sum - diag u) =
// diag (hb' + u) - sum, with hb' = h^2 b / diag, and the restriction applies
// 0.25 / h^2.  Neumann restriction taps fine 2I-1 .. 2I+2 (1/4, 3/4, 3/4, 1/4) of the
// residual extended by r[-1] = r[0] and, for even sizes, r[f] = r[f-1]; Dirichlet taps
// 2I .. 2I+2 (1/2, 1, 1/2).
template <typename T, int NU, int CR, int FLIP, bool EDGE, bool MASK = false>
__device__ __attribute__((always_inline)) void sdown_body(const VTile& t, int I0, int lane, typename Pair<T>::type* hb,
                                                        uint64_t mb = 0) {
  using V2 = typename Pair<T>::type;
  using G = StripGeo<NU>;
  constexpr int OFF = G::kOff, CC = G::kDownCols;
  constexpr int FH = 2 * CR + (NU > 0 ? 4 * NU + 3 : 4);
  const int m = t.m, n = t.n;
  const int gr0 = 2 * I0 - OFF, gc0 = 2 * t.j0 - OFF;
  const bool neu = t.bc == kNeumann;
  LaneCols<T> lc;
  lc.init(gc0, lane, n, neu);
  const T h2 = T(t.h2);
  V2 wt[MASK ? FH : 1];
  if constexpr (MASK) mask_weights<T, FH>(wt, mb, lc, gr0, m, neu);

  // hb' = h^2 b / diag (zero outside the grid) and the first colour-0 half-sweep from
  // the zero guess, u = -hb' on colour 0
  V2 u[FH];
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    if constexpr (MASK) {
      hb[a].x = mask_bit(mb, 2 * a) ? wt[a].x * h2 * hb[a].x : T(0);
      hb[a].y = mask_bit(mb, 2 * a + 1) ? wt[a].y * h2 * hb[a].y : T(0);
    } else {
      const int gi = gr0 + a;
      int vc = 2;
      bool rin = true;
      if (EDGE) {
        rin = gi >= 0 && gi < m;
        vc = row_count(gi, m, neu);
      }
      const T wx = EDGE ? lc.wx(vc) : lc.ix2;
      const T wy = EDGE ? lc.wy(vc) : lc.iy2;
      hb[a].x = (rin && lc.inx) ? wx * h2 * hb[a].x : T(0);
      hb[a].y = (rin && lc.iny) ? wy * h2 * hb[a].y : T(0);
    }
    if (NU > 0 && ((a + FLIP) & 1) == 0) {
      u[a].x = -hb[a].x;
      u[a].y = T(0);
    } else if (NU > 0) {
      u[a].x = T(0);
      u[a].y = -hb[a].y;
    } else {
      u[a] = V2{T(0), T(0)};
    }
  }
  if constexpr (NU > 0)
    strip_sweeps<T, FH, 1, 2 * NU - 1, FLIP, EDGE, MASK>(u, hb, lc, gr0, m, neu, wt);

  // pre-smoothed iterate of the owned fine points (the last strip row / column of the
  // grid owns every remaining fine row / column)
  {
    const int fr0 = 2 * I0, fc0 = 2 * t.j0;
    const int fr1 = (I0 + CR >= t.cm) ? m : min(m, fr0 + 2 * CR);
    const int fc1 = (t.j0 + CC >= t.cn) ? n : min(n, fc0 + 2 * CC);
    GDP_GLOBAL T* v = (GDP_GLOBAL T*)t.v;
    const bool wx = lc.gjx >= fc0 && lc.gjx < fc1;
    const bool wy = lc.gjx + 1 >= fc0 && lc.gjx + 1 < fc1;
#pragma unroll
    for (int a = OFF; a < OFF + 2 * CR + 2 && a < FH; ++a) {
      const int gi = gr0 + a;
      if (gi >= fr1) break;
      GDP_GLOBAL T* row = v + (size_t)gi * t.pitch + lc.gjx;
      if (wx && wy) {
        *(GDP_GLOBAL V2*)row = u[a];
      } else if (wx) {
        row[0] = u[a].x;
      }
    }
  }

  // scaled residual over the rows the restriction reads, in place of hb'
  constexpr int R0 = OFF - 1, R1 = OFF + 2 * CR;  // inclusive
#pragma unroll
  for (int a = R0; a <= R1; ++a) {
    int vc = 2;
    bool rin = true;
    if (EDGE) {
      const int gi = gr0 + a;
      rin = gi >= 0 && gi < m;
      vc = row_count(gi, m, neu);
    }
    const T cx = T(vc) + lc.hx, cy = T(vc) + lc.hy;
    const T nx = ((u[a - 1].x + u[a + 1].x) + dpp_shr1(u[a].y)) + u[a].y;
    const T ny = ((u[a - 1].y + u[a + 1].y) + u[a].x) + dpp_shl1(u[a].x);
    const T rx = fmad(cx, hb[a].x + u[a].x, -nx);
    const T ry = fmad(cy, hb[a].y + u[a].y, -ny);
    hb[a].x = (rin && lc.inx && (!MASK || mask_bit(mb, 2 * a))) ? rx : T(0);
    hb[a].y = (rin && lc.iny && (!MASK || mask_bit(mb, 2 * a + 1))) ? ry : T(0);
  }
  if (neu) {
    // residual extension past the grid: columns first, then rows (corners take both)
    if (gc0 < 0 || gc0 + 128 > n) {
      const bool gl = lc.gjx + 1 == -1;
      const bool gr = lc.gjx == n && (n & 1) == 0;
#pragma unroll
      for (int a = R0; a <= R1; ++a) {
        const T sx = dpp_shl1(hb[a].x);
        const T sy = dpp_shr1(hb[a].y);
        if (gl) hb[a].y = sx;
        if (gr) hb[a].x = sy;
      }
    }
    if (EDGE) {
#pragma unroll
      for (int a = R0; a <= R1; ++a) {
        const int gi = gr0 + a;
        if (gi == -1 && a < R1) hb[a] = hb[a + 1];
        if (gi == m && (m & 1) == 0 && a > R0) hb[a] = hb[a - 1];
      }
    }
  }

  // restriction to the coarse right-hand side
  GDP_GLOBAL T* cb = (GDP_GLOBAL T*)t.cb;
  const T s = T(0.25) * T(t.inv_h2);
  const int jc = lane - OFF / 2;
  const int J = t.j0 + jc;
  const bool wj = jc >= 0 && jc < CC && J < t.cn;
#pragma unroll
  for (int i = 0; i < CR; ++i) {
    const int I = I0 + i;
    T c;
    if (neu) {
      c = restrict_neumann<T>(hb[2 * i + OFF - 1], hb[2 * i + OFF], hb[2 * i + OFF + 1],
                              hb[2 * i + OFF + 2], s);
    } else {
      c = restrict_dirichlet<T>(hb[2 * i + OFF], hb[2 * i + OFF + 1], hb[2 * i + OFF + 2], s);
    }
    if (I < t.cm && wj) cb[(size_t)I * t.cpitch + J] = c;
  }
}

template <typename T, int NU, int CR, bool MASK = false>
__global__ __launch_bounds__(kThreads) void k_sdown(const VTile* __restrict__ tiles,
                                                    ProblemState* __restrict__ st,
                                                    const double2* __restrict__ upart,
                                                    int check_conv) {
  using V2 = typename Pair<T>::type;
  using G = StripGeo<NU>;
  constexpr int OFF = G::kOff;
  constexpr int FH = 2 * CR + (NU > 0 ? 4 * NU + 3 : 4);
  const VTile t = tiles[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int I0 = t.i0 + w * CR;
  const bool work = I0 < t.cm;
  const int gr0 = 2 * I0 - OFF, gc0 = 2 * t.j0 - OFF;
  const int m = t.m, n = t.n;

  // every load of the strip is issued before anything waits on one: branch-free, from
  // clamped addresses inside the padded rows (gjx is even), zeroed outside the grid
  V2 hb[FH];
  {
    const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
    const int gjx = gc0 + 2 * lane;
    const bool lin = gjx >= 0 && gjx < n;
    const int jc = lin ? gjx : 0;
#pragma unroll
    for (int a = 0; a < FH; ++a) {
      const int gi = gr0 + a;
      const bool in = work && lin && gi >= 0 && gi < m;
      const V2 v = *(const GDP_GLOBAL V2*)(b + (size_t)min(max(gi, 0), m - 1) * t.pitch + jc);
      hb[a] = in ? v : V2{T(0), T(0)};
    }
  }
  // a compiler-only fence: it keeps the loads above from being sunk below the state load
  // of the finished-problem test (which would add its round trip before them)
  asm volatile("" ::: "memory");
  if (st[t.problem].done) return;
  if (check_conv == t.level + 1) {
    // convergence test of the preceding PCG update (the loads above are in flight)
    __shared__ double shc[2 * kThreads / 64];
    __shared__ double2 bcc;
    if (check_converged<kThreads>(st, t.problem, upart, SlotRange{t.rbeg, t.rcnt}, t.lead != 0,
                                  shc, &bcc))
      return;
  }
  if (!work) return;
  const bool edge = gr0 < 1 || gr0 + FH > m - 1;
  const int flip = (gr0 + gc0) & 1;
  if constexpr (MASK) {
    // masked levels (kernels of their own, so the unmasked kernels keep their
    // registers): one body for every strip, fixed points held at zero
    const uint64_t mb = load_mask_bits<FH>(t.mask, gr0, gc0 + 2 * lane, m, n, true);
    if (flip) sdown_body<T, NU, CR, 1, true, true>(t, I0, lane, hb, mb);
    else sdown_body<T, NU, CR, 0, true, true>(t, I0, lane, hb, mb);
    return;
  }
  if (edge) {
    if (flip) sdown_body<T, NU, CR, 1, true>(t, I0, lane, hb);
    else sdown_body<T, NU, CR, 0, true>(t, I0, lane, hb);
  } else {
    if (flip) sdown_body<T, NU, CR, 1, false>(t, I0, lane, hb);
    else sdown_body<T, NU, CR, 0, false>(t, I0, lane, hb);
  }
}

// ------------------------------------------------------------------ down, streamed
// The same pre-smoothing + residual + restriction as sdown_body, in wavefront order.  The
// strip's rows are visited top to bottom; at step t row t is loaded (its right-hand side
// prefetched kPrefetch rows ahead) and set to the zero-guess half-sweep, half-sweep s
// (s = 1 .. 2 NU - 1) is applied to row t - s, row t - (2 NU - 1) is then final and
// stored if owned, the residual is formed at the row above it and, once the last residual
// row a coarse row reads is ready, that coarse row is restricted.  A half-sweep of colour
// c at row a reads rows a - 1 and a + 1 only at colour 1 - c, which no half-sweep between
// s - 1 and s changes, so every point sees the values it sees in the sweep order of
// sdown_body: the results are the same bits.  Only a window of about 2 NU + 3 rows is live
// at a time (plus the prefetched loads), so a strip can own CRS = 8 .. 32 coarse rows with
// few registers: the rows recomputed at its ends are 4 NU + 3 per 2 CRS owned.
constexpr int kPrefetch = 8;
// rows in flight of the fused update-down kernel (two FP64 pairs a row)
constexpr int kPrefetchUpd = 4;

// f(integral_constant<int, I>) for I = B .. E - 1, expanded at compile time, so the row
// index of every step is a constant and the row arrays stay in registers however many
// steps a strip has.
template <int B, int E, typename F>
__device__ __attribute__((always_inline)) void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// UPD (finest level of a PCG iteration, the fused update): the right-hand side is the
// updated residual's negative, formed row by row from the FP64 r and q rows (rr, qq,
// prefetched like raw) as rn = r - alpha q, the same fused multiply-add k_update uses;
// the kernel stores the FP32/FP64 right-hand side b0 = -rn of its owned points and adds
// their rn^2 to acc.  r itself is left to the finest up kernel, which owns each point
// alone (a strip's halo rows are other strips' owned rows, so a store here would race
// with their loads).
//
// REV (strips that touch neither grid edge): the rows are visited bottom to top, the same
// schedule mirrored (half-sweep s at row t + s as row t arrives, the residual formed one
// row below the final one, a coarse row restricted once its first residual row is
// ready).  Every point sees the values it sees top to bottom, so the results are the same
// bits.  Alternate strips run mirrored so that two neighbouring strips load the halo rows
// they share at about the same time (both at their start, or both at their end), while
// those rows are still in L2; in one direction for all, one strip loads them first and the
// other last.
template <typename T, int NU, int CRS, int FLIP, bool EDGE, bool UPD = false, bool REV = false>
__device__ __attribute__((always_inline)) void sdown_stream_body(
    const VTile& t, int I0, int lane, typename Pair<T>::type* raw,
    typename Pair<double>::type* rr = nullptr, typename Pair<double>::type* qq = nullptr,
    double alpha = 0.0, double* acc = nullptr) {
  using V2 = typename Pair<T>::type;
  using D2 = typename Pair<double>::type;
  using G = StripGeo<NU>;
  constexpr int PF = UPD ? kPrefetchUpd : kPrefetch;
  constexpr int OFF = G::kOff, CC = G::kDownCols;
  constexpr int FH = 2 * CRS + (NU > 0 ? 4 * NU + 3 : 4);
  constexpr int K = NU > 0 ? 2 * NU - 1 : 0;  // half-sweeps after the zero-guess one
  constexpr int R0 = OFF - 1, R1 = OFF + 2 * CRS;
  static_assert(!(REV && EDGE), "mirrored strips touch no grid edge");
  // the row of step s of the schedule
  auto ro = [](int s) constexpr { return REV ? FH - 1 - s : s; };
  const int m = t.m, n = t.n;
  const int gr0 = 2 * I0 - OFF, gc0 = 2 * t.j0 - OFF;
  const bool neu = t.bc == kNeumann;
  LaneCols<T> lc;
  lc.init(gc0, lane, n, neu);
  const T h2 = T(t.h2);
  const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
  const int gjx = gc0 + 2 * lane;
  const bool lin = gjx >= 0 && gjx < n;
  const int jload = lin ? gjx : 0;
  // stores: owned fine rows / columns of v (the last strip of the grid owns the rest)
  const int fr0 = 2 * I0, fc0 = 2 * t.j0;
  const int fr1 = (I0 + CRS >= t.cm) ? m : min(m, fr0 + 2 * CRS);
  const int fc1 = (t.j0 + CC >= t.cn) ? n : min(n, fc0 + 2 * CC);
  GDP_GLOBAL T* v = (GDP_GLOBAL T*)t.v;
  const bool vx = lc.gjx >= fc0 && lc.gjx < fc1;
  const bool vy = lc.gjx + 1 >= fc0 && lc.gjx + 1 < fc1;
  // restriction
  GDP_GLOBAL T* cb = (GDP_GLOBAL T*)t.cb;
  const T sres = T(0.25) * T(t.inv_h2);
  const int jcr = lane - OFF / 2;
  const int J = t.j0 + jcr;
  const bool wj = jcr >= 0 && jcr < CC && J < t.cn;
  const bool colext = neu && (gc0 < 0 || gc0 + 128 > n);
  const bool gl = lc.gjx + 1 == -1;
  const bool gr = lc.gjx == n && (n & 1) == 0;

  const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)t.kr;
  const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)t.kq;
  GDP_GLOBAL T* b0 = (GDP_GLOBAL T*)t.b;

  V2 u[FH], hb[FH];
  if constexpr (REV) {
    // the first PF rows of the mirrored order (the kernel requested the others' early)
#pragma unroll
    for (int k = 0; k < PF && k < FH; ++k) {
      const int a = ro(k);
      const int gi = gr0 + a;
      const size_t o = (size_t)min(max(gi, 0), m - 1) * t.pitch + jload;
      if constexpr (UPD) {
        rr[a] = *(const GDP_GLOBAL D2*)(kr + o);
        qq[a] = *(const GDP_GLOBAL D2*)(kq + o);
      } else {
        const V2 q = *(const GDP_GLOBAL V2*)(b + o);
        raw[a] = (lin && gi >= 0 && gi < m) ? q : V2{T(0), T(0)};
      }
    }
  }
  static_for<0, FH>([&](auto step) __attribute__((always_inline)) {
    constexpr int st = decltype(step)::value;
    // prefetch the right-hand side PF rows ahead (clamped address, zero outside)
    if (st + PF < FH) {
      const int a = ro(st + PF);
      const int gi = gr0 + a;
      const size_t o = (size_t)min(max(gi, 0), m - 1) * t.pitch + jload;
      if constexpr (UPD) {
        rr[a] = *(const GDP_GLOBAL D2*)(kr + o);
        qq[a] = *(const GDP_GLOBAL D2*)(kq + o);
      } else {
        const V2 q = *(const GDP_GLOBAL V2*)(b + o);
        raw[a] = (lin && gi >= 0 && gi < m) ? q : V2{T(0), T(0)};
      }
    }
    // UPD: the updated residual of row st; its owned points store b0 = -rn and count rn^2
    if constexpr (UPD) {
      const int a = ro(st);
      const int gi = gr0 + a;
      D2 rn;
      rn.x = __builtin_fma(-alpha, qq[a].x, rr[a].x);
      rn.y = __builtin_fma(-alpha, qq[a].y, rr[a].y);
      const V2 nb = V2{T(-rn.x), T(-rn.y)};
      if (a >= OFF && a < OFF + 2 * CRS + 2 && gi < fr1) {
        GDP_GLOBAL T* row = b0 + (size_t)gi * t.pitch + lc.gjx;
        if (vx && vy) {
          *(GDP_GLOBAL V2*)row = nb;
          *acc = __builtin_fma(rn.x, rn.x, *acc);
          *acc = __builtin_fma(rn.y, rn.y, *acc);
        } else if (vx) {
          row[0] = nb.x;
          *acc = __builtin_fma(rn.x, rn.x, *acc);
        }
      }
      raw[a] = (lin && gi >= 0 && gi < m) ? nb : V2{T(0), T(0)};
    }
    // row st: scaled right-hand side and the zero-guess half-sweep (colour 0)
    {
      const int a = ro(st);
      const int gi = gr0 + a;
      int vc = 2;
      bool rin = true;
      if (EDGE) {
        rin = gi >= 0 && gi < m;
        vc = row_count(gi, m, neu);
      }
      const T wx = EDGE ? lc.wx(vc) : lc.ix2;
      const T wy = EDGE ? lc.wy(vc) : lc.iy2;
      hb[a].x = (rin && lc.inx) ? wx * h2 * raw[a].x : T(0);
      hb[a].y = (rin && lc.iny) ? wy * h2 * raw[a].y : T(0);
      if (NU > 0 && ((a + FLIP) & 1) == 0) {
        u[a].x = -hb[a].x;
        u[a].y = T(0);
      } else if (NU > 0) {
        u[a].x = T(0);
        u[a].y = -hb[a].y;
      } else {
        u[a] = V2{T(0), T(0)};
      }
    }
    // half-sweep s at row st - s (mirrored: st + s)
#pragma unroll
    for (int s = 1; s <= K; ++s) {
      const int a = ro(st - s);
      if (a < 1 || a > FH - 2) continue;
      int vc = 2;
      if (EDGE) {
        const int gi = gr0 + a;
        if (gi < 0 || gi >= m) continue;
        vc = row_count(gi, m, neu);
      }
      if (((a + s + FLIP) & 1) == 0) {
        const T nb = ((u[a - 1].x + u[a + 1].x) + dpp_shr1(u[a].y)) + u[a].y;
        const T w = EDGE ? lc.wx(vc) : lc.ix2;
        u[a].x = fmad(w, nb, -hb[a].x);
      } else {
        const T nb = ((u[a - 1].y + u[a + 1].y) + u[a].x) + dpp_shl1(u[a].x);
        const T w = EDGE ? lc.wy(vc) : lc.iy2;
        u[a].y = fmad(w, nb, -hb[a].y);
      }
    }
    // row st - K is final: the pre-smoothed iterate of the owned rows
    {
      const int a = ro(st - K);
      if (a >= OFF && a < OFF + 2 * CRS + 2) {
        const int gi = gr0 + a;
        if (gi < fr1) {
          GDP_GLOBAL T* row = v + (size_t)gi * t.pitch + lc.gjx;
          if (vx && vy) {
            *(GDP_GLOBAL V2*)row = u[a];
          } else if (vx) {
            row[0] = u[a].x;
          }
        }
      }
    }
    // scaled residual of the row above it, in place of hb', then the extensions past the
    // grid (Neumann: columns, then rows) and the restriction of a completed coarse row
    {
      const int a = ro(st - K - 1);
      if (a >= R0 && a <= R1) {
        const int gi = gr0 + a;
        int vc = 2;
        bool rin = true;
        if (EDGE) {
          rin = gi >= 0 && gi < m;
          vc = row_count(gi, m, neu);
        }
        const T cx = T(vc) + lc.hx, cy = T(vc) + lc.hy;
        const T nx = ((u[a - 1].x + u[a + 1].x) + dpp_shr1(u[a].y)) + u[a].y;
        const T ny = ((u[a - 1].y + u[a + 1].y) + u[a].x) + dpp_shl1(u[a].x);
        const T rx = fmad(cx, hb[a].x + u[a].x, -nx);
        const T ry = fmad(cy, hb[a].y + u[a].y, -ny);
        hb[a].x = (rin && lc.inx) ? rx : T(0);
        hb[a].y = (rin && lc.iny) ? ry : T(0);
        if (colext) {
          const T sx = dpp_shl1(hb[a].x);
          const T sy = dpp_shr1(hb[a].y);
          if (gl) hb[a].y = sx;
          if (gr) hb[a].x = sy;
        }
        if (EDGE && neu) {
          if (a > R0 && gi - 1 == -1) hb[a - 1] = hb[a];
          if (gi == m && (m & 1) == 0 && a > R0) hb[a] = hb[a - 1];
        }
        if constexpr (!REV) {
          if (a >= OFF + 2 && ((a - OFF - 2) & 1) == 0) {
            const int i = (a - OFF - 2) / 2;
            const int I = I0 + i;
            T c;
            if (neu) {
              c = restrict_neumann<T>(hb[2 * i + OFF - 1], hb[2 * i + OFF],
                                      hb[2 * i + OFF + 1], hb[2 * i + OFF + 2], sres);
            } else {
              c = restrict_dirichlet<T>(hb[2 * i + OFF], hb[2 * i + OFF + 1],
                                        hb[2 * i + OFF + 2], sres);
            }
            if (I < t.cm && wj) cb[(size_t)I * t.cpitch + J] = c;
          }
        } else {
          // mirrored: coarse row i once its first residual row is ready, 2 i + OFF - 1
          // (Neumann) or 2 i + OFF (Dirichlet)
          const int in = (a - OFF + 1) / 2, id = (a - OFF) / 2;
          if (neu) {
            if (((a - OFF + 1) & 1) == 0 && in < CRS) {
              const T c = restrict_neumann<T>(hb[2 * in + OFF - 1], hb[2 * in + OFF],
                                              hb[2 * in + OFF + 1], hb[2 * in + OFF + 2], sres);
              if (I0 + in < t.cm && wj) cb[(size_t)(I0 + in) * t.cpitch + J] = c;
            }
          } else if (a >= OFF && ((a - OFF) & 1) == 0 && id < CRS) {
            const T c = restrict_dirichlet<T>(hb[2 * id + OFF], hb[2 * id + OFF + 1],
                                              hb[2 * id + OFF + 2], sres);
            if (I0 + id < t.cm && wj) cb[(size_t)(I0 + id) * t.cpitch + J] = c;
          }
        }
      }
    }
  });
}

// UPD: the finest level of a PCG iteration whose update this kernel carries (in place of
// k_update<T, TH, TW, true>): alpha from the <d, q> partials (the same slots and order as
// k_update, so the same bits), the lead block's state record of the update, the right-hand
// side -rn row by row, and the rn^2 partial of each block at its slot in upart; the
// convergence test of that update falls to the tail kernels' entry.
template <typename T, int NU, int CRS, bool UPD = false>
__global__ __launch_bounds__(kThreads) void k_sdown_stream(const VTile* __restrict__ tiles,
                                                           ProblemState* __restrict__ st,
                                                           double2* __restrict__ upart,
                                                           int check_conv,
                                                           const double2* __restrict__ apart,
                                                           int dcur) {
  using V2 = typename Pair<T>::type;
  using D2 = typename Pair<double>::type;
  using G = StripGeo<NU>;
  constexpr int OFF = G::kOff;
  constexpr int FH = 2 * CRS + (NU > 0 ? 4 * NU + 3 : 4);
  const VTile t = tiles[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int I0 = t.i0 + w * CRS;
  const bool work = I0 < t.cm;
  const int gr0 = 2 * I0 - OFF, gc0 = 2 * t.j0 - OFF;
  const int m = t.m, n = t.n;
  const bool edge = gr0 < 1 || gr0 + FH > m - 1;
  // odd strips that touch no grid edge visit their rows bottom to top (wave-uniform)
  const bool rev = !edge && (w & 1);
  // the first rows of the strip's order are requested before the state tests
  // (branch-free, clamped addresses, zero outside the grid); the body requests the rest
  // as it goes
  V2 raw[FH];
  D2 rr[UPD ? FH : 1], qq[UPD ? FH : 1];
  {
    const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
    const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)t.kr;
    const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)t.kq;
    const int gjx = gc0 + 2 * lane;
    const bool lin = gjx >= 0 && gjx < n;
    const int jc = lin ? gjx : 0;
    constexpr int PF = UPD ? kPrefetchUpd : kPrefetch;
    // (a mirrored strip requests its first rows, the last ones, in its body)
    if (!rev) {
#pragma unroll
      for (int a = 0; a < PF && a < FH; ++a) {
        const int gi = gr0 + a;
        const size_t o = (size_t)min(max(gi, 0), m - 1) * t.pitch + jc;
        if constexpr (UPD) {
          rr[a] = *(const GDP_GLOBAL D2*)(kr + o);
          qq[a] = *(const GDP_GLOBAL D2*)(kq + o);
        } else {
          const bool in = work && lin && gi >= 0 && gi < m;
          const V2 q = *(const GDP_GLOBAL V2*)(b + o);
          raw[a] = in ? q : V2{T(0), T(0)};
        }
      }
    }
  }
  asm volatile("" ::: "memory");
  const int flip = (gr0 + gc0) & 1;
  if constexpr (UPD) {
    // alpha as k_update forms it; the finished-problem test after the reduction's barrier
    __shared__ double sh[2 * kThreads / 64];
    __shared__ double2 bc;
    const int p = t.problem;
    const double2 a2 = reduce_slots<kThreads>(apart, SlotRange{t.abeg, t.acnt}, sh, &bc);
    if (st[p].done) {
      if (t.lead && threadIdx.x == 0) st[p].xpend = st[p].xdouble = 0;
      return;
    }
    const double delta = st[p].delta;
    const double alpha = delta / a2.x;
    if (t.lead && threadIdx.x == 0) {
      st[p].alpha = alpha;
      st[p].dq = a2.x;
      st[p].delta_old = delta;
      st[p].iter += 1;
      st[p].restart = 0;
      st[p].xpend = dcur + 1;  // the x half goes to the tail launch or the next apply
      st[p].xdouble = st[p].xhold;  // with the one that apply held back, if it did
    }
    double acc = 0.0;
    if (work) {
      if (edge) {
        if (flip) sdown_stream_body<T, NU, CRS, 1, true, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
        else sdown_stream_body<T, NU, CRS, 0, true, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
      } else if (rev) {
        if (flip) sdown_stream_body<T, NU, CRS, 1, false, true, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
        else sdown_stream_body<T, NU, CRS, 0, false, true, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
      } else {
        if (flip) sdown_stream_body<T, NU, CRS, 1, false, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
        else sdown_stream_body<T, NU, CRS, 0, false, true>(t, I0, lane, raw, rr, qq, alpha, &acc);
      }
    }
    // every wave joins the block sum (the waves past the grid with zero)
    write_partial<kThreads>(acc, 0.0, upart, t.slot, sh);
    return;
  } else {
    if (st[t.problem].done) return;
    if (check_conv == t.level + 1) {
      __shared__ double shc[2 * kThreads / 64];
      __shared__ double2 bcc;
      if (check_converged<kThreads>(st, t.problem, upart, SlotRange{t.rbeg, t.rcnt},
                                    t.lead != 0, shc, &bcc))
        return;
    }
    if (!work) return;
    if (edge) {
      if (flip) sdown_stream_body<T, NU, CRS, 1, true>(t, I0, lane, raw);
      else sdown_stream_body<T, NU, CRS, 0, true>(t, I0, lane, raw);
    } else if (rev) {
      if (flip) sdown_stream_body<T, NU, CRS, 1, false, false, true>(t, I0, lane, raw);
      else sdown_stream_body<T, NU, CRS, 0, false, false, true>(t, I0, lane, raw);
    } else {
      if (flip) sdown_stream_body<T, NU, CRS, 1, false>(t, I0, lane, raw);
      else sdown_stream_body<T, NU, CRS, 0, false>(t, I0, lane, raw);
    }
  }
}

// ------------------------------------------------------------------ up
// Prolongation of the coarse correction + post-smoothing.  A strip owns TR fine rows x
// kUpCols fine columns starting at (i0, j0) and loads a margin of 2 NU rows and
// 2 NU + 2 columns around them.  With DOTS (finest level of a PCG iteration) the block also sums
// <r, z> and <q, z> over its points.
//
// rupd (DOTS only): the iteration's update ran in the fused update-down kernel, which left
// r to this kernel: each owned point's residual becomes r - alpha q (k_update's fused
// multiply-add) before the dot products use it, and is stored.
template <typename T, int NU, int TR, int FLIP, bool EDGE, bool DOTS, bool MASK = false>
__device__ __attribute__((always_inline)) void sup_body(const VTile& t, int i0, int lane, typename Pair<T>::type* u,
                         typename Pair<T>::type* hb, const T* E, double& zr, double& zq,
                         uint64_t mb = 0, int rupd = 0, double alpha = 0.0) {
  using V2 = typename Pair<T>::type;
  using G = StripGeo<NU>;
  constexpr int H = G::kHalo, TC = G::kUpCols;
  constexpr int FH = TR + 2 * H;
  const int m = t.m, n = t.n;
  const int gr0 = i0 - H, gc0 = t.j0 - G::kColHalo;
  const bool neu = t.bc == kNeumann;
  LaneCols<T> lc;
  lc.init(gc0, lane, n, neu);
  const T h2 = T(t.h2);
  V2 wt[MASK ? FH : 1];
  if constexpr (MASK) mask_weights<T, FH>(wt, mb, lc, gr0, m, neu);

  // u = v + P e (zero outside the grid); hb' = h^2 b / diag
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    const int gi = gr0 + a;
    int vc = 2;
    bool rin = true;
    if (EDGE) {
      rin = gi >= 0 && gi < m;
      vc = row_count(gi, m, neu);
    }
    T cx, cy;
    if (neu) {
      const int P = a / 2 + 1;
      const int O = (a & 1) ? P + 1 : P - 1;
      // explicit fused multiply-adds: the same bits in every kernel that interpolates
      const T tv = fmad(T(0.25), E[O], T(0.75) * E[P]);
      cx = fmad(T(0.25), dpp_shr1(tv), T(0.75) * tv);
      cy = fmad(T(0.25), dpp_shl1(tv), T(0.75) * tv);
    } else {
      const T tv = (a & 1) ? E[(a + 1) / 2] : T(0.5) * (E[a / 2] + E[a / 2 + 1]);
      cx = T(0.5) * (dpp_shr1(tv) + tv);
      cy = tv;
    }
    u[a].x = (rin && lc.inx) ? u[a].x + cx : T(0);
    u[a].y = (rin && lc.iny) ? u[a].y + cy : T(0);
    if constexpr (MASK) {
      if (!mask_bit(mb, 2 * a)) u[a].x = T(0);
      if (!mask_bit(mb, 2 * a + 1)) u[a].y = T(0);
      hb[a].x = mask_bit(mb, 2 * a) ? wt[a].x * h2 * hb[a].x : T(0);
      hb[a].y = mask_bit(mb, 2 * a + 1) ? wt[a].y * h2 * hb[a].y : T(0);
    } else {
      const T wx = EDGE ? lc.wx(vc) : lc.ix2;
      const T wy = EDGE ? lc.wy(vc) : lc.iy2;
      hb[a].x = (rin && lc.inx) ? wx * h2 * hb[a].x : T(0);
      hb[a].y = (rin && lc.iny) ? wy * h2 * hb[a].y : T(0);
    }
  }
  strip_sweeps<T, FH, 0, 2 * NU, FLIP, EDGE, MASK>(u, hb, lc, gr0, m, neu, wt);

  GDP_GLOBAL T* uo = (GDP_GLOBAL T*)t.u;
  const int fc1 = min(n, t.j0 + TC);
  const bool wx = lc.gjx >= t.j0 && lc.gjx < fc1;
  const bool wy = lc.gjx + 1 >= t.j0 && lc.gjx + 1 < fc1;
  const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)t.kr;
  const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)t.kq;
  // DOTS: the FP64 residual and q of every owned row (the finest level shares the Krylov
  // grid's pitch), loaded together before any is used, so their latency is paid once:
  // branch-free, from clamped addresses inside the padded rows (gjx is even)
  using D2 = typename Pair<double>::type;
  D2 rq[DOTS ? 2 * TR : 1];
  if constexpr (DOTS) {
    const int jc = (lc.gjx >= 0 && lc.gjx < n) ? lc.gjx : 0;
#pragma unroll
    for (int a = 0; a < TR; ++a) {
      const size_t o = (size_t)min(gr0 + H + a, m - 1) * t.pitch + jc;
      rq[2 * a] = *(const GDP_GLOBAL D2*)(kr + o);
      rq[2 * a + 1] = *(const GDP_GLOBAL D2*)(kq + o);
    }
  }
#pragma unroll
  for (int a = H; a < H + TR; ++a) {
    const int gi = gr0 + a;
    if (gi >= m) break;
    const size_t o = (size_t)gi * t.pitch + lc.gjx;
    if (wx && wy) {
      *(GDP_GLOBAL V2*)(uo + o) = u[a];
    } else if (wx) {
      uo[o] = u[a].x;
    }
    if constexpr (DOTS) {
      D2 r2 = rq[2 * (a - H)];
      const D2 q2 = rq[2 * (a - H) + 1];
      if (rupd) {
        r2.x = __builtin_fma(-alpha, q2.x, r2.x);
        r2.y = __builtin_fma(-alpha, q2.y, r2.y);
        GDP_GLOBAL double* rw = (GDP_GLOBAL double*)t.kr;
        if (wx && wy) {
          *(GDP_GLOBAL D2*)(rw + o) = r2;
        } else if (wx) {
          rw[o] = r2.x;
        }
      }
      const double rx = wx ? r2.x : 0.0, ry = (wx && wy) ? r2.y : 0.0;
      const double qx = wx ? q2.x : 0.0, qy = (wx && wy) ? q2.y : 0.0;
      zr += double(u[a].x) * rx + double(u[a].y) * ry;
      zq += double(u[a].x) * qx + double(u[a].y) * qy;
    }
  }
}

template <typename T, int NU, int TR, bool DOTS, bool MASK = false>
__global__ __launch_bounds__(kThreads) void k_sup(const VTile* __restrict__ tiles,
                                                  const ProblemState* __restrict__ st,
                                                  double2* __restrict__ zpart, int order,
                                                  int rupd) {
  using V2 = typename Pair<T>::type;
  using G = StripGeo<NU>;
  constexpr int H = G::kHalo;
  constexpr int FH = TR + 2 * H;
  constexpr int NE = FH / 2 + 2;
  __shared__ double sh[2 * kThreads / 64];
  const VTile t = tiles[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int i0 = t.i0 + w * TR;
  const bool work = i0 < t.m;
  const int gr0 = i0 - H, gc0 = t.j0 - G::kColHalo;
  const int m = t.m, n = t.n;
  const bool neu = t.bc == kNeumann;

  // loads: v and b of the strip, and the coarse correction rows it interpolates from
  // (edge-replicated for Neumann, zero outside for Dirichlet)
  V2 u[FH], hb[FH];
  T E[NE];
  {
    const GDP_GLOBAL T* vv = (const GDP_GLOBAL T*)t.v;
    const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
    const GDP_GLOBAL T* e = (const GDP_GLOBAL T*)t.cu;
    const int gjx = gc0 + 2 * lane;
    const bool lin = gjx >= 0 && gjx < n;
    const int cr0 = gr0 / 2 - 1;
    int K = gc0 / 2 + lane;
    bool kin = K >= 0 && K < t.cn;
    if (neu) {
      K = min(max(K, 0), t.cn - 1);
      kin = true;
    }
    // branch-free: every coarse value comes from a clamped address and is zeroed where
    // it lies outside the coarse grid (Dirichlet), so no load waits before the next
    const int Kc = min(max(K, 0), t.cn - 1);
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int I = cr0 + k;
      const bool in = work && (neu || (kin && I >= 0 && I < t.cm));
      const T ev = e[(size_t)min(max(I, 0), t.cm - 1) * t.cpitch + Kc];
      E[k] = in ? ev : T(0);
    }
    // branch-free: every pair is loaded from a clamped address (inside the grid's padded
    // rows: the pitch is even) and zeroed outside the grid, so the strip arrays are not
    // copied at each conditional load
    const int jc = lin ? gjx : 0;
#pragma unroll
    for (int a = 0; a < FH; ++a) {
      const int gi = gr0 + a;
      const bool in = work && lin && gi >= 0 && gi < m;
      const size_t o = (size_t)min(max(gi, 0), m - 1) * t.pitch + jc;
      const V2 uv = *(const GDP_GLOBAL V2*)(vv + o);
      const V2 bv = *(const GDP_GLOBAL V2*)(b + o);
      u[a] = in ? uv : V2{T(0), T(0)};
      hb[a] = in ? bv : V2{T(0), T(0)};
    }
  }
  if (st[t.problem].done) return;
  const double alpha = (DOTS && rupd) ? st[t.problem].alpha : 0.0;
  double zr = 0.0, zq = 0.0;
  if (work) {
    const bool edge = gr0 < 1 || gr0 + FH > m - 1;
    const int flip = ((gr0 + gc0) + order) & 1;
    if constexpr (MASK) {
      // masked levels (kernels of their own): one body for every strip, fixed points
      // held at zero
      const uint64_t mb = load_mask_bits<FH>(t.mask, gr0, gc0 + 2 * lane, m, n, true);
      if (flip) sup_body<T, NU, TR, 1, true, DOTS, true>(t, i0, lane, u, hb, E, zr, zq, mb);
      else sup_body<T, NU, TR, 0, true, DOTS, true>(t, i0, lane, u, hb, E, zr, zq, mb);
    } else if (edge) {
      if (flip) sup_body<T, NU, TR, 1, true, DOTS>(t, i0, lane, u, hb, E, zr, zq, 0, rupd, alpha);
      else sup_body<T, NU, TR, 0, true, DOTS>(t, i0, lane, u, hb, E, zr, zq, 0, rupd, alpha);
    } else {
      if (flip) sup_body<T, NU, TR, 1, false, DOTS>(t, i0, lane, u, hb, E, zr, zq, 0, rupd, alpha);
      else sup_body<T, NU, TR, 0, false, DOTS>(t, i0, lane, u, hb, E, zr, zq, 0, rupd, alpha);
    }
  }
  if (DOTS) write_partial<kThreads>(zr, zq, zpart, t.slot, sh);
}

}  // namespace gdp
