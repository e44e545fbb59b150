// Block-tiled level kernels.  Down: the pre-smoothing + residual + restriction of sdown_body
// (gdp_strip.h) for one tall strip per block, its rows shared among the block's four
// waves instead of recomputed by each.
//
// A block owns CB coarse rows x kDownCols coarse columns.  Its fine rows are those of a
// strip kernel with CB coarse rows, gr0 = 2 I0 - kOff .. gr0 + FH - 1 (FH = 2 CB + 4 NU
// + 3), rounded up to 4 RW rows: wave w holds rows w RW .. w RW + RW - 1 of them in
// registers, lane l the column pair (2l, 2l+1) of each, as a strip wave does.  A
// half-sweep reads the rows just above and below a wave's rows from the neighbouring
// waves: each wave publishes its first and last row to a double-buffered LDS slot and
// reads its neighbours' after one barrier.  The margin rows of the block (kOff above,
// 2 NU + 1 below the owned rows, 4 NU + 3 per 2 CB owned) are the only rows computed
// twice, against 4 NU + 3 per 2 CR for a strip of CR rows: at CB = 16 and NU = 2 the
// block loads 1.34 rows per owned row, the 2-row strip 3.75.
//
// Every point is updated by the same expression as in sdown_body (explicit fused
// multiply-adds, the same neighbour order), and the rows and columns the block writes
// lie inside the region a whole-grid sweep agrees with, so the results are the same
// bits as the strip kernels'.  The residual rows the restriction reads go through LDS,
// where the Neumann row extension (r[-1] = r[0], r[m] = r[m-1] for even m) is applied
// as they are read.
//
// Arithmetic: gradient-domain/src/gradient_domain/multigrid.py smoother :147-166,
// residual :210, restriction 0.25 P^T r P :169-174.
#pragma once
#include "gdp_strip.h"

namespace gdp {

constexpr int kBlkWaves = kThreads / 64;  // 4

template <typename T, int NU, int CB>
struct BlkGeo {
  static constexpr int kOff = StripGeo<NU>::kOff;
  static constexpr int kFH = 2 * CB + (NU > 0 ? 4 * NU + 3 : 4);
  // rows per wave, even (a row's colour is then its index's parity within the wave)
  static constexpr int kRW = 2 * ((kFH + 2 * kBlkWaves - 1) / (2 * kBlkWaves));
  static constexpr int kR0 = kOff - 1, kR1 = kOff + 2 * CB;  // residual rows (inclusive)
  static constexpr int kRes = kR1 - kR0 + 1;
  // LDS: the halo slots (2 buffers x 4 waves x 2 rows of 64 pairs) and the residual rows
  static constexpr int kHaloBytes = 2 * kBlkWaves * 2 * 64 * int(sizeof(T)) * 2;
  static constexpr int kLdsBytes = kHaloBytes + kRes * 64 * int(sizeof(T)) * 2;
};

// Exchange of wave edge rows: publish this wave's first and last row, return the row
// above its first (the previous wave's last) and below its last (the next wave's first);
// zero at the block's ends.
template <typename V2>
__device__ __attribute__((always_inline)) void blk_halo(V2* xh, int& hbuf, int w, int lane,
                                                      const V2& top, const V2& bot, V2& above,
                                                      V2& below) {
  V2* h = xh + hbuf * (kBlkWaves * 2 * 64);
  h[(w * 2 + 0) * 64 + lane] = top;
  h[(w * 2 + 1) * 64 + lane] = bot;
  __syncthreads();
  above = w > 0 ? h[((w - 1) * 2 + 1) * 64 + lane] : V2{0, 0};
  below = w < kBlkWaves - 1 ? h[((w + 1) * 2 + 0) * 64 + lane] : V2{0, 0};
  hbuf ^= 1;
}

// Half-sweep S of the wave's rows (block rows a = w RW + r; the colour of x is
// (r + S + FLIP) even, RW being even).  EDGE: rows may lie outside the grid or on its
// first / last row.
template <typename T, int RW, int S, int FLIP, bool EDGE>
__device__ __attribute__((always_inline)) void blk_halfsweep(typename Pair<T>::type* u,
                                                           const typename Pair<T>::type* hb,
                                                           const LaneCols<T>& lc, int gw0, int m,
                                                           bool neu,
                                                           const typename Pair<T>::type& above,
                                                           const typename Pair<T>::type& below) {
  using V2 = typename Pair<T>::type;
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    int vc = 2;
    if (EDGE) {
      const int gi = gw0 + r;
      if (gi < 0 || gi >= m) continue;
      vc = row_count(gi, m, neu);
    }
    const V2 up = r > 0 ? u[r - 1] : above;
    const V2 dn = r < RW - 1 ? u[r + 1] : below;
    if (((r + S + FLIP) & 1) == 0) {
      const T nb = ((up.x + dn.x) + dpp_shr1(u[r].y)) + u[r].y;
      const T w = EDGE ? lc.wx(vc) : lc.ix2;
      u[r].x = fmad(w, nb, -hb[r].x);
    } else {
      const T nb = ((up.y + dn.y) + u[r].x) + dpp_shl1(u[r].x);
      const T w = EDGE ? lc.wy(vc) : lc.iy2;
      u[r].y = fmad(w, nb, -hb[r].y);
    }
  }
}

template <typename T, int RW, int S, int NS, int FLIP, bool EDGE>
__device__ __attribute__((always_inline)) void blk_sweeps(typename Pair<T>::type* u,
                                                        const typename Pair<T>::type* hb,
                                                        const LaneCols<T>& lc, int gw0, int m,
                                                        bool neu, typename Pair<T>::type* xh,
                                                        int& hbuf, int w, int lane) {
  using V2 = typename Pair<T>::type;
  if constexpr (NS > 0) {
    V2 above, below;
    blk_halo(xh, hbuf, w, lane, u[0], u[RW - 1], above, below);
    blk_halfsweep<T, RW, S, FLIP, EDGE>(u, hb, lc, gw0, m, neu, above, below);
    blk_sweeps<T, RW, S + 1, NS - 1, FLIP, EDGE>(u, hb, lc, gw0, m, neu, xh, hbuf, w, lane);
  }
}

// The body after the loads: hb holds the wave's rows of the right-hand side (zero
// outside the grid).
template <typename T, int NU, int CB, int FLIP, bool EDGE>
__device__ __attribute__((always_inline)) void sdown_blk_body(const VTile& t, int w, int lane,
                                                            typename Pair<T>::type* hb,
                                                            unsigned char* lds) {
  using V2 = typename Pair<T>::type;
  using B = BlkGeo<T, NU, CB>;
  constexpr int OFF = B::kOff, RW = B::kRW, R0 = B::kR0, R1 = B::kR1;
  constexpr int CC = StripGeo<NU>::kDownCols;
  const int m = t.m, n = t.n;
  const int I0 = t.i0;
  const int gr0 = 2 * I0 - OFF, gc0 = 2 * t.j0 - OFF;
  const int a0 = w * RW;       // the wave's first block row
  const int gw0 = gr0 + a0;    // and its grid row
  const bool neu = t.bc == kNeumann;
  LaneCols<T> lc;
  lc.init(gc0, lane, n, neu);
  const T h2 = T(t.h2);
  V2* xh = reinterpret_cast<V2*>(lds);
  V2* rbuf = reinterpret_cast<V2*>(lds + B::kHaloBytes);
  int hbuf = 0;

  // hb' = h^2 b / diag and the zero-guess half-sweep of colour 0, u = -hb' there
  V2 u[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int gi = gw0 + r;
    int vc = 2;
    bool rin = true;
    if (EDGE) {
      rin = gi >= 0 && gi < m;
      vc = row_count(gi, m, neu);
    }
    const T wx = EDGE ? lc.wx(vc) : lc.ix2;
    const T wy = EDGE ? lc.wy(vc) : lc.iy2;
    hb[r].x = (rin && lc.inx) ? wx * h2 * hb[r].x : T(0);
    hb[r].y = (rin && lc.iny) ? wy * h2 * hb[r].y : T(0);
    if (NU > 0 && ((r + FLIP) & 1) == 0) {
      u[r].x = -hb[r].x;
      u[r].y = T(0);
    } else if (NU > 0) {
      u[r].x = T(0);
      u[r].y = -hb[r].y;
    } else {
      u[r] = V2{T(0), T(0)};
    }
  }
  if constexpr (NU > 0)
    blk_sweeps<T, RW, 1, 2 * NU - 1, FLIP, EDGE>(u, hb, lc, gw0, m, neu, xh, hbuf, w, lane);

  // pre-smoothed iterate of the owned fine points (the grid's last block row / strip
  // column owns every remaining fine row / column)
  {
    const int fr0 = 2 * I0, fc0 = 2 * t.j0;
    const int fr1 = (I0 + CB >= t.cm) ? m : min(m, fr0 + 2 * CB);
    const int fc1 = (t.j0 + CC >= t.cn) ? n : min(n, fc0 + 2 * CC);
    GDP_GLOBAL T* v = (GDP_GLOBAL T*)t.v;
    const bool wx = lc.gjx >= fc0 && lc.gjx < fc1;
    const bool wy = lc.gjx + 1 >= fc0 && lc.gjx + 1 < fc1;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int a = a0 + r;
      const int gi = gw0 + r;
      if (a < OFF || a >= OFF + 2 * CB + 2 || gi >= fr1) continue;
      GDP_GLOBAL T* row = v + (size_t)gi * t.pitch + lc.gjx;
      if (wx && wy) {
        *(GDP_GLOBAL V2*)row = u[r];
      } else if (wx) {
        row[0] = u[r].x;
      }
    }
  }

  // scaled residual of the rows the restriction reads (R0 .. R1), with the Neumann
  // column extension, to LDS
  {
    V2 above, below;
    blk_halo(xh, hbuf, w, lane, u[0], u[RW - 1], above, below);
    const bool colext = neu && (gc0 < 0 || gc0 + 128 > n);
    const bool gl = lc.gjx + 1 == -1;
    const bool gr = lc.gjx == n && (n & 1) == 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int a = a0 + r;
      if (a < R0 || a > R1) continue;
      int vc = 2;
      bool rin = true;
      if (EDGE) {
        const int gi = gw0 + r;
        rin = gi >= 0 && gi < m;
        vc = row_count(gi, m, neu);
      }
      const V2 up = r > 0 ? u[r - 1] : above;
      const V2 dn = r < RW - 1 ? u[r + 1] : below;
      const T cx = T(vc) + lc.hx, cy = T(vc) + lc.hy;
      const T nx = ((up.x + dn.x) + dpp_shr1(u[r].y)) + u[r].y;
      const T ny = ((up.y + dn.y) + u[r].x) + dpp_shl1(u[r].x);
      const T rx = fmad(cx, hb[r].x + u[r].x, -nx);
      const T ry = fmad(cy, hb[r].y + u[r].y, -ny);
      V2 res;
      res.x = (rin && lc.inx) ? rx : T(0);
      res.y = (rin && lc.iny) ? ry : T(0);
      if (colext) {
        const T sx = dpp_shl1(res.x);
        const T sy = dpp_shr1(res.y);
        if (gl) res.y = sx;
        if (gr) res.x = sy;
      }
      rbuf[(a - R0) * 64 + lane] = res;
    }
  }
  __syncthreads();

  // restriction of the wave's coarse rows (CB / 4 of them) to the coarse right-hand side
  GDP_GLOBAL T* cb = (GDP_GLOBAL T*)t.cb;
  const T s = T(0.25) * T(t.inv_h2);
  const int jc = lane - OFF / 2;
  const int J = t.j0 + jc;
  const bool wj = jc >= 0 && jc < CC && J < t.cn;
  // residual row a (R0 .. R1) as the restriction reads it: Neumann rows past the grid
  // take the edge row's value (r[-1] = r[0]; r[m] = r[m-1] for even m).  The rows a wave
  // restricts from are not its own register rows, so this test does not follow EDGE.
  auto res_row = [&](int a) {
    if (neu) {
      const int gi = gr0 + a;
      if (gi == -1 && a < R1) a += 1;
      else if (gi == m && (m & 1) == 0 && a > R0) a -= 1;
    }
    return rbuf[(a - R0) * 64 + lane];
  };
  constexpr int CW = CB / kBlkWaves;
  static_assert(CB % kBlkWaves == 0, "coarse rows per block must split evenly over the waves");
#pragma unroll
  for (int k = 0; k < CW; ++k) {
    const int i = w * CW + k;
    const int I = I0 + i;
    T c;
    if (neu) {
      c = restrict_neumann<T>(res_row(2 * i + OFF - 1), res_row(2 * i + OFF),
                              res_row(2 * i + OFF + 1), res_row(2 * i + OFF + 2), s);
    } else {
      c = restrict_dirichlet<T>(res_row(2 * i + OFF), res_row(2 * i + OFF + 1),
                                res_row(2 * i + OFF + 2), s);
    }
    if (I < t.cm && wj) cb[(size_t)I * t.cpitch + J] = c;
  }
}

// UPD: the finest level of a PCG iteration whose update this kernel carries, as the
// streamed kernel's UPD form does (gdp_strip.h): alpha from the <d, q> partials, the lead
// block's state record of the update, the right-hand side -rn of the wave's rows from the
// FP64 r and q rows (rn = r - alpha q, one rounding), b0 = -rn stored at the owned points
// and their rn^2 summed into the block's slot of upart; r is stored by the finest up
// kernel (the block's margin rows are other blocks' owned rows).
template <typename T, int NU, int CB, bool UPD = false>
__global__ __launch_bounds__(kThreads) void k_sdown_blk(const VTile* __restrict__ tiles,
                                                        ProblemState* __restrict__ st,
                                                        double2* __restrict__ upart,
                                                        int check_conv,
                                                        const double2* __restrict__ apart,
                                                        int dcur) {
  using V2 = typename Pair<T>::type;
  using D2 = typename Pair<double>::type;
  using B = BlkGeo<T, NU, CB>;
  constexpr int OFF = B::kOff, RW = B::kRW;
  __shared__ __align__(16) unsigned char lds[B::kLdsBytes];
  const VTile t = tiles[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int gr0 = 2 * t.i0 - OFF, gc0 = 2 * t.j0 - OFF;
  const int gw0 = gr0 + w * RW;
  const int m = t.m, n = t.n;
  const int gjx = gc0 + 2 * lane;
  const bool lin = gjx >= 0 && gjx < n;
  const int jl = lin ? gjx : 0;

  // every load of the wave's rows is issued before anything waits on one: branch-free,
  // from clamped addresses inside the padded rows (gjx is even), zeroed outside the grid
  V2 hb[RW];
  D2 rr[UPD ? RW : 1], qq[UPD ? RW : 1];
  if constexpr (UPD) {
    const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)t.kr;
    const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)t.kq;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const size_t o = (size_t)min(max(gw0 + r, 0), m - 1) * t.pitch + jl;
      rr[r] = *(const GDP_GLOBAL D2*)(kr + o);
      qq[r] = *(const GDP_GLOBAL D2*)(kq + o);
    }
  } else {
    const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int gi = gw0 + r;
      const bool in = lin && gi >= 0 && gi < m;
      const V2 v = *(const GDP_GLOBAL V2*)(b + (size_t)min(max(gi, 0), m - 1) * t.pitch + jl);
      hb[r] = in ? v : V2{T(0), T(0)};
    }
  }
  // a compiler-only fence: it keeps the loads above from being sunk below the state load
  // of the finished-problem test
  asm volatile("" ::: "memory");
  // every wave of the block takes the same exits (block-uniform tests), so the barriers
  // below see all four waves
  __shared__ double shc[2 * kThreads / 64];
  __shared__ double2 bcc;
  double acc = 0.0;
  if constexpr (UPD) {
    // alpha as k_update forms it; the finished-problem test after the reduction's barrier
    const int p = t.problem;
    const double2 a2 = reduce_slots<kThreads>(apart, SlotRange{t.abeg, t.acnt}, shc, &bcc);
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
    // the updated residual of the wave's rows: b0 = -rn and rn^2 at the owned points
    // (rows and columns as the pre-smoothed iterate's), -rn into the strip rows
    const int fr0 = 2 * t.i0, fc0 = 2 * t.j0;
    constexpr int CC = StripGeo<NU>::kDownCols;
    const int fr1 = (t.i0 + CB >= t.cm) ? m : min(m, fr0 + 2 * CB);
    const int fc1 = (t.j0 + CC >= t.cn) ? n : min(n, fc0 + 2 * CC);
    const bool vx = gjx >= fc0 && gjx < fc1;
    const bool vy = gjx + 1 >= fc0 && gjx + 1 < fc1;
    GDP_GLOBAL T* b0 = (GDP_GLOBAL T*)t.b;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int a = w * RW + r;
      const int gi = gw0 + r;
      D2 rn;
      rn.x = __builtin_fma(-alpha, qq[r].x, rr[r].x);
      rn.y = __builtin_fma(-alpha, qq[r].y, rr[r].y);
      const V2 nb = V2{T(-rn.x), T(-rn.y)};
      if (a >= OFF && a < OFF + 2 * CB + 2 && gi < fr1) {
        GDP_GLOBAL T* row = b0 + (size_t)gi * t.pitch + gjx;
        if (vx && vy) {
          *(GDP_GLOBAL V2*)row = nb;
          acc = __builtin_fma(rn.x, rn.x, acc);
          acc = __builtin_fma(rn.y, rn.y, acc);
        } else if (vx) {
          row[0] = nb.x;
          acc = __builtin_fma(rn.x, rn.x, acc);
        }
      }
      hb[r] = (lin && gi >= 0 && gi < m) ? nb : V2{T(0), T(0)};
    }
  } else {
    if (st[t.problem].done) return;
    if (check_conv == t.level + 1) {
      if (check_converged<kThreads>(st, t.problem, upart, SlotRange{t.rbeg, t.rcnt},
                                    t.lead != 0, shc, &bcc))
        return;
    }
  }
  const bool edge = gw0 < 1 || gw0 + RW > m - 1;
  const int flip = (gr0 + gc0) & 1;
  if (edge) {
    if (flip) sdown_blk_body<T, NU, CB, 1, true>(t, w, lane, hb, lds);
    else sdown_blk_body<T, NU, CB, 0, true>(t, w, lane, hb, lds);
  } else {
    if (flip) sdown_blk_body<T, NU, CB, 1, false>(t, w, lane, hb, lds);
    else sdown_blk_body<T, NU, CB, 0, false>(t, w, lane, hb, lds);
  }
  if constexpr (UPD) write_partial<kThreads>(acc, 0.0, upart, t.slot, shc);
}

// ------------------------------------------------------------------ up
// Block-tiled up kernel: the prolongation of the coarse correction + post-smoothing of
// sup_body (gdp_strip.h) for a block owning TB fine rows x kUpCols fine columns, its
// TB + 4 NU rows (rounded up to 4 RW, RW even) shared among the four waves like the down
// kernel's: each wave interpolates its own rows from the coarse rows it reads, and the
// 2 NU half-sweeps exchange the waves' edge rows through LDS.  The margins (2 NU rows
// above and below the owned rows) are the only rows computed twice: 1.5x at TB = 16 and
// NU = 2, against 3x for a strip of 4 rows.  Same expressions, same bits.  DOTS: the
// <r, z>, <q, z> partials of the owned points (finest level of a PCG iteration).
template <typename T, int NU, int TB>
struct UpBlkGeo {
  static constexpr int kH = StripGeo<NU>::kHalo;  // 2 NU
  static constexpr int kFH = TB + 2 * kH;
  static constexpr int kRW = 2 * ((kFH + 2 * kBlkWaves - 1) / (2 * kBlkWaves));
  static constexpr int kNE = kRW / 2 + 2;  // coarse rows a wave interpolates from
  static constexpr int kLdsBytes = 2 * kBlkWaves * 2 * 64 * int(sizeof(T)) * 2;
};

template <typename T, int NU, int TB, int FLIP, bool EDGE, bool DOTS>
__device__ __attribute__((always_inline)) void sup_blk_body(const VTile& t, int w, int lane,
                                                          typename Pair<T>::type* u,
                                                          typename Pair<T>::type* hb,
                                                          const T* E, double& zr, double& zq,
                                                          unsigned char* lds, int rupd,
                                                          double alpha) {
  using V2 = typename Pair<T>::type;
  using B = UpBlkGeo<T, NU, TB>;
  using G = StripGeo<NU>;
  constexpr int H = B::kH, RW = B::kRW, TC = G::kUpCols;
  const int m = t.m, n = t.n;
  const int gr0 = t.i0 - H, gc0 = t.j0 - G::kColHalo;
  const int a0 = w * RW, gw0 = gr0 + a0;
  const bool neu = t.bc == kNeumann;
  LaneCols<T> lc;
  lc.init(gc0, lane, n, neu);
  const T h2 = T(t.h2);
  V2* xh = reinterpret_cast<V2*>(lds);
  int hbuf = 0;

  // u = v + P e (zero outside the grid); hb' = h^2 b / diag
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int gi = gw0 + r;
    int vc = 2;
    bool rin = true;
    if (EDGE) {
      rin = gi >= 0 && gi < m;
      vc = row_count(gi, m, neu);
    }
    T cx, cy;
    if (neu) {
      const int P = r / 2 + 1;
      const int O = (r & 1) ? P + 1 : P - 1;
      // explicit fused multiply-adds: the same bits in every kernel that interpolates
      const T tv = fmad(T(0.25), E[O], T(0.75) * E[P]);
      cx = fmad(T(0.25), dpp_shr1(tv), T(0.75) * tv);
      cy = fmad(T(0.25), dpp_shl1(tv), T(0.75) * tv);
    } else {
      const T tv = (r & 1) ? E[(r + 1) / 2] : T(0.5) * (E[r / 2] + E[r / 2 + 1]);
      cx = T(0.5) * (dpp_shr1(tv) + tv);
      cy = tv;
    }
    u[r].x = (rin && lc.inx) ? u[r].x + cx : T(0);
    u[r].y = (rin && lc.iny) ? u[r].y + cy : T(0);
    const T wx = EDGE ? lc.wx(vc) : lc.ix2;
    const T wy = EDGE ? lc.wy(vc) : lc.iy2;
    hb[r].x = (rin && lc.inx) ? wx * h2 * hb[r].x : T(0);
    hb[r].y = (rin && lc.iny) ? wy * h2 * hb[r].y : T(0);
  }
  blk_sweeps<T, RW, 0, 2 * NU, FLIP, EDGE>(u, hb, lc, gw0, m, neu, xh, hbuf, w, lane);

  GDP_GLOBAL T* uo = (GDP_GLOBAL T*)t.u;
  const int fc1 = min(n, t.j0 + TC);
  const bool wx = lc.gjx >= t.j0 && lc.gjx < fc1;
  const bool wy = lc.gjx + 1 >= t.j0 && lc.gjx + 1 < fc1;
  const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)t.kr;
  const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)t.kq;
  // DOTS: the FP64 residual and q of the wave's rows, loaded together (branch-free,
  // clamped addresses inside the padded rows) before any is used
  using D2 = typename Pair<double>::type;
  D2 rq[DOTS ? 2 * RW : 1];
  if constexpr (DOTS) {
    const int jc = (lc.gjx >= 0 && lc.gjx < n) ? lc.gjx : 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const size_t o = (size_t)min(max(gw0 + r, 0), m - 1) * t.pitch + jc;
      rq[2 * r] = *(const GDP_GLOBAL D2*)(kr + o);
      rq[2 * r + 1] = *(const GDP_GLOBAL D2*)(kq + o);
    }
  }
#pragma unroll
  for (int r = 0; r < RW; ++r) {
    const int a = a0 + r;
    const int gi = gw0 + r;
    if (a < H || a >= H + TB || gi >= m) continue;
    const size_t o = (size_t)gi * t.pitch + lc.gjx;
    if (wx && wy) {
      *(GDP_GLOBAL V2*)(uo + o) = u[r];
    } else if (wx) {
      uo[o] = u[r].x;
    }
    if constexpr (DOTS) {
      // rupd: the residual update left by the fused update-down kernel (as sup_body)
      D2 r2 = rq[2 * r];
      const D2 q2 = rq[2 * r + 1];
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
      zr += double(u[r].x) * rx + double(u[r].y) * ry;
      zq += double(u[r].x) * qx + double(u[r].y) * qy;
    }
  }
}

template <typename T, int NU, int TB, bool DOTS>
__global__ __launch_bounds__(kThreads) void k_sup_blk(const VTile* __restrict__ tiles,
                                                      const ProblemState* __restrict__ st,
                                                      double2* __restrict__ zpart, int order,
                                                      int rupd) {
  using V2 = typename Pair<T>::type;
  using B = UpBlkGeo<T, NU, TB>;
  using G = StripGeo<NU>;
  constexpr int H = B::kH, RW = B::kRW, NE = B::kNE;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ __align__(16) unsigned char lds[B::kLdsBytes];
  const VTile t = tiles[blockIdx.x];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int gr0 = t.i0 - H, gc0 = t.j0 - G::kColHalo;
  const int gw0 = gr0 + w * RW;
  const int m = t.m, n = t.n;
  const bool neu = t.bc == kNeumann;

  // loads: v and b of the wave's rows, and the coarse correction rows they interpolate
  // from (edge-replicated for Neumann, zero outside for Dirichlet), all branch-free from
  // clamped addresses before any is waited on
  V2 u[RW], hb[RW];
  T E[NE];
  {
    const GDP_GLOBAL T* vv = (const GDP_GLOBAL T*)t.v;
    const GDP_GLOBAL T* b = (const GDP_GLOBAL T*)t.b;
    const GDP_GLOBAL T* e = (const GDP_GLOBAL T*)t.cu;
    const int gjx = gc0 + 2 * lane;
    const bool lin = gjx >= 0 && gjx < n;
    const int cr0 = gw0 / 2 - 1;
    int K = gc0 / 2 + lane;
    bool kin = K >= 0 && K < t.cn;
    if (neu) {
      K = min(max(K, 0), t.cn - 1);
      kin = true;
    }
    const int Kc = min(max(K, 0), t.cn - 1);
#pragma unroll
    for (int k = 0; k < NE; ++k) {
      const int I = cr0 + k;
      const bool in = neu || (kin && I >= 0 && I < t.cm);
      const T ev = e[(size_t)min(max(I, 0), t.cm - 1) * t.cpitch + Kc];
      E[k] = in ? ev : T(0);
    }
    const int jc = lin ? gjx : 0;
#pragma unroll
    for (int r = 0; r < RW; ++r) {
      const int gi = gw0 + r;
      const bool in = lin && gi >= 0 && gi < m;
      const size_t o = (size_t)min(max(gi, 0), m - 1) * t.pitch + jc;
      const V2 uv = *(const GDP_GLOBAL V2*)(vv + o);
      const V2 bv = *(const GDP_GLOBAL V2*)(b + o);
      u[r] = in ? uv : V2{T(0), T(0)};
      hb[r] = in ? bv : V2{T(0), T(0)};
    }
  }
  // every wave takes the same exit (a block-uniform test), so the barriers see all four
  if (st[t.problem].done) return;
  const double alpha = (DOTS && rupd) ? st[t.problem].alpha : 0.0;
  double zr = 0.0, zq = 0.0;
  const bool edge = gw0 < 1 || gw0 + RW > m - 1;
  const int flip = ((gr0 + gc0) + order) & 1;
  if (edge) {
    if (flip) sup_blk_body<T, NU, TB, 1, true, DOTS>(t, w, lane, u, hb, E, zr, zq, lds, rupd, alpha);
    else sup_blk_body<T, NU, TB, 0, true, DOTS>(t, w, lane, u, hb, E, zr, zq, lds, rupd, alpha);
  } else {
    if (flip) sup_blk_body<T, NU, TB, 1, false, DOTS>(t, w, lane, u, hb, E, zr, zq, lds, rupd, alpha);
    else sup_blk_body<T, NU, TB, 0, false, DOTS>(t, w, lane, u, hb, E, zr, zq, lds, rupd, alpha);
  }
  if (DOTS) write_partial<kThreads>(zr, zq, zpart, t.slot, sh);
}

}  // namespace gdp
