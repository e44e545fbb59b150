// MI355X (gfx950) kernels of the multigrid-preconditioned CG Poisson solver.
//
// Every kernel is batched: blockIdx.x indexes a host-built Tile table whose entries
// name a level (or Krylov) descriptor and a tile origin, so one launch covers every
// problem of a plan (RGB channels, a batch of images, the requests of a serving
// step).  Per-problem scalars (alpha, beta, delta, ||r||) are reduced on the device:
// each block publishes one partial with a relaxed agent-scope atomic store, the last
// block of a problem (atomic arrival counter) collects the partials in slot order and
// updates the ProblemState.  No block issues a release fence (on gfx950 that writes
// the whole L2 back), and the order of every sum is fixed, so a problem's result
// does not depend on what else is in the batch.
//
// Arithmetic follows gradient-domain/src/gradient_domain/:
//   smoother      multigrid.py:147-166  u <- (sum of neighbours - h^2 b) / diag, colour by colour
//   residual      multigrid.py:210      r = b - laplacian(u, h)     (operators.py:86-107)
//   restriction   multigrid.py:169-174  0.25 P_r^T r P_c
//   prolongation  multigrid.py:177-181  P_r e P_c^T, added to u (multigrid.py:214)
//   coarsest      multigrid.py:206-207  coarsest_sweeps sweeps (Dirichlet)
//   PCG           solvers.py:134-160    q = -L d, alpha = delta / <d, q>, ...
// The Neumann variants (diag = number of in-grid neighbours, cell-centred transfers,
// exact DCT coarsest solve) have no counterpart in the reference.
#pragma once
#include "gdp_device.h"

namespace gdp {

constexpr int kThreads = 256;

// ------------------------------------------------------------------ reductions
// Sum over the block; the result is valid in thread 0.  Fixed order: shuffle tree
// inside each wave, then waves in index order.
template <int NT>
__device__ inline double block_sum(double v, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) t += sh[i];
  }
  return t;
}

// Deterministic sum of the partial slots [begin, begin + count) written by an earlier
// kernel, broadcast to every thread of the block (each thread sums a contiguous run,
// then the block tree): every block that calls it gets bit-identical totals.
template <int NT>
__device__ inline double2 reduce_slots(const double2* __restrict__ partial, SlotRange rg,
                                       double* sh, double2* bc) {
  const int per = (rg.count + NT - 1) / NT;
  const int lo = rg.begin + threadIdx.x * per;
  const int hi = min(lo + per, rg.begin + rg.count);
  double sx = 0.0, sy = 0.0;
  for (int i = lo; i < hi; ++i) {
    const double2 v = partial[i];
    sx += v.x;
    sy += v.y;
  }
  const double tx = block_sum<NT>(sx, sh);
  const double ty = block_sum<NT>(sy, sh + NT / 64);
  if (threadIdx.x == 0) *bc = make_double2(tx, ty);
  __syncthreads();
  return *bc;
}

// This block's contribution (a, b) to a per-problem sum, for a later kernel to reduce.
template <int NT>
__device__ inline void write_partial(double a, double b, double2* partial, int slot, double* sh) {
  const double ba = block_sum<NT>(a, sh);
  const double bb = block_sum<NT>(b, sh + NT / 64);
  if (threadIdx.x == 0) partial[slot] = make_double2(ba, bb);
}

// Convergence test after a PCG update (solvers.py:154): every block of the first kernel
// that touches problem p reaches the same decision from the update's partial sums of
// ||r||^2; the lead block records it.  Returns true when the problem is done.
template <int NT>
__device__ inline bool check_converged(ProblemState* st, int p, const double2* upart,
                                       SlotRange rg, bool lead, double* sh, double2* bc) {
  const double2 t = reduce_slots<NT>(upart, rg, sh, bc);
  const bool conv = t.x <= st[p].target2;
  const bool done = conv || st[p].iter >= st[p].max_iter;
  if (lead && threadIdx.x == 0) {
    st[p].rr = t.x;
    st[p].converged = conv;
    st[p].done = done;
  }
  return done;
}

__device__ inline int neumann_count(int gi, int gj, int m, int n) {
  return (gi > 0) + (gi < m - 1) + (gj > 0) + (gj < n - 1);
}

// One Gauss-Seidel update: (sum of neighbours - h^2 b) / diag.
template <typename T>
__device__ inline T smooth_value(T nb, T b, T h2, bool neu, int gi, int gj, int m, int n) {
  const T val = nb - h2 * b;
  if (neu) {
    // 1/count by selection: a true division expands to a long scaled-division sequence
    const int cnt = neumann_count(gi, gj, m, n);
    const T inv = cnt == 4 ? T(0.25) : (cnt == 3 ? T(1.0 / 3.0) : (cnt == 2 ? T(0.5) : T(1)));
    return val * inv;
  }
  return val * T(0.25);
}

template <typename T>
__device__ inline T diag_of(bool neu, int gi, int gj, int m, int n) {
  return neu ? T(neumann_count(gi, gj, m, n)) : T(4);
}

// Row taps of the prolongation at fine index i (coarse indices ia, ib, weights wa, wb).
// Dirichlet (multigrid.py:111-133): odd i -> coarse (i-1)/2; even i -> mean of coarse
// i/2-1 and i/2, zero outside.  Neumann: 3/4 parent i/2 + 1/4 next-nearest, with the
// coarse grid edge-replicated so a missing neighbour reflects onto the parent.
template <typename T>
__device__ inline void ptaps(bool neu, int i, int& ia, int& ib, T& wa, T& wb) {
  if (neu) {
    ia = i >> 1;
    ib = (i & 1) ? ia + 1 : ia - 1;
    wa = T(0.75);
    wb = T(0.25);
  } else if (i & 1) {
    ia = ib = (i - 1) >> 1;
    wa = T(1);
    wb = T(0);
  } else {
    ia = (i >> 1) - 1;
    ib = i >> 1;
    wa = wb = T(0.5);
  }
}

// ------------------------------------------------------------------ interior fast paths
// A tile whose whole LDS region (halo included) lies strictly inside the grid needs no
// bounds tests: every point has four in-grid neighbours (diag = 4 for Dirichlet and
// Neumann alike) and the transfers never reach a boundary.  Threads map to the region
// as (row = wave + 4 i, column = lane + 64 j), so no index arithmetic is divided.

template <typename T, int FH, int FW>
__device__ inline void interior_halfsweep(T* su, const T* sb, int s, int parity, int par0,
                                          T h2) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int lo = 1 + s, hr = FH - 1 - s, hc = FW - 1 - s;
  constexpr int NRR = (FH + 3) / 4;
  constexpr int NCC = (FW + 127) / 128;
  T val[NRR][NCC];
  int at[NRR][NCC];
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr) {
    const int a = lo + wv + 4 * rr;
    const int c0 = lo + ((parity + a + par0 + lo) & 1);
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int c = c0 + 2 * lane + 128 * cc;
      const bool ok = a < hr && c < hc;
      const int idx = ok ? a * FW + c : FW + 1;
      at[rr][cc] = ok ? idx : -1;
      const T nb = ((su[idx - FW] + su[idx + FW]) + su[idx - 1]) + su[idx + 1];
      val[rr][cc] = (nb - h2 * sb[idx]) * T(0.25);
    }
  }
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc)
      if (at[rr][cc] >= 0) su[at[rr][cc]] = val[rr][cc];
  __syncthreads();
}

template <typename T, int CH, int CW, int NU>
__device__ void down_interior(const VTile& t, ProblemState* st, const double2* upart,
                              int check_conv, T* su, T* sb) {
  constexpr int OFF = 2 + 2 * NU;
  constexpr int FH = 2 * CH + 2 + 2 * OFF, FW = 2 * CW + 2 + 2 * OFF;
  constexpr int NRR = (FH + 3) / 4, NCC = (FW + 63) / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gr0 = 2 * t.i0 - OFF, gc0 = 2 * t.j0 - OFF;
  const int pitch = t.pitch;
  const int par0 = (gr0 + gc0) & 1;
  const T h2 = T(t.h2), ih2 = T(t.inv_h2);
  const T* __restrict__ b = static_cast<const T*>(t.b) + (size_t)gr0 * pitch + gc0;
  T bl[NRR][NCC];
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr) {
    const int a = wv + 4 * rr;
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int c = lane + 64 * cc;
      const bool in = a >= 1 && a < FH - 1 && c >= 1 && c < FW - 1;
      bl[rr][cc] = in ? b[(size_t)a * pitch + c] : T(0);
    }
  }
  if (st[t.problem].done) return;
  if (check_conv && t.level == 0) {
    __shared__ double shc[2 * kThreads / 64];
    __shared__ double2 bcc;
    if (check_converged<kThreads>(st, t.problem, upart, SlotRange{t.rbeg, t.rcnt}, t.lead != 0,
                                  shc, &bcc))
      return;
  }
  // b into LDS; the first colour-0 half-sweep from u = 0 folded in
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr) {
    const int a = wv + 4 * rr;
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int c = lane + 64 * cc;
      if (a < FH && c < FW) {
        const int k = a * FW + c;
        const T bv = bl[rr][cc];
        sb[k] = bv;
        const bool upd = NU > 0 && a >= 1 && a < FH - 1 && c >= 1 && c < FW - 1 &&
                         ((a + c + par0) & 1) == 0;
        su[k] = upd ? -h2 * bv * T(0.25) : T(0);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 1; s < 2 * NU; ++s) interior_halfsweep<T, FH, FW>(su, sb, s, s & 1, par0, h2);
  // residual over fine rows / cols [fr0-1, fr0+2CH+2) in place of b; owned iterate out
  constexpr int RRH = 2 * CH + 3, RRW = 2 * CW + 3;
  constexpr int NR2 = (RRH + 3) / 4, NC2 = (RRW + 63) / 64;
  T* __restrict__ v = static_cast<T*>(t.v) + (size_t)gr0 * pitch + gc0;
  T rv[NR2][NC2];
#pragma unroll
  for (int rr = 0; rr < NR2; ++rr) {
    const int a = OFF - 1 + wv + 4 * rr;
#pragma unroll
    for (int cc = 0; cc < NC2; ++cc) {
      const int c = OFF - 1 + lane + 64 * cc;
      const bool ok = a < OFF - 1 + RRH && c < OFF - 1 + RRW;
      const int idx = ok ? a * FW + c : FW + 1;
      const T uc = su[idx];
      const T nb = ((su[idx - FW] + su[idx + FW]) + su[idx - 1]) + su[idx + 1];
      rv[rr][cc] = sb[idx] - (nb - T(4) * uc) * ih2;
      if (ok && a >= OFF && a < OFF + 2 * CH && c >= OFF && c < OFF + 2 * CW)
        v[(size_t)a * pitch + c] = uc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int rr = 0; rr < NR2; ++rr) {
    const int a = OFF - 1 + wv + 4 * rr;
#pragma unroll
    for (int cc = 0; cc < NC2; ++cc) {
      const int c = OFF - 1 + lane + 64 * cc;
      if (a < OFF - 1 + RRH && c < OFF - 1 + RRW) sb[a * FW + c] = rv[rr][cc];
    }
  }
  __syncthreads();
  const bool neu = t.bc == kNeumann;
  T* __restrict__ cb = static_cast<T*>(t.cb);
  for (int k = threadIdx.x; k < CH * CW; k += kThreads) {
    const int ci = k / CW, cj = k - ci * CW;
    T acc;
    if (neu) {
      const T* r0 = &sb[(2 * ci + OFF - 1) * FW + 2 * cj + OFF - 1];
      T col[4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
        col[x] = T(0.25) * r0[x] + T(0.75) * r0[FW + x] + T(0.75) * r0[2 * FW + x] +
                 T(0.25) * r0[3 * FW + x];
      acc = T(0.25) * col[0] + T(0.75) * col[1] + T(0.75) * col[2] + T(0.25) * col[3];
    } else {
      const T* r0 = &sb[(2 * ci + OFF) * FW + 2 * cj + OFF];
      T col[3];
#pragma unroll
      for (int x = 0; x < 3; ++x)
        col[x] = T(0.5) * r0[x] + r0[FW + x] + T(0.5) * r0[2 * FW + x];
      acc = T(0.5) * col[0] + col[1] + T(0.5) * col[2];
    }
    cb[(size_t)(t.i0 + ci) * t.cpitch + t.j0 + cj] = T(0.25) * acc;
  }
}

template <typename T, int TH, int TW, int NU, bool DOTS>
__device__ void up_interior(const VTile& t, const ProblemState* st, double2* zpart, int order,
                            T* su, T* sb, T* se, double* sh) {
  constexpr int H = 2 * NU;
  constexpr int RH = TH + 2 * H, RW = TW + 2 * H;
  constexpr int CRH = RH / 2 + 4, CRW = RW / 2 + 4;
  constexpr int NRR = (RH + 3) / 4, NCC = (RW + 63) / 64;
  constexpr int NER = (CRH + 3) / 4, NEC = (CRW + 63) / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gr0 = t.i0 - H, gc0 = t.j0 - H;
  const int cr0 = (gr0 >> 1) - 1, cc0 = (gc0 >> 1) - 1;
  const int pitch = t.pitch;
  const int par0 = (gr0 + gc0) & 1;
  const T h2 = T(t.h2);
  const T* __restrict__ e = static_cast<const T*>(t.cu) + (size_t)cr0 * t.cpitch + cc0;
  const T* __restrict__ vv = static_cast<const T*>(t.v) + (size_t)gr0 * pitch + gc0;
  const T* __restrict__ b = static_cast<const T*>(t.b) + (size_t)gr0 * pitch + gc0;
  T el[NER][NEC], vl[NRR][NCC], bl[NRR][NCC];
#pragma unroll
  for (int rr = 0; rr < NER; ++rr)
#pragma unroll
    for (int cc = 0; cc < NEC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      el[rr][cc] = (a < CRH && c < CRW) ? e[(size_t)a * t.cpitch + c] : T(0);
    }
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      const bool in = a < RH && c < RW;
      vl[rr][cc] = in ? vv[(size_t)a * pitch + c] : T(0);
      bl[rr][cc] = in ? b[(size_t)a * pitch + c] : T(0);
    }
  if (st[t.problem].done) return;
#pragma unroll
  for (int rr = 0; rr < NER; ++rr)
#pragma unroll
    for (int cc = 0; cc < NEC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      if (a < CRH && c < CRW) se[a * CRW + c] = el[rr][cc];
    }
  __syncthreads();
  const bool neu = t.bc == kNeumann;
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      if (a < RH && c < RW) {
        int ia, ib, ja, jb;
        T wa, wb, va, vb;
        ptaps<T>(neu, gr0 + a, ia, ib, wa, wb);
        ptaps<T>(neu, gc0 + c, ja, jb, va, vb);
        const int oa = (ia - cr0) * CRW - cc0, ob = (ib - cr0) * CRW - cc0;
        const T corr = va * (wa * se[oa + ja] + wb * se[ob + ja]) +
                       vb * (wa * se[oa + jb] + wb * se[ob + jb]);
        su[a * RW + c] = vl[rr][cc] + corr;
        sb[a * RW + c] = bl[rr][cc];
      }
    }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 2 * NU; ++s) {
    const int parity = order == 0 ? (s & 1) : 1 - (s & 1);
    interior_halfsweep<T, RH, RW>(su, sb, s, parity, par0, h2);
  }
  T* __restrict__ uo = static_cast<T*>(t.u) + (size_t)t.i0 * pitch + t.j0;
  constexpr int NOR = (TH + 3) / 4, NOC = (TW + 63) / 64;
  double zr = 0.0, zq = 0.0;
  double rl[DOTS ? NOR : 1][DOTS ? NOC : 1], ql[DOTS ? NOR : 1][DOTS ? NOC : 1];
  if (DOTS) {
    const double* kr = t.kr + (size_t)t.i0 * pitch + t.j0;
    const double* kq = t.kq + (size_t)t.i0 * pitch + t.j0;
#pragma unroll
    for (int rr = 0; rr < NOR; ++rr)
#pragma unroll
      for (int cc = 0; cc < NOC; ++cc) {
        const int a = wv + 4 * rr, c = lane + 64 * cc;
        const bool in = a < TH && c < TW;
        rl[rr][cc] = in ? kr[(size_t)a * pitch + c] : 0.0;
        ql[rr][cc] = in ? kq[(size_t)a * pitch + c] : 0.0;
      }
  }
#pragma unroll
  for (int rr = 0; rr < NOR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NOC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      if (a < TH && c < TW) {
        const T z = su[(a + H) * RW + c + H];
        uo[(size_t)a * pitch + c] = z;
        if (DOTS) {
          zr += double(z) * rl[rr][cc];
          zq += double(z) * ql[rr][cc];
        }
      }
    }
  if (DOTS) write_partial<kThreads>(zr, zq, zpart, t.slot, sh);
}

// ------------------------------------------------------------------ down kernel
// Pre-smoothing (NU sweeps from a zero guess) + residual + restriction, fused.
// One block produces a CH x CW tile of the coarse right-hand side and the
// pre-smoothed fine iterate of the fine points it owns.
//
// Restriction 0.25 P^T r: Dirichlet taps 2I..2I+2 with weights (1/2, 1, 1/2); Neumann
// taps 2I-1..2I+2 with (1/4, 3/4, 3/4, 1/4) applied to the residual extended by
// r[-1] = r[0] and, for even sizes, r[f] = r[f-1] (odd sizes: zero), which is exactly
// the transpose of the reflected cell-centred interpolation.
//
// With check_conv the finest-level launch of a PCG iteration first applies the
// convergence test of the preceding update (check_converged).
template <typename T, int CH, int CW, int NU>
__global__ __launch_bounds__(kThreads) void k_down(const VTile* __restrict__ tiles,
                                                   ProblemState* __restrict__ st,
                                                   const double2* __restrict__ upart,
                                                   int check_conv) {
  const VTile t = tiles[blockIdx.x];
  constexpr int OFF = 2 + 2 * NU;
  constexpr int FH = 2 * CH + 2 + 2 * OFF, FW = 2 * CW + 2 + 2 * OFF;
  constexpr int RRH = 2 * CH + 3, RRW = 2 * CW + 3;  // residual rows fr0-1 .. fr0+2CH+1
  __shared__ T su[FH * FW];
  __shared__ T sb[FH * FW];  // b, then the residual (in place)
  const int m = t.m, n = t.n, pitch = t.pitch;
  const int fr0 = 2 * t.i0, fc0 = 2 * t.j0;
  const int gr0 = fr0 - OFF, gc0 = fc0 - OFF;
  const bool neu = t.bc == kNeumann;
  const T h2 = T(t.h2), ih2 = T(t.inv_h2);
  const T* __restrict__ b = static_cast<const T*>(t.b);
  if (gr0 >= 1 && gr0 + FH <= m - 1 && gc0 >= 1 && gc0 + FW <= n - 1 && t.i0 + CH < t.cm &&
      t.j0 + CW < t.cn) {
    down_interior<T, CH, CW, NU>(t, st, upart, check_conv, su, sb);
    return;
  }

  // load b; the first colour-0 half-sweep from u = 0 is folded into the load.  All of a
  // thread's loads are issued before any is consumed (fixed trip count, predicated).
  constexpr int NL = (FH * FW + kThreads - 1) / kThreads;
  T bl[NL];
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / FW, c = k - a * FW;
    const int gi = gr0 + a, gj = gc0 + c;
    const bool in = k < FH * FW && a >= 1 && a < FH - 1 && c >= 1 && c < FW - 1 && gi >= 0 &&
                    gi < m && gj >= 0 && gj < n;
    bl[it] = in ? b[(size_t)gi * pitch + gj] : T(0);
  }
  if (st[t.problem].done) return;
  if (check_conv && t.level == 0) {
    // convergence test of the preceding PCG update (the loads above are in flight)
    __shared__ double shc[2 * kThreads / 64];
    __shared__ double2 bcc;
    if (check_converged<kThreads>(st, t.problem, upart, SlotRange{t.rbeg, t.rcnt}, t.lead != 0,
                                  shc, &bcc))
      return;
  }
#pragma unroll
  for (int it = 0; it < NL; ++it) {
    const int k = threadIdx.x + it * kThreads;
    if (k >= FH * FW) break;
    const int a = k / FW, c = k - a * FW;
    const int gi = gr0 + a, gj = gc0 + c;
    T uv = T(0);
    if (NU > 0 && a >= 1 && a < FH - 1 && c >= 1 && c < FW - 1 && gi >= 0 && gi < m &&
        gj >= 0 && gj < n && ((gi + gj) & 1) == 0)
      uv = smooth_value<T>(T(0), bl[it], h2, neu, gi, gj, m, n);
    sb[k] = bl[it];
    su[k] = uv;
  }
  __syncthreads();

  // After half-sweep s the points of [s+1, F-1-s) hold exactly what a whole-grid sweep
  // would give; the ring outside that region is only read, never trusted.
#pragma unroll
  for (int s = 1; s < 2 * NU; ++s) {
    const int parity = s & 1;
    const int lo = 1 + s, hr = FH - 1 - s, hc = FW - 1 - s;
    const int rows = hr - lo, half = (hc - lo + 1) >> 1;
    for (int k = threadIdx.x; k < rows * half; k += kThreads) {
      const int a = lo + k / half;
      int c = lo + 2 * (k % half);
      const int gi = gr0 + a;
      int gj = gc0 + c;
      if (((gi + gj) & 1) != parity) { ++c; ++gj; }
      if (c >= hc || gi < 0 || gi >= m || gj < 0 || gj >= n) continue;
      const int idx = a * FW + c;
      const T nb = ((su[idx - FW] + su[idx + FW]) + su[idx - 1]) + su[idx + 1];
      su[idx] = smooth_value<T>(nb, sb[idx], h2, neu, gi, gj, m, n);
    }
    __syncthreads();
  }

  // residual on fine rows/cols [fr0-1, fr0+2CH+2), written over b (the residual of a
  // point reads b only at that point), and the pre-smoothed iterate of the owned points
  const int fr1 = (t.i0 + CH >= t.cm) ? m : min(m, fr0 + 2 * CH);
  const int fc1 = (t.j0 + CW >= t.cn) ? n : min(n, fc0 + 2 * CW);
  T* __restrict__ v = static_cast<T*>(t.v);
  constexpr int NRR = (RRH * RRW + kThreads - 1) / kThreads;
  T rv[NRR];
#pragma unroll
  for (int it = 0; it < NRR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int ra = k / RRW, rc = k - ra * RRW;
    const int si = fr0 - 1 + ra, sj = fc0 - 1 + rc;
    T r = T(0);
    if (k < RRH * RRW && si >= 0 && si < m && sj >= 0 && sj < n) {
      const int idx = (si - gr0) * FW + (sj - gc0);
      const T uc = su[idx];
      const T nb = ((su[idx - FW] + su[idx + FW]) + su[idx - 1]) + su[idx + 1];
      r = sb[idx] - (nb - diag_of<T>(neu, si, sj, m, n) * uc) * ih2;
      if (si < fr1 && sj < fc1 && si >= fr0 && sj >= fc0) v[(size_t)si * pitch + sj] = uc;
    }
    rv[it] = r;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NRR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int ra = k / RRW, rc = k - ra * RRW;
    if (k < RRH * RRW) sb[(ra + OFF - 1) * FW + rc + OFF - 1] = rv[it];
  }
  __syncthreads();
  if (neu) {
    // Neumann extension of the residual past the grid: r[-1] = r[0], r[f] = r[f-1] for
    // even f, zero otherwise (corners take both rules)
#pragma unroll
    for (int it = 0; it < NRR; ++it) {
      const int k = threadIdx.x + it * kThreads;
      const int ra = k / RRW, rc = k - ra * RRW;
      const int si = fr0 - 1 + ra, sj = fc0 - 1 + rc;
      if (k >= RRH * RRW || (si >= 0 && si < m && sj >= 0 && sj < n)) continue;
      int ti = si, tj = sj;
      bool ok = true;
      if (ti == -1) ti = 0;
      else if (ti == m && (m & 1) == 0) ti = m - 1;
      else if (ti < 0 || ti >= m) ok = false;
      if (tj == -1) tj = 0;
      else if (tj == n && (n & 1) == 0) tj = n - 1;
      else if (tj < 0 || tj >= n) ok = false;
      rv[it] = ok ? sb[(ti - gr0) * FW + (tj - gc0)] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < NRR; ++it) {
      const int k = threadIdx.x + it * kThreads;
      const int ra = k / RRW, rc = k - ra * RRW;
      const int si = fr0 - 1 + ra, sj = fc0 - 1 + rc;
      if (k >= RRH * RRW || (si >= 0 && si < m && sj >= 0 && sj < n)) continue;
      sb[(ra + OFF - 1) * FW + rc + OFF - 1] = rv[it];
    }
    __syncthreads();
  }

  T* __restrict__ cb = static_cast<T*>(t.cb);
  for (int k = threadIdx.x; k < CH * CW; k += kThreads) {
    const int ci = k / CW, cj = k - ci * CW;
    const int I = t.i0 + ci, J = t.j0 + cj;
    if (I >= t.cm || J >= t.cn) continue;
    T acc;
    if (neu) {
      const T* r0 = &sb[(2 * ci + OFF - 1) * FW + 2 * cj + OFF - 1];  // fine (2I-1, 2J-1)
      T col[4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
        col[x] = T(0.25) * r0[x] + T(0.75) * r0[FW + x] + T(0.75) * r0[2 * FW + x] +
                 T(0.25) * r0[3 * FW + x];
      acc = T(0.25) * col[0] + T(0.75) * col[1] + T(0.75) * col[2] + T(0.25) * col[3];
    } else {
      const T* r0 = &sb[(2 * ci + OFF) * FW + 2 * cj + OFF];  // fine (2I, 2J)
      T col[3];
#pragma unroll
      for (int x = 0; x < 3; ++x)
        col[x] = T(0.5) * r0[x] + r0[FW + x] + T(0.5) * r0[2 * FW + x];
      acc = T(0.5) * col[0] + col[1] + T(0.5) * col[2];
    }
    cb[(size_t)I * t.cpitch + J] = T(0.25) * acc;
  }
}

// ------------------------------------------------------------------ up kernel
// Prolongation of the coarse correction + post-smoothing (NU sweeps), fused.  With
// DOTS (finest level of a PCG iteration) the block also accumulates <r, z> and <q, z>
// and the problem's last block turns them into delta and beta.
template <typename T, int TH, int TW, int NU, bool DOTS>
__global__ __launch_bounds__(kThreads) void k_up(const VTile* __restrict__ tiles,
                                                 const ProblemState* __restrict__ st,
                                                 double2* __restrict__ zpart, int order) {
  const VTile t = tiles[blockIdx.x];
  const int p = t.problem;
  constexpr int H = 2 * NU;
  constexpr int RH = TH + 2 * H, RW = TW + 2 * H;
  constexpr int CRH = RH / 2 + 4, CRW = RW / 2 + 4;
  __shared__ T su[RH * RW];
  __shared__ T sb[RH * RW];
  __shared__ T se[CRH * CRW];
  __shared__ double sh[2 * kThreads / 64];
  const int m = t.m, n = t.n, pitch = t.pitch;
  const int gr0 = t.i0 - H, gc0 = t.j0 - H;
  const int cr0 = (gr0 >> 1) - 1, cc0 = (gc0 >> 1) - 1;
  const bool neu = t.bc == kNeumann;
  const T h2 = T(t.h2);
  const T* __restrict__ e = static_cast<const T*>(t.cu);
  const T* __restrict__ vv = static_cast<const T*>(t.v);
  const T* __restrict__ b = static_cast<const T*>(t.b);
  const int Cm = t.cm, Cn = t.cn, Cp = t.cpitch;
  if (gr0 >= 1 && gr0 + RH <= m - 1 && gc0 >= 1 && gc0 + RW <= n - 1 && cr0 >= 0 &&
      cr0 + CRH <= Cm && cc0 >= 0 && cc0 + CRW <= Cn) {
    up_interior<T, TH, TW, NU, DOTS>(t, st, zpart, order, su, sb, se, sh);
    return;
  }
  // Issue every global load of the block first: the coarse correction (with its
  // padding: zero outside for Dirichlet, edge values for Neumann) and the fine v and b.
  constexpr int NE = (CRH * CRW + kThreads - 1) / kThreads;
  constexpr int NR = (RH * RW + kThreads - 1) / kThreads;
  T el[NE], vl[NR], bl[NR];
#pragma unroll
  for (int it = 0; it < NE; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / CRW, c = k - a * CRW;
    int I = cr0 + a, J = cc0 + c;
    bool in = k < CRH * CRW;
    if (neu) {
      I = min(max(I, 0), Cm - 1);
      J = min(max(J, 0), Cn - 1);
    } else {
      in = in && I >= 0 && I < Cm && J >= 0 && J < Cn;
    }
    el[it] = in ? e[(size_t)I * Cp + J] : T(0);
  }
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    const bool in = k < RH * RW && gi >= 0 && gi < m && gj >= 0 && gj < n;
    const size_t o = in ? (size_t)gi * pitch + gj : 0;
    vl[it] = in ? vv[o] : T(0);
    bl[it] = in ? b[o] : T(0);
  }
  if (st[p].done) return;
#pragma unroll
  for (int it = 0; it < NE; ++it) {
    const int k = threadIdx.x + it * kThreads;
    if (k < CRH * CRW) se[k] = el[it];
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    if (k >= RH * RW) break;
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    T uv = T(0);
    if (gi >= 0 && gi < m && gj >= 0 && gj < n) {
      int ia, ib, ja, jb;
      T wa, wb, va, vb;
      ptaps<T>(neu, gi, ia, ib, wa, wb);
      ptaps<T>(neu, gj, ja, jb, va, vb);
      const int oa = (ia - cr0) * CRW - cc0, ob = (ib - cr0) * CRW - cc0;
      const T corr = va * (wa * se[oa + ja] + wb * se[ob + ja]) +
                     vb * (wa * se[oa + jb] + wb * se[ob + jb]);
      uv = vl[it] + corr;
    }
    su[k] = uv;
    sb[k] = bl[it];
  }
  __syncthreads();

#pragma unroll
  for (int s = 0; s < 2 * NU; ++s) {
    const int parity = order == 0 ? (s & 1) : 1 - (s & 1);
    const int lo = 1 + s, hr = RH - 1 - s, hc = RW - 1 - s;
    const int rows = hr - lo, half = (hc - lo + 1) >> 1;
    for (int k = threadIdx.x; k < rows * half; k += kThreads) {
      const int a = lo + k / half;
      int c = lo + 2 * (k % half);
      const int gi = gr0 + a;
      int gj = gc0 + c;
      if (((gi + gj) & 1) != parity) { ++c; ++gj; }
      if (c >= hc || gi < 0 || gi >= m || gj < 0 || gj >= n) continue;
      const int idx = a * RW + c;
      const T nb = ((su[idx - RW] + su[idx + RW]) + su[idx - 1]) + su[idx + 1];
      su[idx] = smooth_value<T>(nb, sb[idx], h2, neu, gi, gj, m, n);
    }
    __syncthreads();
  }

  T* __restrict__ uo = static_cast<T*>(t.u);
  double zr = 0.0, zq = 0.0;
  constexpr int NO = (TH * TW + kThreads - 1) / kThreads;
  double rl[DOTS ? NO : 1], ql[DOTS ? NO : 1];
  if (DOTS) {
    // the finest level shares the Krylov grid's pitch
#pragma unroll
    for (int it = 0; it < NO; ++it) {
      const int k = threadIdx.x + it * kThreads;
      const int a = k / TW, c = k - a * TW;
      const int gi = t.i0 + a, gj = t.j0 + c;
      const bool in = k < TH * TW && gi < m && gj < n;
      const size_t o = in ? (size_t)gi * pitch + gj : 0;
      rl[it] = in ? t.kr[o] : 0.0;
      ql[it] = in ? t.kq[o] : 0.0;
    }
  }
#pragma unroll
  for (int it = 0; it < NO; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (k >= TH * TW || gi >= m || gj >= n) continue;
    const T z = su[(a + H) * RW + c + H];
    uo[(size_t)gi * pitch + gj] = z;
    if (DOTS) {
      zr += double(z) * rl[it];
      zq += double(z) * ql[it];
    }
  }
  if (DOTS) write_partial<kThreads>(zr, zq, zpart, t.slot, sh);
}

// ------------------------------------------------------------------ coarse tail
// One block per problem runs every level from TailDesc.first to the coarsest in LDS:
// pre-smooth / residual / restrict down, FP64 coarsest solve, prolong / post-smooth up.
// Loops run rows over waves and columns over lanes (no integer division).
constexpr int kTailWaves = kTailThreads / 64;

// Tail loops put rows on waves and columns on lanes (no division), and a wave with no
// row of the level branches over the whole pass, so a small level costs only the
// waves it needs: on one CU every executed wave instruction costs the SIMD 4 cycles
// whether or not its lanes are active.

// Red-black half-sweep of colour `parity`; with `from_zero` the iterate is known to be
// zero, so colour `parity` gets (-h^2 b)/diag and the other colour is cleared.
template <typename T>
__device__ void lds_halfsweep(T* u, const T* b, int m, int n, T h2, bool neu, int parity,
                              bool from_zero) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < m; i += kTailWaves) {
    T* ur = u + i * n;
    const T* br = b + i * n;
    if (from_zero) {
      for (int j = lane; j < n; j += 64)
        ur[j] = ((i + j) & 1) == parity ? smooth_value<T>(T(0), br[j], h2, neu, i, j, m, n)
                                        : T(0);
      continue;
    }
    for (int j = ((i + parity) & 1) + 2 * lane; j < n; j += 128) {
      const T up = i > 0 ? ur[j - n] : T(0);
      const T dn = i < m - 1 ? ur[j + n] : T(0);
      const T lf = j > 0 ? ur[j - 1] : T(0);
      const T rt = j < n - 1 ? ur[j + 1] : T(0);
      ur[j] = smooth_value<T>(((up + dn) + lf) + rt, br[j], h2, neu, i, j, m, n);
    }
  }
  __syncthreads();
}

// r = b - A u (r may not alias u or b)
template <typename T>
__device__ void lds_residual(const T* u, const T* b, T* r, int m, int n, T ih2, bool neu) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < m; i += kTailWaves)
    for (int j = lane; j < n; j += 64) {
      const int o = i * n + j;
      const T up = i > 0 ? u[o - n] : T(0);
      const T dn = i < m - 1 ? u[o + n] : T(0);
      const T lf = j > 0 ? u[o - 1] : T(0);
      const T rt = j < n - 1 ? u[o + 1] : T(0);
      r[o] = b[o] - ((((up + dn) + lf) + rt) - diag_of<T>(neu, i, j, m, n) * u[o]) * ih2;
    }
  __syncthreads();
}

// Neumann residual extension used by the restriction (see k_down).
__device__ inline int nsrc(int i, int f, bool& ok) {
  if (i == -1) return 0;
  if (i >= f) {
    if (i == f && (f & 1) == 0) return f - 1;
    ok = false;
    return 0;
  }
  return i;
}

template <typename T>
__device__ void lds_restrict(const T* r, int m, int n, T* bc, int mc, int nc, bool neu) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int I = wave; I < mc; I += kTailWaves)
    for (int J = lane; J < nc; J += 64) {
      T acc = T(0);
      if (neu) {
        const T w[4] = {T(0.25), T(0.75), T(0.75), T(0.25)};
        int ii[4], jj[4];
        bool oi[4], oj[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          oi[x] = true;
          oj[x] = true;
          ii[x] = nsrc(2 * I - 1 + x, m, oi[x]);
          jj[x] = nsrc(2 * J - 1 + x, n, oj[x]);
        }
#pragma unroll
        for (int y = 0; y < 4; ++y) {
          T col = T(0);
#pragma unroll
          for (int x = 0; x < 4; ++x)
            col += (oi[x] && oj[y]) ? w[x] * r[ii[x] * n + jj[y]] : T(0);
          acc += w[y] * col;
        }
      } else {
        const T w[3] = {T(0.5), T(1), T(0.5)};
#pragma unroll
        for (int y = 0; y < 3; ++y) {
          const int j = 2 * J + y;
          T col = T(0);
#pragma unroll
          for (int x = 0; x < 3; ++x) col += w[x] * r[(2 * I + x) * n + j];
          acc += w[y] * col;
        }
      }
      bc[I * nc + J] = T(0.25) * acc;
    }
  __syncthreads();
}

template <typename T>
__device__ inline T lds_coarse(const T* e, int mc, int nc, int I, int J, bool neu) {
  if (neu) {
    I = min(max(I, 0), mc - 1);
    J = min(max(J, 0), nc - 1);
    return e[I * nc + J];
  }
  return (I >= 0 && I < mc && J >= 0 && J < nc) ? e[I * nc + J] : T(0);
}

template <typename T>
__device__ void lds_prolong_add(T* u, int m, int n, const T* e, int mc, int nc, bool neu) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < m; i += kTailWaves) {
    int ia, ib;
    T wa, wb;
    ptaps<T>(neu, i, ia, ib, wa, wb);
    for (int j = lane; j < n; j += 64) {
      int ja, jb;
      T va, vb;
      ptaps<T>(neu, j, ja, jb, va, vb);
      const T corr = va * (wa * lds_coarse(e, mc, nc, ia, ja, neu) +
                           wb * lds_coarse(e, mc, nc, ib, ja, neu)) +
                     vb * (wa * lds_coarse(e, mc, nc, ia, jb, neu) +
                           wb * lds_coarse(e, mc, nc, ib, jb, neu));
      u[i * n + j] += corr;
    }
  }
  __syncthreads();
}

// Exact solve of the coarsest level by an orthonormal transform that diagonalises its
// operator: U = Cm^T ((Cm B Cn^T) ./ lambda) Cn.  Neumann: DCT-II (the solvers.py:224-262
// construction, zero-mean solution: 1/lambda of the constant mode is 0).  Dirichlet:
// DST-I.  Tables live in LDS: Cm (m*m), Cn (n*n), 1/lambda (m*n).  Each pass puts
// output rows on waves and output columns on lanes.
__device__ void lds_dct_solve(double* w0, double* w1, int m, int n, const double* tab) {
  const double* Cm = tab;
  const double* Cn = tab + m * m;
  const double* il = tab + m * m + n * n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = wave; k < m; k += kTailWaves)  // w1 = Cm w0
    for (int j = lane; j < n; j += 64) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += Cm[k * m + i] * w0[i * n + j];
      w1[k * n + j] = s;
    }
  __syncthreads();
  for (int k = wave; k < m; k += kTailWaves)  // w0 = (w1 Cn^T) ./ lambda
    for (int l = lane; l < n; l += 64) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += w1[k * n + j] * Cn[l * n + j];
      w0[k * n + l] = s * il[k * n + l];
    }
  __syncthreads();
  for (int i = wave; i < m; i += kTailWaves)  // w1 = Cm^T w0
    for (int l = lane; l < n; l += 64) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) s += Cm[k * m + i] * w0[k * n + l];
      w1[i * n + l] = s;
    }
  __syncthreads();
  for (int i = wave; i < m; i += kTailWaves)  // w0 = w1 Cn
    for (int j = lane; j < n; j += 64) {
      double s = 0.0;
      for (int l = 0; l < n; ++l) s += w1[i * n + l] * Cn[l * n + j];
      w0[i * n + j] = s;
    }
  __syncthreads();
}

// Many sweeps on a level of at most 64 x 8 points: one wave does them all, keeping its
// points in registers and exchanging through LDS with wave-level ordering only (LDS
// operations of a wave complete in order), then the block synchronises once.
__device__ void wave_sweeps(double* u, const double* b, int m, int n, double h2, int sweeps) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    for (int s = 0; s < sweeps; ++s)
      for (int parity = 0; parity < 2; ++parity) {
        for (int o = lane; o < m * n; o += 64) {
          const int i = o / n, j = o - i * n;
          if (((i + j) & 1) != parity) continue;
          const double up = i > 0 ? u[o - n] : 0.0;
          const double dn = i < m - 1 ? u[o + n] : 0.0;
          const double lf = j > 0 ? u[o - 1] : 0.0;
          const double rt = j < n - 1 ? u[o + 1] : 0.0;
          u[o] = ((((up + dn) + lf) + rt) - h2 * b[o]) * 0.25;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      }
  }
  __syncthreads();
}

template <typename T, bool DOTS>
__global__ __launch_bounds__(kTailThreads) void k_tail(const TailDesc* __restrict__ tails,
                                                       const LevelDesc* __restrict__ levels,
                                                       ProblemState* __restrict__ st,
                                                       const KrylovDesc* __restrict__ kds,
                                                       const double2* __restrict__ upart,
                                                       const SlotRange* __restrict__ rk,
                                                       double2* __restrict__ zpart,
                                                       const SlotRange* __restrict__ rz,
                                                       int check_conv, int pre, int post,
                                                       int coarsest_sweeps, int order) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ TailDesc D;
  __shared__ LevelDesc LV[kMaxTailLevels];
  __shared__ double sh[2 * kTailThreads / 64];
  __shared__ double2 bcc;
  // the descriptors are read once, into LDS
  {
    const int* src = reinterpret_cast<const int*>(&tails[blockIdx.x]);
    int* dst = reinterpret_cast<int*>(&D);
    for (int w = threadIdx.x; w < int(sizeof(TailDesc) / 4); w += kTailThreads) dst[w] = src[w];
    __syncthreads();
    const int* ls = reinterpret_cast<const int*>(&levels[D.first]);
    int* ld = reinterpret_cast<int*>(LV);
    const int nw = D.nlev * int(sizeof(LevelDesc) / 4);
    for (int w = threadIdx.x; w < nw; w += kTailThreads) ld[w] = ls[w];
    __syncthreads();
  }
  const int p = D.problem;
  if (st[p].done) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // a problem whose finest level is in the tail gets its convergence test here
  if (check_conv && D.krylov >= 0 &&
      check_converged<kTailThreads>(st, p, upart, rk[p], true, sh, &bcc))
    return;

  {  // right-hand side of the first tail level, and the DCT tables
    const LevelDesc& L = LV[0];
    T* b = reinterpret_cast<T*>(smem + D.lds_off[0][1]);
    const T* g = static_cast<const T*>(L.b);
    for (int i = wave; i < L.m; i += kTailWaves)
      for (int j = lane; j < L.n; j += 64) b[i * L.n + j] = g[(size_t)i * L.pitch + j];
    if (D.dct) {
      const LevelDesc& Lc = LV[D.nlev - 1];
      const int nt = Lc.m * Lc.m + Lc.n * Lc.n + Lc.m * Lc.n;
      double* tb = reinterpret_cast<double*>(smem + D.tab_off);
      for (int o = threadIdx.x; o < nt; o += kTailThreads) tb[o] = D.dct[o];
    }
    __syncthreads();
  }
  T* scratch = reinterpret_cast<T*>(smem + D.scratch_off);
  for (int l = 0; l + 1 < D.nlev; ++l) {
    const LevelDesc& L = LV[l];
    const LevelDesc& C = LV[l + 1];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    T* b = reinterpret_cast<T*>(smem + D.lds_off[l][1]);
    T* bc = reinterpret_cast<T*>(smem + D.lds_off[l + 1][1]);
    const bool neu = L.bc == kNeumann;
    const T h2 = T(L.h2);
    if (pre == 0) {
      for (int i = wave; i < L.m; i += kTailWaves)
        for (int j = lane; j < L.n; j += 64) u[i * L.n + j] = T(0);
      __syncthreads();
    }
    for (int s = 0; s < pre; ++s) {
      lds_halfsweep<T>(u, b, L.m, L.n, h2, neu, 0, s == 0);
      lds_halfsweep<T>(u, b, L.m, L.n, h2, neu, 1, false);
    }
    lds_residual<T>(u, b, scratch, L.m, L.n, T(L.inv_h2), neu);
    lds_restrict<T>(scratch, L.m, L.n, bc, C.m, C.n, neu);
  }
  {  // coarsest level, FP64
    const int l = D.nlev - 1;
    const LevelDesc& L = LV[l];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    const T* b = reinterpret_cast<const T*>(smem + D.lds_off[l][1]);
    double* w0 = reinterpret_cast<double*>(smem + D.coarse_off);
    double* w1 = w0 + L.m * L.n;
    const int N = L.m * L.n;
    if (D.dct) {
      // exact: DCT-II (Neumann) or DST-I (Dirichlet, coarsest_sweeps == 0) diagonalisation
      for (int o = threadIdx.x; o < N; o += kTailThreads) w0[o] = double(b[o]) * L.h2;
      __syncthreads();
      lds_dct_solve(w0, w1, L.m, L.n, reinterpret_cast<const double*>(smem + D.tab_off));
    } else {
      for (int o = threadIdx.x; o < N; o += kTailThreads) {
        w0[o] = 0.0;
        w1[o] = double(b[o]);
      }
      __syncthreads();
      if (N <= 512) {
        wave_sweeps(w0, w1, L.m, L.n, L.h2, coarsest_sweeps);
      } else {
        for (int s = 0; s < coarsest_sweeps; ++s) {
          lds_halfsweep<double>(w0, w1, L.m, L.n, L.h2, false, 0, false);
          lds_halfsweep<double>(w0, w1, L.m, L.n, L.h2, false, 1, false);
        }
      }
    }
    for (int o = threadIdx.x; o < N; o += kTailThreads) u[o] = T(w0[o]);
    __syncthreads();
  }
  for (int l = D.nlev - 2; l >= 0; --l) {
    const LevelDesc& L = LV[l];
    const LevelDesc& C = LV[l + 1];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    const T* b = reinterpret_cast<const T*>(smem + D.lds_off[l][1]);
    const T* e = reinterpret_cast<const T*>(smem + D.lds_off[l + 1][0]);
    const bool neu = L.bc == kNeumann;
    lds_prolong_add<T>(u, L.m, L.n, e, C.m, C.n, neu);
    for (int s = 0; s < post; ++s) {
      const int p0 = order == 0 ? 0 : 1;
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, p0, false);
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, 1 - p0, false);
    }
  }
  // write the first tail level's iterate
  const LevelDesc& L = LV[0];
  const T* u = reinterpret_cast<const T*>(smem + D.lds_off[0][0]);
  T* g = static_cast<T*>(L.u);
  double zr = 0.0, zq = 0.0;
  const bool dots = DOTS && D.krylov >= 0;
  const double* kr = dots ? kds[D.krylov].r : nullptr;
  const double* kq = dots ? kds[D.krylov].q : nullptr;
  for (int i = wave; i < L.m; i += kTailWaves)
    for (int j = lane; j < L.n; j += 64) {
      const size_t go = (size_t)i * L.pitch + j;
      const T z = u[i * L.n + j];
      g[go] = z;
      if (dots) {
        zr += double(z) * kr[go];
        zq += double(z) * kq[go];
      }
    }
  if (dots) write_partial<kTailThreads>(zr, zq, zpart, rz[p].begin, sh);
}

// ------------------------------------------------------------------ Krylov kernels
// FP64 vectors of the fine grid, tiles of TH x TW points (TH * TW a multiple of the
// block size).  Scalar data flow of one PCG iteration (solvers.py:149-159):
//   k_apply   reads the V-cycle's <r,z>, <q,z> partials -> delta, beta;
//             d = z + beta d, q = -A d, partial <d, q>
//   k_update  reads the <d, q> partials -> alpha = delta / <d, q>;
//             x += alpha d, r -= alpha q, V-cycle RHS -r, partial ||r||^2
//   first V-cycle kernel: convergence test from the ||r||^2 partials
// Each kernel reads only state fields written by earlier kernels and its lead block
// writes only fields that no block of the same kernel reads.

template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_apply(const KTile* __restrict__ tiles,
                                                    ProblemState* __restrict__ st,
                                                    const double2* __restrict__ zpart,
                                                    double2* __restrict__ apart, int din,
                                                    int jacobi) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  constexpr int RH = TH + 2, RW = TW + 2;
  constexpr int NRR = (RH + 3) / 4, NCC = (RW + 63) / 64;
  __shared__ double sd[RH * RW];
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = t.m, n = t.n, pitch = t.pitch;
  const int gr0 = t.i0 - 1, gc0 = t.j0 - 1;
  // the region (halo included) inside the grid: no tests, every diagonal is 4
  const bool interior = gr0 >= 0 && gr0 + RH <= m && gc0 >= 0 && gc0 + RW <= n;
  const T* __restrict__ z = static_cast<const T*>(t.z);
  const double* __restrict__ dold = din ? t.d1 : t.d0;
  double* __restrict__ dnew = din ? t.d0 : t.d1;
  // vector loads first; the scalar reduction below overlaps them
  T zl[NRR][NCC];
  double dl[NRR][NCC];
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      const int gi = gr0 + a, gj = gc0 + c;
      const bool in = a < RH && c < RW && (interior || (gi >= 0 && gi < m && gj >= 0 && gj < n));
      const size_t o = in ? (size_t)gi * pitch + gj : 0;
      zl[rr][cc] = in ? z[o] : T(0);
      dl[rr][cc] = in ? dold[o] : 0.0;
    }
  if (st[p].done) return;
  const double2 z2 = reduce_slots<kThreads>(
      zpart, jacobi ? SlotRange{t.kbeg, t.kcnt} : SlotRange{t.zbeg, t.zcnt}, sh, &bc);
  double beta;
  if (st[p].iter == 0 || st[p].restart) beta = 0.0;
  else if (st[p].flexible) beta = -st[p].alpha * z2.y / st[p].delta_old;
  else beta = z2.x / st[p].delta_old;
  if (t.lead && threadIdx.x == 0) {
    st[p].delta = z2.x;
    st[p].beta = beta;
  }
#pragma unroll
  for (int rr = 0; rr < NRR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NCC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      if (a < RH && c < RW) sd[a * RW + c] = double(zl[rr][cc]) + beta * dl[rr][cc];
    }
  __syncthreads();
  const bool neu = t.bc == kNeumann;
  constexpr int NOR = (TH + 3) / 4, NOC = (TW + 63) / 64;
  double acc = 0.0;
#pragma unroll
  for (int rr = 0; rr < NOR; ++rr)
#pragma unroll
    for (int cc = 0; cc < NOC; ++cc) {
      const int a = wv + 4 * rr, c = lane + 64 * cc;
      const int gi = t.i0 + a, gj = t.j0 + c;
      if (a >= TH || c >= TW || gi >= m || gj >= n) continue;
      const int idx = (a + 1) * RW + c + 1;
      const double dc = sd[idx];
      const double nb = ((sd[idx - RW] + sd[idx + RW]) + sd[idx - 1]) + sd[idx + 1];
      const double dg = interior ? 4.0 : diag_of<double>(neu, gi, gj, m, n);
      const double q = -(nb - dg * dc) * t.inv_h2;
      const size_t o = (size_t)gi * pitch + gj;
      dnew[o] = dc;
      t.q[o] = q;
      acc += dc * q;
    }
  write_partial<kThreads>(acc, 0.0, apart, t.slot, sh);
}

template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_update(const KTile* __restrict__ tiles,
                                                     ProblemState* __restrict__ st,
                                                     const double2* __restrict__ apart,
                                                     double2* __restrict__ upart, int dcur) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  const int m = t.m, n = t.n, pitch = t.pitch;
  T* __restrict__ b0 = static_cast<T*>(t.b0);
  const double* __restrict__ d = dcur ? t.d1 : t.d0;
  constexpr int NO = (TH * TW) / kThreads;
  double xl[NO], dl[NO], rl[NO], ql[NO];
  size_t ol[NO];
  bool il[NO];
#pragma unroll
  for (int it = 0; it < NO; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    il[it] = gi < m && gj < n;
    ol[it] = il[it] ? (size_t)gi * pitch + gj : 0;
    xl[it] = il[it] ? t.x[ol[it]] : 0.0;
    dl[it] = il[it] ? d[ol[it]] : 0.0;
    rl[it] = il[it] ? t.r[ol[it]] : 0.0;
    ql[it] = il[it] ? t.q[ol[it]] : 0.0;
  }
  if (st[p].done) return;
  const double2 a2 = reduce_slots<kThreads>(apart, SlotRange{t.kbeg, t.kcnt}, sh, &bc);
  const double delta = st[p].delta;
  const double alpha = delta / a2.x;
  if (t.lead && threadIdx.x == 0) {
    st[p].alpha = alpha;
    st[p].dq = a2.x;
    st[p].delta_old = delta;
    st[p].iter += 1;
    st[p].restart = 0;
  }
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < NO; ++it) {
    if (!il[it]) continue;
    const double r = rl[it] - alpha * ql[it];
    t.x[ol[it]] = xl[it] + alpha * dl[it];
    t.r[ol[it]] = r;
    b0[ol[it]] = T(-r);
    acc += r * r;
  }
  write_partial<kThreads>(acc, 0.0, upart, t.slot, sh);
}

// Jacobi preconditioner of the reference "cg" method (solvers.py:114-117):
// z = r / (4 / h^2) and its <r, z> partials (consumed by k_apply with beta = delta'/delta).
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_jacobi(const KTile* __restrict__ tiles,
                                                     ProblemState* __restrict__ st,
                                                     const double2* __restrict__ upart,
                                                     double2* __restrict__ zpart,
                                                     int check_conv) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  if (st[p].done) return;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  if (check_conv && check_converged<kThreads>(st, p, upart, SlotRange{t.kbeg, t.kcnt},
                                              t.lead != 0, sh, &bc))
    return;
  T* __restrict__ z = static_cast<T*>(t.z);
  const double diag = 4.0 * t.inv_h2;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    const size_t o = (size_t)gi * t.pitch + gj;
    const double r = t.r[o];
    const double zz = r / diag;
    z[o] = T(zz);
    acc += r * double(T(zz));
  }
  write_partial<kThreads>(acc, 0.0, zpart, t.slot, sh);
}

// Set-up pass 1: partial sums of b (mean for the Neumann projection onto compatible
// right-hand sides -- the constant that solvers.py:224-262 drops).
template <int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_init_stats(const KTile* __restrict__ tiles,
                                                         const IoDesc* __restrict__ io,
                                                         double2* __restrict__ spart) {
  const KTile t = tiles[blockIdx.x];
  __shared__ double sh[2 * kThreads / 64];
  const double* __restrict__ b = io[t.problem].b;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    acc += (gi < t.m && gj < t.n) ? b[(size_t)gi * t.n + gj] : 0.0;
  }
  write_partial<kThreads>(acc, 0.0, spart, t.slot, sh);
}

// Set-up pass 2.  mode 0 (PCG): r = -(b - mean), x = d = q = 0, V-cycle RHS b - mean.
// mode 1 (MG iteration): x = initial guess (or 0).  Partial sums of (b - mean)^2.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_init(const KTile* __restrict__ tiles,
                                                   const IoDesc* __restrict__ io,
                                                   const double* const* __restrict__ initial,
                                                   const ProblemState* __restrict__ st,
                                                   double2* __restrict__ spart, int mode) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  __shared__ double sh[2 * kThreads / 64];
  const double mean = st[p].bmean;
  const double* __restrict__ b = io[p].b;
  const double* __restrict__ x0 = initial ? initial[p] : nullptr;
  T* __restrict__ b0 = static_cast<T*>(t.b0);
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    const size_t o = (size_t)gi * t.pitch + gj;
    const double bb = b[(size_t)gi * t.n + gj] - mean;
    acc += bb * bb;
    if (mode == 0) {
      t.r[o] = -bb;
      t.x[o] = 0.0;
      t.d0[o] = 0.0;
      t.d1[o] = 0.0;
      t.q[o] = 0.0;
      b0[o] = T(bb);
    } else {
      t.x[o] = x0 ? x0[(size_t)gi * t.n + gj] : 0.0;
    }
  }
  write_partial<kThreads>(acc, 0.0, spart, t.slot, sh);
}

// r_true = (b - mean) - L x (FP64): partial sums of ||r_true||^2 and of x; the V-cycle
// RHS becomes r_true and, in mode 0 (PCG), r = -r_true so an unconverged problem can
// restart from it.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_residual(const KTile* __restrict__ tiles,
                                                       const IoDesc* __restrict__ io,
                                                       const ProblemState* __restrict__ st,
                                                       double2* __restrict__ spart, int mode,
                                                       int respect_done) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  if (respect_done && st[p].done) return;
  constexpr int RH = TH + 2, RW = TW + 2;
  __shared__ double sx[RH * RW];
  __shared__ double sh[2 * kThreads / 64];
  const int m = t.m, n = t.n, pitch = t.pitch;
  const double mean = st[p].bmean;
  const double* __restrict__ b = io[p].b;
  T* __restrict__ b0 = static_cast<T*>(t.b0);
  const int gr0 = t.i0 - 1, gc0 = t.j0 - 1;
  constexpr int NR = (RH * RW + kThreads - 1) / kThreads;
  double xl[NR];
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    const bool in = k < RH * RW && gi >= 0 && gi < m && gj >= 0 && gj < n;
    xl[it] = in ? t.x[(size_t)gi * pitch + gj] : 0.0;
  }
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const int k = threadIdx.x + it * kThreads;
    if (k < RH * RW) sx[k] = xl[it];
  }
  __syncthreads();
  const bool neu = t.bc == kNeumann;
  double rr = 0.0, sum = 0.0;
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const int idx = (a + 1) * RW + c + 1;
    const double xc = sx[idx];
    const double nb = ((sx[idx - RW] + sx[idx + RW]) + sx[idx - 1]) + sx[idx + 1];
    const double lx = (nb - diag_of<double>(neu, gi, gj, m, n) * xc) * t.inv_h2;
    const double res = (b[(size_t)gi * n + gj] - mean) - lx;
    const size_t o = (size_t)gi * pitch + gj;
    if (mode == 0) t.r[o] = -res;
    b0[o] = T(res);
    rr += res * res;
    sum += xc;
  }
  write_partial<kThreads>(rr, sum, spart, t.slot, sh);
}

// One block per problem: reduce a set-up / residual pass and update the state.
//   kind 0 (after k_init_stats): bmean
//   kind 1 (after k_init): ||b||^2, target, fresh iteration state
//   kind 2 (after k_residual): true ||r||^2, sum(x); MG modes also record the history
//          and test rel < tolerance (multigrid.py:250-254)
__global__ __launch_bounds__(kThreads) void k_scalar(const KrylovDesc* __restrict__ kds,
                                                     ProblemState* __restrict__ st,
                                                     const double2* __restrict__ spart,
                                                     const SlotRange* __restrict__ rk, int kind,
                                                     int mode, int project, double tol,
                                                     int max_iter, int flexible, int zero_done,
                                                     int respect_done, double* history,
                                                     int hist_stride) {
  const int p = blockIdx.x;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  if (kind == 2 && respect_done && st[p].done) return;
  const double2 tot = reduce_slots<kThreads>(spart, rk[p], sh, &bc);
  if (threadIdx.x != 0) return;
  ProblemState& S = st[p];
  const KrylovDesc& K = kds[p];
  if (kind == 0) {
    S.bmean = (project && K.bc == kNeumann) ? tot.x / (double(K.m) * double(K.n)) : 0.0;
  } else if (kind == 1) {
    S.bnorm2 = tot.x;
    S.target2 = tol * tol * tot.x;
    S.iter = 0;
    S.converged = zero_done && tot.x == 0.0;
    S.done = S.converged || max_iter <= 0;
    S.restart = 0;
    S.alpha = S.beta = S.delta = S.delta_old = 0.0;
    S.rr = tot.x;
    S.max_iter = max_iter;
    S.flexible = flexible;
  } else {
    S.true_rr = tot.x;
    S.sumx = tot.y;
    if (mode >= 1) {
      const double rel = S.bnorm2 > 0.0 ? sqrt(tot.x / S.bnorm2) : 0.0;
      if (mode == 2) S.iter += 1;
      if (history) history[(size_t)p * hist_stride + S.iter] = rel;
      S.rr = tot.x;
      if (mode == 2) {
        S.converged = rel < tol;
        S.done = S.converged || S.iter >= S.max_iter;
      }
    }
  }
}

// MG iteration: x += z (the V-cycle correction).
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_correct(const KTile* __restrict__ tiles,
                                                      const ProblemState* __restrict__ st) {
  const KTile t = tiles[blockIdx.x];
  if (st[t.problem].done) return;
  const T* __restrict__ z = static_cast<const T*>(t.z);
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    const size_t o = (size_t)gi * t.pitch + gj;
    t.x[o] += double(z[o]);
  }
}

// Caller output: u = x - mean(x) for Neumann problems when mean_fix, else u = x.
template <int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_finalize(const KTile* __restrict__ tiles,
                                                       const IoDesc* __restrict__ io,
                                                       const ProblemState* __restrict__ st,
                                                       int mean_fix) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  const double mean = (mean_fix && t.bc == kNeumann)
                          ? st[p].sumx / (double(t.m) * double(t.n))
                          : 0.0;
  double* __restrict__ u = io[p].u;
#pragma unroll
  for (int it = 0; it < (TH * TW) / kThreads; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    u[(size_t)gi * t.n + gj] = t.x[(size_t)gi * t.pitch + gj] - mean;
  }
}

// Restart bookkeeping after the FP64 true-residual check: a problem whose recursive
// residual converged but whose true residual did not is reopened with a fresh Krylov
// sequence (residual replacement) if it has iterations left.
__global__ void k_reopen(ProblemState* __restrict__ st, int count, int* __restrict__ reopened) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= count) return;
  ProblemState& S = st[p];
  const bool bad = S.true_rr > S.target2 && S.bnorm2 > 0.0;
  if (bad && S.iter < S.max_iter) {
    S.converged = 0;
    S.done = 0;
    S.restart = 1;
    atomicAdd(reopened, 1);
  } else if (bad) {
    S.converged = 0;
  }
}

}  // namespace gdp
