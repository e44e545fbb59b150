// MI355X (gfx950) kernels of the multigrid-preconditioned CG Poisson solver.
//
// Every kernel is batched: blockIdx.x indexes a host-built Tile table whose entries
// name a level (or Krylov) descriptor and a tile origin, so one launch covers every
// problem of a plan (RGB channels, a batch of images, the requests of a serving
// step).  Per-problem scalars (alpha, beta, delta, ||r||) are reduced on the device:
// each block writes one partial, the last block of a problem (atomic arrival
// counter) sums the partials in slot order and updates the ProblemState.  The order
// of every sum is fixed, so a problem's result does not depend on what else is in
// the batch.
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

// Publish this block's partial and report whether it is the problem's last block.
__device__ inline bool arrive_last(int* counter, int total, int* s_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = atomicAdd(counter, 1);
    *s_flag = (prev == total - 1);
  }
  __syncthreads();
  const bool last = *s_flag != 0;
  if (last) __threadfence();
  return last;
}

// Sum of partial slots [begin, begin + count) in slot order (last block only).
template <int NT>
__device__ inline double2 gather_partials(const double2* partial, SlotRange rg, double* sh) {
  // Each thread sums a contiguous run so the order is a fixed function of count.
  const int per = (rg.count + NT - 1) / NT;
  const int lo = rg.begin + threadIdx.x * per;
  const int hi = min(lo + per, rg.begin + rg.count);
  double sx = 0.0, sy = 0.0;
  for (int i = lo; i < hi; ++i) {
    const double2 p = partial[i];
    sx += p.x;
    sy += p.y;
  }
  const double tx = block_sum<NT>(sx, sh);
  const double ty = block_sum<NT>(sy, sh + NT / 64);
  return make_double2(tx, ty);
}

__device__ inline int neumann_count(int gi, int gj, int m, int n) {
  return (gi > 0) + (gi < m - 1) + (gj > 0) + (gj < n - 1);
}

template <typename T>
__device__ inline T smooth_value(T nb, T b, T h2, bool neu, int gi, int gj, int m, int n) {
  T val = nb - h2 * b;
  if (neu) {
    const int cnt = neumann_count(gi, gj, m, n);
    return cnt == 4 ? val * T(0.25) : val / T(cnt);
  }
  return val * T(0.25);
}

template <typename T>
__device__ inline T diag_of(bool neu, int gi, int gj, int m, int n) {
  return neu ? T(neumann_count(gi, gj, m, n)) : T(4);
}

// 1-D transfer weights.  Dirichlet (multigrid.py:111-133): coarse I sits on fine
// 2I+1; even fine points average their two coarse neighbours; zero outside.
// Neumann: cell-centred, 3/4 parent + 1/4 next-nearest, reflected at the ends.
struct Taps {
  int idx[2];
  double w[2];
  int count;
};

__device__ inline Taps prolong_taps(bool neu, int i, int mc) {
  Taps t;
  t.count = 0;
  if (!neu) {
    if (i & 1) {
      const int s = (i - 1) >> 1;
      if (s >= 0 && s < mc) { t.idx[t.count] = s; t.w[t.count] = 1.0; ++t.count; }
    } else {
      const int s0 = (i >> 1) - 1, s1 = i >> 1;
      if (s0 >= 0 && s0 < mc) { t.idx[t.count] = s0; t.w[t.count] = 0.5; ++t.count; }
      if (s1 >= 0 && s1 < mc) { t.idx[t.count] = s1; t.w[t.count] = 0.5; ++t.count; }
    }
  } else {
    const int p = i >> 1;
    const int o = (i & 1) ? p + 1 : p - 1;
    if (o < 0 || o >= mc) { t.idx[0] = p; t.w[0] = 1.0; t.count = 1; }
    else { t.idx[0] = p; t.w[0] = 0.75; t.idx[1] = o; t.w[1] = 0.25; t.count = 2; }
  }
  return t;
}

// Restriction weight of fine point (base + k), k = 0..3, for coarse I; base = 2I (Dirichlet,
// taps k = 0..2) or 2I - 1 (Neumann, taps k = 0..3).  f = fine size, c = coarse size.
__device__ inline double restrict_weight(bool neu, int I, int k, int f, int c) {
  if (!neu) return k == 1 ? 1.0 : (k == 3 ? 0.0 : 0.5);
  const int i = 2 * I - 1 + k;
  if (i < 0 || i >= f) return 0.0;
  if (k == 0 || k == 3) return 0.25;
  if (k == 1) return I == 0 ? 1.0 : 0.75;
  return I == c - 1 ? 1.0 : 0.75;
}

// ------------------------------------------------------------------ down kernel
// Pre-smoothing (NU sweeps from a zero guess) + residual + restriction, fused.
// One block produces a CH x CW tile of the coarse right-hand side and the
// pre-smoothed fine iterate of the fine points it owns.
template <typename T, int CH, int CW, int NU>
__global__ __launch_bounds__(kThreads) void k_down(const Tile* __restrict__ tiles,
                                                   const LevelDesc* __restrict__ levels,
                                                   const ProblemState* __restrict__ st) {
  const Tile t = tiles[blockIdx.x];
  const LevelDesc L = levels[t.desc];
  if (st[L.problem].done) return;
  const LevelDesc C = levels[L.coarse];
  constexpr int OFF = 2 + 2 * NU;
  constexpr int FH = 2 * CH + 2 + 2 * OFF, FW = 2 * CW + 2 + 2 * OFF;
  __shared__ T su[FH * FW];
  __shared__ T sb[FH * FW];
  __shared__ T sr[FH * FW];
  const int m = L.m, n = L.n, pitch = L.pitch;
  const int fr0 = 2 * t.i0, fc0 = 2 * t.j0;
  const int gr0 = fr0 - OFF, gc0 = fc0 - OFF;
  const bool neu = L.bc == kNeumann;
  const T h2 = T(L.h2), ih2 = T(L.inv_h2);
  const T* __restrict__ b = static_cast<const T*>(L.b);

  for (int k = threadIdx.x; k < FH * FW; k += kThreads) {
    const int a = k / FW, c = k - a * FW;
    const int gi = gr0 + a, gj = gc0 + c;
    T bv = T(0);
    if (a >= 1 && a < FH - 1 && c >= 1 && c < FW - 1 && gi >= 0 && gi < m && gj >= 0 && gj < n)
      bv = b[(size_t)gi * pitch + gj];
    sb[k] = bv;
    su[k] = T(0);
  }
  __syncthreads();

  // After half-sweep s the points of [s+1, F-1-s) hold exactly what a whole-grid sweep
  // would give; the ring outside that region is only read, never trusted.
#pragma unroll
  for (int s = 0; s < 2 * NU; ++s) {
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

  // Residual on the core [OFF-1, OFF+2CH+2) and the pre-smoothed iterate on owned points.
  const int fr1 = (t.i0 + CH >= C.m) ? m : min(m, fr0 + 2 * CH);
  const int fc1 = (t.j0 + CW >= C.n) ? n : min(n, fc0 + 2 * CW);
  T* __restrict__ v = static_cast<T*>(L.v);
  constexpr int CR = 2 * CH + 3, CC = 2 * CW + 3;
  for (int k = threadIdx.x; k < CR * CC; k += kThreads) {
    const int ca = k / CC, cc = k - ca * CC;
    const int a = OFF - 1 + ca, c = OFF - 1 + cc;
    const int gi = gr0 + a, gj = gc0 + c;
    const int idx = a * FW + c;
    T r = T(0);
    if (gi >= 0 && gi < m && gj >= 0 && gj < n) {
      const T uc = su[idx];
      const T nb = ((su[idx - FW] + su[idx + FW]) + su[idx - 1]) + su[idx + 1];
      r = sb[idx] - (nb - diag_of<T>(neu, gi, gj, m, n) * uc) * ih2;
      if (gi >= fr0 && gi < fr1 && gj >= fc0 && gj < fc1) v[(size_t)gi * pitch + gj] = uc;
    }
    sr[idx] = r;
  }
  __syncthreads();

  T* __restrict__ cb = static_cast<T*>(C.b);
  const int ntap = neu ? 4 : 3;
  const int tap0 = neu ? -1 : 0;
  for (int k = threadIdx.x; k < CH * CW; k += kThreads) {
    const int I = t.i0 + k / CW, J = t.j0 + k % CW;
    if (I >= C.m || J >= C.n) continue;
    const int a0 = 2 * (I - t.i0) + OFF + tap0, c0 = 2 * (J - t.j0) + OFF + tap0;
    double acc = 0.0;
    for (int cj = 0; cj < ntap; ++cj) {
      const double wc = restrict_weight(neu, J, cj, n, C.n);
      if (wc == 0.0) continue;
      double col = 0.0;
      for (int ri = 0; ri < ntap; ++ri) {
        const double wr = restrict_weight(neu, I, ri, m, C.m);
        if (wr == 0.0) continue;
        col += wr * double(sr[(a0 + ri) * FW + c0 + cj]);
      }
      acc += wc * col;
    }
    cb[(size_t)I * C.pitch + J] = T(0.25 * acc);
  }
}

// ------------------------------------------------------------------ up kernel
// Prolongation of the coarse correction + post-smoothing (NU sweeps), fused.  With
// DOTS (finest level of a PCG iteration) the block also accumulates <r, z> and <q, z>
// and the problem's last block turns them into delta and beta.
template <typename T, int TH, int TW, int NU, bool DOTS>
__global__ __launch_bounds__(kThreads) void k_up(const Tile* __restrict__ tiles,
                                                 const LevelDesc* __restrict__ levels,
                                                 ProblemState* __restrict__ st,
                                                 const KrylovDesc* __restrict__ kds,
                                                 double2* __restrict__ partial,
                                                 int* __restrict__ counters,
                                                 const SlotRange* __restrict__ ranges,
                                                 int order) {
  const Tile t = tiles[blockIdx.x];
  const LevelDesc L = levels[t.desc];
  const int p = L.problem;
  if (st[p].done) return;
  const LevelDesc C = levels[L.coarse];
  constexpr int H = 2 * NU;
  constexpr int RH = TH + 2 * H, RW = TW + 2 * H;
  constexpr int CRH = RH / 2 + 4, CRW = RW / 2 + 4;
  __shared__ T su[RH * RW];
  __shared__ T sb[RH * RW];
  __shared__ T se[CRH * CRW];
  __shared__ double sh[2 * kThreads / 64];
  __shared__ int s_flag;
  const int m = L.m, n = L.n, pitch = L.pitch;
  const int gr0 = t.i0 - H, gc0 = t.j0 - H;
  const int cr0 = (gr0 >> 1) - 1, cc0 = (gc0 >> 1) - 1;
  const bool neu = L.bc == kNeumann;
  const T h2 = T(L.h2);
  const T* __restrict__ e = static_cast<const T*>(C.u);
  const T* __restrict__ vv = static_cast<const T*>(L.v);
  const T* __restrict__ b = static_cast<const T*>(L.b);

  for (int k = threadIdx.x; k < CRH * CRW; k += kThreads) {
    const int a = k / CRW, c = k - a * CRW;
    const int I = cr0 + a, J = cc0 + c;
    se[k] = (I >= 0 && I < C.m && J >= 0 && J < C.n) ? e[(size_t)I * C.pitch + J] : T(0);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < RH * RW; k += kThreads) {
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    T uv = T(0), bv = T(0);
    if (gi >= 0 && gi < m && gj >= 0 && gj < n) {
      const Taps tr = prolong_taps(neu, gi, C.m);
      const Taps tc = prolong_taps(neu, gj, C.n);
      double acc = 0.0;
      for (int y = 0; y < tc.count; ++y) {
        double col = 0.0;
        for (int x = 0; x < tr.count; ++x)
          col += tr.w[x] * double(se[(tr.idx[x] - cr0) * CRW + (tc.idx[y] - cc0)]);
        acc += tc.w[y] * col;
      }
      uv = vv[(size_t)gi * pitch + gj] + T(acc);
      bv = b[(size_t)gi * pitch + gj];
    }
    su[k] = uv;
    sb[k] = bv;
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

  T* __restrict__ uo = static_cast<T*>(L.u);
  double zr = 0.0, zq = 0.0;
  const KrylovDesc* K = DOTS ? &kds[p] : nullptr;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const T z = su[(a + H) * RW + c + H];
    uo[(size_t)gi * pitch + gj] = z;
    if (DOTS) {
      const size_t o = (size_t)gi * pitch + gj;
      zr += double(z) * K->r[o];
      zq += double(z) * K->q[o];
    }
  }
  if (DOTS) {
    const double br = block_sum<kThreads>(zr, sh);
    const double bq = block_sum<kThreads>(zq, sh + kThreads / 64);
    if (threadIdx.x == 0) partial[t.slot] = make_double2(br, bq);
    if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
      const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
      if (threadIdx.x == 0) {
        ProblemState& S = st[p];
        double beta;
        if (S.iter == 0 || S.restart) beta = 0.0;
        else if (S.flexible) beta = -S.alpha * tot.y / S.delta;
        else beta = tot.x / S.delta;
        S.beta = beta;
        S.delta = tot.x;
        S.restart = 0;
        counters[p] = 0;
      }
    }
  }
}

// ------------------------------------------------------------------ coarse tail
// One block per problem runs every level from TailDesc.first to the coarsest in LDS:
// pre-smooth / residual / restrict down, FP64 coarsest solve, prolong / post-smooth up.
template <typename T>
__device__ void lds_halfsweep(T* u, const T* b, int m, int n, T h2, bool neu, int parity) {
  const int half = (n + 1) >> 1;
  for (int k = threadIdx.x; k < m * half; k += kTailThreads) {
    const int i = k / half;
    const int j = 2 * (k - i * half) + ((i + parity) & 1);
    if (j >= n) continue;
    const int o = i * n + j;
    const T up = i > 0 ? u[o - n] : T(0);
    const T dn = i < m - 1 ? u[o + n] : T(0);
    const T lf = j > 0 ? u[o - 1] : T(0);
    const T rt = j < n - 1 ? u[o + 1] : T(0);
    u[o] = smooth_value<T>(((up + dn) + lf) + rt, b[o], h2, neu, i, j, m, n);
  }
  __syncthreads();
}

template <typename T>
__device__ void lds_residual(const T* u, const T* b, T* r, int m, int n, T ih2, bool neu) {
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int i = o / n, j = o - i * n;
    const T up = i > 0 ? u[o - n] : T(0);
    const T dn = i < m - 1 ? u[o + n] : T(0);
    const T lf = j > 0 ? u[o - 1] : T(0);
    const T rt = j < n - 1 ? u[o + 1] : T(0);
    r[o] = b[o] - ((((up + dn) + lf) + rt) - diag_of<T>(neu, i, j, m, n) * u[o]) * ih2;
  }
  __syncthreads();
}

template <typename T>
__device__ void lds_restrict(const T* r, int m, int n, T* bc, int mc, int nc, bool neu) {
  const int ntap = neu ? 4 : 3, tap0 = neu ? -1 : 0;
  for (int o = threadIdx.x; o < mc * nc; o += kTailThreads) {
    const int I = o / nc, J = o - I * nc;
    double acc = 0.0;
    for (int cj = 0; cj < ntap; ++cj) {
      const double wc = restrict_weight(neu, J, cj, n, nc);
      if (wc == 0.0) continue;
      const int j = 2 * J + tap0 + cj;
      double col = 0.0;
      for (int ri = 0; ri < ntap; ++ri) {
        const double wr = restrict_weight(neu, I, ri, m, mc);
        if (wr == 0.0) continue;
        col += wr * double(r[(2 * I + tap0 + ri) * n + j]);
      }
      acc += wc * col;
    }
    bc[o] = T(0.25 * acc);
  }
  __syncthreads();
}

template <typename T>
__device__ void lds_prolong_add(T* u, int m, int n, const T* e, int mc, int nc, bool neu) {
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int i = o / n, j = o - i * n;
    const Taps tr = prolong_taps(neu, i, mc);
    const Taps tc = prolong_taps(neu, j, nc);
    double acc = 0.0;
    for (int y = 0; y < tc.count; ++y) {
      double col = 0.0;
      for (int x = 0; x < tr.count; ++x) col += tr.w[x] * double(e[tr.idx[x] * nc + tc.idx[y]]);
      acc += tc.w[y] * col;
    }
    u[o] = u[o] + T(acc);
  }
  __syncthreads();
}

// Exact zero-mean solve of the Neumann coarsest level by the orthonormal DCT-II
// (the solvers.py:224-262 construction): U = Cm^T ((Cm B Cn^T) ./ lambda) Cn.
__device__ void lds_dct_solve(double* w0, double* w1, int m, int n, const double* tab) {
  const double* Cm = tab;
  const double* Cn = tab + m * m;
  const double* il = tab + m * m + n * n;
  // w1 = Cm * w0
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int k = o / n, j = o - k * n;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += Cm[k * m + i] * w0[i * n + j];
    w1[o] = s;
  }
  __syncthreads();
  // w0 = (w1 * Cn^T) ./ lambda
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int k = o / n, l = o - k * n;
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += w1[k * n + j] * Cn[l * n + j];
    w0[o] = s * il[o];
  }
  __syncthreads();
  // w1 = Cm^T * w0
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int i = o / n, l = o - i * n;
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += Cm[k * m + i] * w0[k * n + l];
    w1[o] = s;
  }
  __syncthreads();
  // w0 = w1 * Cn
  for (int o = threadIdx.x; o < m * n; o += kTailThreads) {
    const int i = o / n, j = o - i * n;
    double s = 0.0;
    for (int l = 0; l < n; ++l) s += w1[i * n + l] * Cn[l * n + j];
    w0[o] = s;
  }
  __syncthreads();
}

template <typename T, bool DOTS>
__global__ __launch_bounds__(kTailThreads) void k_tail(const TailDesc* __restrict__ tails,
                                                       const LevelDesc* __restrict__ levels,
                                                       ProblemState* __restrict__ st,
                                                       const KrylovDesc* __restrict__ kds,
                                                       int pre, int post, int coarsest_sweeps,
                                                       int order) {
  extern __shared__ __align__(16) unsigned char smem[];
  const TailDesc& D = tails[blockIdx.x];
  const int p = D.problem;
  if (st[p].done) return;
  __shared__ double sh[2 * kTailThreads / 64];

  // load the right-hand side of the first tail level
  {
    const LevelDesc L = levels[D.first];
    T* b = reinterpret_cast<T*>(smem + D.lds_off[0][1]);
    const T* g = static_cast<const T*>(L.b);
    for (int o = threadIdx.x; o < L.m * L.n; o += kTailThreads) {
      const int i = o / L.n, j = o - i * L.n;
      b[o] = g[(size_t)i * L.pitch + j];
    }
    __syncthreads();
  }
  T* scratch = reinterpret_cast<T*>(smem + D.scratch_off);
  for (int l = 0; l + 1 < D.nlev; ++l) {
    const LevelDesc L = levels[D.first + l];
    const LevelDesc C = levels[D.first + l + 1];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    T* b = reinterpret_cast<T*>(smem + D.lds_off[l][1]);
    T* bc = reinterpret_cast<T*>(smem + D.lds_off[l + 1][1]);
    const bool neu = L.bc == kNeumann;
    for (int o = threadIdx.x; o < L.m * L.n; o += kTailThreads) u[o] = T(0);
    __syncthreads();
    for (int s = 0; s < pre; ++s) {
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, 0);
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, 1);
    }
    lds_residual<T>(u, b, scratch, L.m, L.n, T(L.inv_h2), neu);
    lds_restrict<T>(scratch, L.m, L.n, bc, C.m, C.n, neu);
  }
  // coarsest level, FP64
  {
    const int l = D.nlev - 1;
    const LevelDesc L = levels[D.first + l];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    const T* b = reinterpret_cast<const T*>(smem + D.lds_off[l][1]);
    double* w0 = reinterpret_cast<double*>(smem + D.coarse_off);
    double* w1 = w0 + L.m * L.n;
    const int N = L.m * L.n;
    if (L.bc == kNeumann) {
      for (int o = threadIdx.x; o < N; o += kTailThreads) w0[o] = double(b[o]) * L.h2;
      __syncthreads();
      lds_dct_solve(w0, w1, L.m, L.n, D.dct);
    } else {
      for (int o = threadIdx.x; o < N; o += kTailThreads) {
        w0[o] = 0.0;
        w1[o] = double(b[o]);
      }
      __syncthreads();
      for (int s = 0; s < coarsest_sweeps; ++s) {
        lds_halfsweep<double>(w0, w1, L.m, L.n, L.h2, false, 0);
        lds_halfsweep<double>(w0, w1, L.m, L.n, L.h2, false, 1);
      }
    }
    for (int o = threadIdx.x; o < N; o += kTailThreads) u[o] = T(w0[o]);
    __syncthreads();
  }
  for (int l = D.nlev - 2; l >= 0; --l) {
    const LevelDesc L = levels[D.first + l];
    const LevelDesc C = levels[D.first + l + 1];
    T* u = reinterpret_cast<T*>(smem + D.lds_off[l][0]);
    const T* b = reinterpret_cast<const T*>(smem + D.lds_off[l][1]);
    const T* e = reinterpret_cast<const T*>(smem + D.lds_off[l + 1][0]);
    const bool neu = L.bc == kNeumann;
    lds_prolong_add<T>(u, L.m, L.n, e, C.m, C.n, neu);
    for (int s = 0; s < post; ++s) {
      const int p0 = order == 0 ? 0 : 1;
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, p0);
      lds_halfsweep<T>(u, b, L.m, L.n, T(L.h2), neu, 1 - p0);
    }
  }
  // write the first tail level's iterate
  const LevelDesc L = levels[D.first];
  const T* u = reinterpret_cast<const T*>(smem + D.lds_off[0][0]);
  T* g = static_cast<T*>(L.u);
  double zr = 0.0, zq = 0.0;
  const bool dots = DOTS && D.krylov >= 0;
  for (int o = threadIdx.x; o < L.m * L.n; o += kTailThreads) {
    const int i = o / L.n, j = o - i * L.n;
    const size_t go = (size_t)i * L.pitch + j;
    g[go] = u[o];
    if (dots) {
      const KrylovDesc& K = kds[D.krylov];
      zr += double(u[o]) * K.r[go];
      zq += double(u[o]) * K.q[go];
    }
  }
  if (dots) {
    const double br = block_sum<kTailThreads>(zr, sh);
    const double bq = block_sum<kTailThreads>(zq, sh + kTailThreads / 64);
    if (threadIdx.x == 0) {
      ProblemState& S = st[p];
      double beta;
      if (S.iter == 0 || S.restart) beta = 0.0;
      else if (S.flexible) beta = -S.alpha * bq / S.delta;
      else beta = br / S.delta;
      S.beta = beta;
      S.delta = br;
      S.restart = 0;
    }
  }
}

// ------------------------------------------------------------------ Krylov kernels
// Tiles of TH x TW fine points; KrylovDesc and level-0 LevelDesc share the pitch.
template <int TH, int TW>
struct KTile {
  static constexpr int N = TH * TW;
};

// d_new = z + beta d_old ; q = -A d_new ; <d_new, q>  -> alpha = delta / <d, q>
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_apply(const Tile* __restrict__ tiles,
                                                    const KrylovDesc* __restrict__ kds,
                                                    const LevelDesc* __restrict__ levels,
                                                    ProblemState* __restrict__ st,
                                                    double2* __restrict__ partial,
                                                    int* __restrict__ counters,
                                                    const SlotRange* __restrict__ ranges,
                                                    int din) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  if (st[p].done) return;
  constexpr int RH = TH + 2, RW = TW + 2;
  __shared__ double sd[RH * RW];
  __shared__ double sh[kThreads / 64];
  __shared__ int s_flag;
  const double beta = st[p].beta;
  const int m = K.m, n = K.n, pitch = K.pitch;
  const T* __restrict__ z = static_cast<const T*>(levels[K.level0].u);
  const double* __restrict__ dold = K.d[din];
  double* __restrict__ dnew = K.d[din ^ 1];
  const int gr0 = t.i0 - 1, gc0 = t.j0 - 1;
  for (int k = threadIdx.x; k < RH * RW; k += kThreads) {
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    double dv = 0.0;
    if (gi >= 0 && gi < m && gj >= 0 && gj < n) {
      const size_t o = (size_t)gi * pitch + gj;
      dv = double(z[o]) + beta * dold[o];
    }
    sd[k] = dv;
  }
  __syncthreads();
  const bool neu = K.bc == kNeumann;
  double acc = 0.0;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const int idx = (a + 1) * RW + c + 1;
    const double dc = sd[idx];
    const double nb = ((sd[idx - RW] + sd[idx + RW]) + sd[idx - 1]) + sd[idx + 1];
    const double q = -(nb - diag_of<double>(neu, gi, gj, m, n) * dc) * K.inv_h2;
    const size_t o = (size_t)gi * pitch + gj;
    dnew[o] = dc;
    K.q[o] = q;
    acc += dc * q;
  }
  const double bs = block_sum<kThreads>(acc, sh);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(bs, 0.0);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      st[p].dq = tot.x;
      st[p].alpha = st[p].delta / tot.x;
      counters[p] = 0;
    }
  }
}

// x += alpha d ; r -= alpha q ; V-cycle RHS = -r ; ||r||^2 -> convergence
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_update(const Tile* __restrict__ tiles,
                                                     const KrylovDesc* __restrict__ kds,
                                                     const LevelDesc* __restrict__ levels,
                                                     ProblemState* __restrict__ st,
                                                     double2* __restrict__ partial,
                                                     int* __restrict__ counters,
                                                     const SlotRange* __restrict__ ranges,
                                                     int dcur) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  if (st[p].done) return;
  __shared__ double sh[kThreads / 64];
  __shared__ int s_flag;
  const double alpha = st[p].alpha;
  const int m = K.m, n = K.n, pitch = K.pitch;
  T* __restrict__ b0 = static_cast<T*>(levels[K.level0].b);
  const double* __restrict__ d = K.d[dcur];
  double acc = 0.0;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const size_t o = (size_t)gi * pitch + gj;
    K.x[o] += alpha * d[o];
    const double r = K.r[o] - alpha * K.q[o];
    K.r[o] = r;
    b0[o] = T(-r);
    acc += r * r;
  }
  const double bs = block_sum<kThreads>(acc, sh);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(bs, 0.0);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      ProblemState& S = st[p];
      S.rr = tot.x;
      S.iter += 1;
      S.converged = tot.x <= S.target2;
      S.done = S.converged || S.iter >= S.max_iter;
      counters[p] = 0;
    }
  }
}

// Jacobi preconditioner of the reference "cg" method (solvers.py:114-117):
// z = r / (4 / h^2), delta' = <r, z>, beta = delta' / delta.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_jacobi(const Tile* __restrict__ tiles,
                                                     const KrylovDesc* __restrict__ kds,
                                                     const LevelDesc* __restrict__ levels,
                                                     ProblemState* __restrict__ st,
                                                     double2* __restrict__ partial,
                                                     int* __restrict__ counters,
                                                     const SlotRange* __restrict__ ranges) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  if (st[p].done) return;
  __shared__ double sh[kThreads / 64];
  __shared__ int s_flag;
  const int m = K.m, n = K.n, pitch = K.pitch;
  T* __restrict__ z = static_cast<T*>(levels[K.level0].u);
  const double diag = 4.0 * K.inv_h2;
  double acc = 0.0;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const size_t o = (size_t)gi * pitch + gj;
    const double r = K.r[o];
    const double zz = r / diag;
    z[o] = T(zz);
    acc += r * double(T(zz));
  }
  const double bs = block_sum<kThreads>(acc, sh);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(bs, 0.0);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      ProblemState& S = st[p];
      S.beta = (S.iter == 0 || S.restart) ? 0.0 : tot.x / S.delta;
      S.delta = tot.x;
      S.restart = 0;
      counters[p] = 0;
    }
  }
}

// Pass 1 of a solve's set-up: mean of b for the Neumann projection onto the
// compatible (zero-mean) right-hand sides (the constant that solvers.py:224-262 drops).
template <int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_init_stats(const Tile* __restrict__ tiles,
                                                         const KrylovDesc* __restrict__ kds,
                                                         const IoDesc* __restrict__ io,
                                                         ProblemState* __restrict__ st,
                                                         double2* __restrict__ partial,
                                                         int* __restrict__ counters,
                                                         const SlotRange* __restrict__ ranges,
                                                         int project) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  __shared__ double sh[kThreads / 64];
  __shared__ int s_flag;
  const double* __restrict__ b = io[p].b;
  double acc = 0.0;
  if (project && K.bc == kNeumann) {
    for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
      const int a = k / TW, c = k - a * TW;
      const int gi = t.i0 + a, gj = t.j0 + c;
      if (gi >= K.m || gj >= K.n) continue;
      acc += b[(size_t)gi * K.n + gj];
    }
  }
  const double bs = block_sum<kThreads>(acc, sh);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(bs, 0.0);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      st[p].bmean = (project && K.bc == kNeumann) ? tot.x / (double(K.m) * double(K.n)) : 0.0;
      counters[p] = 0;
    }
  }
}

// Pass 2 of a solve's set-up.  mode 0 (PCG): r = -(b - mean), x = d = q = 0, V-cycle
// RHS = b - mean.  mode 1 (MG iteration): x = initial guess (or 0).  Sets ||b||^2 and
// the per-problem state.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_init(const Tile* __restrict__ tiles,
                                                   const KrylovDesc* __restrict__ kds,
                                                   const LevelDesc* __restrict__ levels,
                                                   const IoDesc* __restrict__ io,
                                                   const double* const* __restrict__ initial,
                                                   ProblemState* __restrict__ st,
                                                   double2* __restrict__ partial,
                                                   int* __restrict__ counters,
                                                   const SlotRange* __restrict__ ranges,
                                                   int mode, double tol2, int max_iter,
                                                   int flexible, int zero_done) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  __shared__ double sh[kThreads / 64];
  __shared__ int s_flag;
  const double mean = st[p].bmean;
  const double* __restrict__ b = io[p].b;
  const double* __restrict__ x0 = initial ? initial[p] : nullptr;
  T* __restrict__ b0 = static_cast<T*>(levels[K.level0].b);
  double acc = 0.0;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= K.m || gj >= K.n) continue;
    const size_t o = (size_t)gi * K.pitch + gj;
    const double bb = b[(size_t)gi * K.n + gj] - mean;
    acc += bb * bb;
    if (mode == 0) {
      K.r[o] = -bb;
      K.x[o] = 0.0;
      K.d[0][o] = 0.0;
      K.d[1][o] = 0.0;
      K.q[o] = 0.0;
      b0[o] = T(bb);
    } else {
      K.x[o] = x0 ? x0[(size_t)gi * K.n + gj] : 0.0;
    }
  }
  const double bs = block_sum<kThreads>(acc, sh);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(bs, 0.0);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      ProblemState& S = st[p];
      S.bnorm2 = tot.x;
      S.target2 = tol2 * tot.x;
      S.iter = 0;
      S.converged = zero_done && tot.x == 0.0;
      S.done = S.converged || max_iter <= 0;
      S.restart = 0;
      S.alpha = 0.0;
      S.beta = 0.0;
      S.delta = 0.0;
      S.rr = tot.x;
      S.max_iter = max_iter;
      S.flexible = flexible;
      counters[p] = 0;
    }
  }
}

// r_true = (b - mean) - L x (FP64), ||r_true||^2 and sum(x); the V-cycle RHS becomes
// r_true.  mode 0 (PCG): also store r = -r_true so an unconverged problem can restart.
// mode 1 (MG, before the first cycle): record history[0].  mode 2 (MG, after a cycle,
// multigrid.py:250-254): count the cycle, record it, test rel < tolerance.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_residual(const Tile* __restrict__ tiles,
                                                       const KrylovDesc* __restrict__ kds,
                                                       const LevelDesc* __restrict__ levels,
                                                       const IoDesc* __restrict__ io,
                                                       ProblemState* __restrict__ st,
                                                       double2* __restrict__ partial,
                                                       int* __restrict__ counters,
                                                       const SlotRange* __restrict__ ranges,
                                                       int mode, double tol, double* history,
                                                       int hist_stride, int respect_done) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  if (respect_done && st[p].done) return;
  constexpr int RH = TH + 2, RW = TW + 2;
  __shared__ double sx[RH * RW];
  __shared__ double sh[2 * kThreads / 64];
  __shared__ int s_flag;
  const int m = K.m, n = K.n, pitch = K.pitch;
  const double mean = st[p].bmean;
  const double* __restrict__ b = io[p].b;
  T* __restrict__ b0 = static_cast<T*>(levels[K.level0].b);
  const int gr0 = t.i0 - 1, gc0 = t.j0 - 1;
  for (int k = threadIdx.x; k < RH * RW; k += kThreads) {
    const int a = k / RW, c = k - a * RW;
    const int gi = gr0 + a, gj = gc0 + c;
    sx[k] = (gi >= 0 && gi < m && gj >= 0 && gj < n) ? K.x[(size_t)gi * pitch + gj] : 0.0;
  }
  __syncthreads();
  const bool neu = K.bc == kNeumann;
  double rr = 0.0, sum = 0.0;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= m || gj >= n) continue;
    const int idx = (a + 1) * RW + c + 1;
    const double xc = sx[idx];
    const double nb = ((sx[idx - RW] + sx[idx + RW]) + sx[idx - 1]) + sx[idx + 1];
    const double lx = (nb - diag_of<double>(neu, gi, gj, m, n) * xc) * K.inv_h2;
    const double res = (b[(size_t)gi * n + gj] - mean) - lx;
    const size_t o = (size_t)gi * pitch + gj;
    if (mode == 0) K.r[o] = -res;
    b0[o] = T(res);
    rr += res * res;
    sum += xc;
  }
  const double br = block_sum<kThreads>(rr, sh);
  const double bx = block_sum<kThreads>(sum, sh + kThreads / 64);
  if (threadIdx.x == 0) partial[t.slot] = make_double2(br, bx);
  if (arrive_last(&counters[p], ranges[p].count, &s_flag)) {
    const double2 tot = gather_partials<kThreads>(partial, ranges[p], sh);
    if (threadIdx.x == 0) {
      ProblemState& S = st[p];
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
      counters[p] = 0;
    }
  }
}

// MG iteration: x += z (the V-cycle correction), one more cycle done.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_correct(const Tile* __restrict__ tiles,
                                                      const KrylovDesc* __restrict__ kds,
                                                      const LevelDesc* __restrict__ levels,
                                                      ProblemState* __restrict__ st) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  if (st[p].done) return;
  const T* __restrict__ z = static_cast<const T*>(levels[K.level0].u);
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= K.m || gj >= K.n) continue;
    const size_t o = (size_t)gi * K.pitch + gj;
    K.x[o] += double(z[o]);
  }
}

}  // namespace gdp

namespace gdp {

// Caller output: u = x - mean(x) for Neumann problems when mean_fix, else u = x.
template <int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_finalize(const Tile* __restrict__ tiles,
                                                       const KrylovDesc* __restrict__ kds,
                                                       const IoDesc* __restrict__ io,
                                                       const ProblemState* __restrict__ st,
                                                       int mean_fix) {
  const Tile t = tiles[blockIdx.x];
  const KrylovDesc K = kds[t.desc];
  const int p = t.desc;
  const double mean = (mean_fix && K.bc == kNeumann)
                          ? st[p].sumx / (double(K.m) * double(K.n))
                          : 0.0;
  double* __restrict__ u = io[p].u;
  for (int k = threadIdx.x; k < TH * TW; k += kThreads) {
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= K.m || gj >= K.n) continue;
    u[(size_t)gi * K.n + gj] = K.x[(size_t)gi * K.pitch + gj] - mean;
  }
}

// Restart bookkeeping after the FP64 true-residual check: a problem whose recursive
// residual converged but whose true residual did not is reopened with a fresh
// Krylov sequence (residual replacement), if it has iterations left.
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
