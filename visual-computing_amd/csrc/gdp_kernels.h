// MI355X (gfx950) kernels of the multigrid-preconditioned CG Poisson solver: the
// coarse tail (every level at or below 64 x 64 of one problem in one block, in LDS) and
// the FP64 Krylov kernels.  The smoothing kernels of the larger levels are in
// gdp_strip.h.
//
// Every kernel is batched: blockIdx.x indexes a host-built table whose records carry
// everything a block needs (pointers, sizes, tile origin, partial-sum slot), so one
// launch covers every problem of a plan (RGB channels, a batch of images).  Per-problem
// scalars (alpha, beta, delta, ||r||) are reduced on the device without atomics
// (gdp_reduce.h): a problem's result does not depend on what else is in the batch.
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
#include "gdp_dct.h"
#include "gdp_reduce.h"

namespace gdp {

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

// ------------------------------------------------------------------ coarse tail
// One block per problem runs every level from TailDesc.first to the coarsest in LDS:
// pre-smooth / residual / restrict down, FP64 coarsest solve, prolong / post-smooth up.
// Loops run rows over waves and columns over lanes (no integer division).

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
  // the convergence test of the update: here for a problem whose finest level is in the
  // tail (check_conv 1, k_update's partials), and for every problem after the fused
  // update-down kernel (check_conv 2, that kernel's partials)
  if ((check_conv == 2 || (check_conv == 1 && D.krylov >= 0)) &&
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
//   k_apply_s (gdp_apply.h) reads the V-cycle's <r,z>, <q,z> partials -> delta, beta;
//             d = z + beta d, q = -A d, partial <d, q>
//   k_update  reads the <d, q> partials -> alpha = delta / <d, q>;
//             x += alpha d, r -= alpha q, V-cycle RHS -r, partial ||r||^2
//   first V-cycle kernel: convergence test from the ||r||^2 partials
// With defer_x the update leaves x += alpha d out (and says so in xpend) for the extra
// blocks of the same iteration's tail launch (st_xupdate in gdp_stail.h), which run on
// the CUs that the tail's one block per problem leaves idle, or, in a plan whose tails
// leave too few CUs idle, for the next iteration's apply (k_apply_s<..., XUPD>), which
// reads that direction anyway; d stays valid until the next iteration's apply writes
// the other direction buffer.  That apply adds it to x every other iteration only and
// holds it back otherwise (xhold, xalpha): the update after it then reads the held
// direction from the buffer it does not write, and adds both to x itself or, deferred
// again, leaves both to the next apply (xdouble).
// Each kernel reads only state fields written by earlier kernels and its lead block
// writes only fields that no block of the same kernel reads (the apply's lead writes
// xalpha only for a problem whose blocks do not read it in that launch).

template <typename T, int TH, int TW, bool DEFER_X>
__global__ __launch_bounds__(kThreads) void k_update(const KTile* __restrict__ tiles,
                                                     ProblemState* __restrict__ st,
                                                     const double2* __restrict__ apart,
                                                     double2* __restrict__ upart, int dcur) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  const int m = t.m, n = t.n, pitch = t.pitch;
  GDP_GLOBAL T* b0 = (GDP_GLOBAL T*)t.b0;
  GDP_GLOBAL double* x = (GDP_GLOBAL double*)t.x;
  GDP_GLOBAL double* r = (GDP_GLOBAL double*)t.r;
  const GDP_GLOBAL double* q = (const GDP_GLOBAL double*)t.q;
  const GDP_GLOBAL double* d = (const GDP_GLOBAL double*)(dcur ? t.d1 : t.d0);
  // an x update the apply held back: xalpha times the other direction buffer, added
  // before this iteration's (k_apply_s in gdp_apply.h)
  const GDP_GLOBAL double* dh = (const GDP_GLOBAL double*)(dcur ? t.d0 : t.d1);
  const bool held = !DEFER_X && st[p].xhold;
  constexpr int NO = (TH * TW) / kThreads;
  double xl[NO], dl[NO], rl[NO], ql[NO], hl[NO];
  size_t ol[NO];
  bool il[NO];
  // branch-free loads (clamped in-bounds addresses; the points outside are not used), so
  // every load is issued before the first wait
#pragma unroll
  for (int it = 0; it < NO; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    il[it] = gi < m && gj < n;
    ol[it] = (size_t)min(gi, m - 1) * pitch + min(gj, n - 1);
    if constexpr (!DEFER_X) {
      xl[it] = x[ol[it]];
      dl[it] = d[ol[it]];
    }
    rl[it] = r[ol[it]];
    ql[it] = q[ol[it]];
  }
  if (held) {
#pragma unroll
    for (int it = 0; it < NO; ++it) hl[it] = dh[ol[it]];
  }
  // the finished-problem test comes after the reduction's barrier, which keeps the
  // loads above in front of it (a test before it lets the compiler sink them below the
  // state load)
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
    st[p].xpend = DEFER_X ? dcur + 1 : 0;
    // a deferred update after an apply that held one back leaves both to the next apply
    st[p].xdouble = DEFER_X ? st[p].xhold : 0;
  }
  const double ah = held ? st[p].xalpha : 0.0;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < NO; ++it) {
    if (!il[it]) continue;
    // one rounding, as the fused update-down and finest up kernels form it
    const double rn = __builtin_fma(-alpha, ql[it], rl[it]);
    if constexpr (!DEFER_X)
      x[ol[it]] = __builtin_fma(alpha, dl[it], held ? __builtin_fma(ah, hl[it], xl[it]) : xl[it]);
    r[ol[it]] = rn;
    b0[ol[it]] = T(-rn);
    acc += rn * rn;
  }
  write_partial<kThreads>(acc, 0.0, upart, t.slot, sh);
}


// Jacobi preconditioner of the reference "cg" method (solvers.py:114-117):
// z = r / (4 / h^2) and its <r, z> partials (consumed by k_apply_s with beta = delta'/delta).
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
  // branch-free loads (clamped in-bounds addresses, the points outside not used), all
  // issued before the first wait; summed in the same order as one at a time
  constexpr int NI = (TH * TW) / kThreads;
  double v[NI];
  bool in[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    in[it] = gi < t.m && gj < t.n;
    v[it] = b[(size_t)min(gi, t.m - 1) * t.n + min(gj, t.n - 1)];
  }
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < NI; ++it) acc += in[it] ? v[it] : 0.0;
  write_partial<kThreads>(acc, 0.0, spart, t.slot, sh);
}

// Set-up pass 2.  mode 0 (PCG): r = -(b - mean), x = d = q = 0, V-cycle RHS b - mean.
// mode 1 (MG iteration): x = initial guess (or 0).  Partial sums of (b - mean)^2.
// The mean (project: of a Neumann problem, else 0) comes from k_init_stats' partial sums
// (bpart, slot ranges rk), reduced by every block of the problem; the lead records it.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_init(const KTile* __restrict__ tiles,
                                                   const IoDesc* __restrict__ io,
                                                   const double* const* __restrict__ initial,
                                                   ProblemState* __restrict__ st,
                                                   double2* __restrict__ spart, int mode,
                                                   const double2* __restrict__ bpart,
                                                   const SlotRange* __restrict__ rk,
                                                   int project) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  const double* __restrict__ b = io[p].b;
  const double* __restrict__ x0 = initial ? initial[p] : nullptr;
  T* __restrict__ b0 = static_cast<T*>(t.b0);
  // branch-free loads of b (clamped in-bounds addresses), all issued before the first
  // wait, the reduction of the mean's partial sums
  constexpr int NI = (TH * TW) / kThreads;
  double bv[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    bv[it] = b[(size_t)min(t.i0 + a, t.m - 1) * t.n + min(t.j0 + c, t.n - 1)];
  }
  const double2 bs = reduce_slots<kThreads>(bpart, rk[p], sh, &bc);
  const double mean =
      (project && t.bc == kNeumann) ? bs.x / (double(t.m) * double(t.n)) : 0.0;
  if (t.lead && threadIdx.x == 0) st[p].bmean = mean;
  double acc = 0.0;
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    const size_t o = (size_t)gi * t.pitch + gj;
    const double bb = bv[it] - mean;
    acc += bb * bb;
    if (mode == 0) {
      // d and q need no clearing: the first direction is z itself (beta = 0) and q is
      // written by the operator apply before any kernel reads it
      t.r[o] = -bb;
      t.x[o] = 0.0;
      b0[o] = T(bb);
    } else {
      t.x[o] = x0 ? x0[(size_t)gi * t.n + gj] : 0.0;
    }
  }
  write_partial<kThreads>(acc, 0.0, spart, t.slot, sh);
}

// Which problems a set-up / check kernel processes: respect 0 every problem, 1 those
// not done, 2 those done whose true residual has not been checked yet.
__device__ inline bool skip_problem(const ProblemState& S, int respect) {
  return respect == 1 ? S.done != 0 : (respect == 2 ? (!S.done || S.checked) : false);
}

// r_true = (b - mean) - L x (FP64): partial sums of ||r_true||^2 and of x; the V-cycle
// RHS becomes r_true and, in mode 0 (PCG), r = -r_true so an unconverged problem can
// restart from it.
// check_conv (with respect_done 2, after an update whose preconditioner is not run): the
// update's convergence test first, in every block of a problem that was open, from the
// ||r||^2 partials (slot ranges rk); the lead block records it, and the problems it
// finishes get their true residual in the same launch.  Whether a block tests is decided
// once for the block from thread 0's read of the state: another block's lead may set
// done meanwhile, to the value this block's own test gives.
template <typename T, int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_residual(const KTile* __restrict__ tiles,
                                                       const IoDesc* __restrict__ io,
                                                       ProblemState* __restrict__ st,
                                                       double2* __restrict__ spart, int mode,
                                                       int respect_done,
                                                       const double2* __restrict__ upart,
                                                       const SlotRange* __restrict__ rk,
                                                       int check_conv) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  constexpr int RH = TH + 2, RW = TW + 2;
  __shared__ double sx[RH * RW];
  __shared__ double sh[2 * kThreads / 64];
  if (check_conv) {
    __shared__ double2 bc;
    __shared__ int flags;
    if (threadIdx.x == 0) flags = (st[p].done ? 1 : 0) | (st[p].checked ? 2 : 0);
    __syncthreads();
    const int f = flags;
    if (f & 2) return;
    if (!(f & 1) && !check_converged<kThreads>(st, p, upart, rk[p], t.lead, sh, &bc)) return;
  } else if (skip_problem(st[p], respect_done)) {
    return;
  }
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
    double res = (b[(size_t)gi * n + gj] - mean) - lx;
    if (t.mask && !t.mask[(size_t)gi * n + gj]) res = 0.0;  // fixed point of a masked domain
    const size_t o = (size_t)gi * pitch + gj;
    if (mode == 0) t.r[o] = -res;
    b0[o] = T(res);
    rr += res * res;
    sum += xc;
  }
  write_partial<kThreads>(rr, sum, spart, t.slot, sh);
}

// One block per problem: reduce a set-up / residual pass and update the state.
//   kind 1 (after k_init): ||b||^2, target, fresh iteration state
//   kind 2 (after k_residual): true ||r||^2, sum(x); MG modes also record the history
//          and test rel < tolerance (multigrid.py:250-254); the check of a PCG batch
//          (reopen 1) also takes the restart decision: a problem whose recursive
//          residual converged but whose true residual did not is reopened with a fresh
//          Krylov sequence (residual replacement) if it has iterations and restarts
//          (kMaxRestarts per problem, so a problem's course does not depend on its batch)
//          left; every other one is marked checked
__global__ __launch_bounds__(kThreads) void k_scalar(const KrylovDesc* __restrict__ kds,
                                                     ProblemState* __restrict__ st,
                                                     const double2* __restrict__ spart,
                                                     const SlotRange* __restrict__ rk, int kind,
                                                     int mode, double tol,
                                                     int max_iter, int flexible, int zero_done,
                                                     int respect_done, double* history,
                                                     int hist_stride, int reopen) {
  const int p = blockIdx.x;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  if (kind == 2 && skip_problem(st[p], respect_done)) return;
  const double2 tot = reduce_slots<kThreads>(spart, rk[p], sh, &bc);
  if (threadIdx.x != 0) return;
  ProblemState& S = st[p];
  const KrylovDesc& K = kds[p];
  if (kind == 1) {
    S.bnorm2 = tot.x;
    S.target2 = tol * tol * tot.x;
    S.iter = 0;
    S.converged = zero_done && tot.x == 0.0;
    S.done = S.converged || max_iter <= 0;
    S.restart = 0;
    S.checked = 0;
    S.xpend = S.xhold = S.xdouble = S.restarts = 0;
    S.alpha = S.beta = S.delta = S.delta_old = S.xalpha = 0.0;
    S.rr = tot.x;
    S.max_iter = max_iter;
    S.flexible = flexible;
  } else {
    S.true_rr = tot.x;
    S.sumx = tot.y;
    if (reopen) {
      // a done, unchecked problem (respect_done 2)
      const bool bad = tot.x > S.target2 && S.bnorm2 > 0.0;
      if (bad && S.iter < S.max_iter && S.restarts < kMaxRestarts) {
        S.converged = 0;
        S.done = 0;
        S.restart = 1;
        S.restarts += 1;
      } else {
        if (bad) S.converged = 0;
        S.checked = 1;
      }
    }
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

// Caller output: u = x - mean(x) for Neumann problems when mean_fix, else u = x.  With
// only_checked, only problems whose true residual has been checked are written.
template <int TH, int TW>
__global__ __launch_bounds__(kThreads) void k_finalize(const KTile* __restrict__ tiles,
                                                       const IoDesc* __restrict__ io,
                                                       const ProblemState* __restrict__ st,
                                                       int mean_fix, int only_checked) {
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  // x first (branch-free, clamped in-bounds addresses), the state test after
  constexpr int NI = (TH * TW) / kThreads;
  double xv[NI];
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    xv[it] = t.x[(size_t)min(t.i0 + a, t.m - 1) * t.pitch + min(t.j0 + c, t.n - 1)];
  }
  asm volatile("" ::: "memory");  // keeps the loads above ahead of the state load
  if (only_checked && !st[p].checked) return;
  const double mean = (mean_fix && t.bc == kNeumann)
                          ? st[p].sumx / (double(t.m) * double(t.n))
                          : 0.0;
  double* __restrict__ u = io[p].u;
#pragma unroll
  for (int it = 0; it < NI; ++it) {
    const int k = threadIdx.x + it * kThreads;
    const int a = k / TW, c = k - a * TW;
    const int gi = t.i0 + a, gj = t.j0 + c;
    if (gi >= t.m || gj >= t.n) continue;
    u[(size_t)gi * t.n + gj] = xv[it] - mean;
  }
}

}  // namespace gdp
