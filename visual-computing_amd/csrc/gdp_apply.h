// The PCG direction update and operator apply of one iteration (solvers.py:149-151:
// d = z + beta d, q = A d, <d, q>) in strip form: no LDS and no barrier.
//
// XUPD: the preceding k_update left its x += alpha d out (ProblemState.xpend) because
// the strip-tail launch could not absorb it (a plan with about as many coarse tails as
// CUs).  The apply reads that old direction anyway, and pays x's traffic every other
// iteration only: it holds a single pending update back (xhold; the next update keeps
// that direction, which the apply after it would overwrite, and either adds it itself
// or leaves both to that apply, xdouble), and adds a double one to x of the points it
// owns, first xalpha times the held direction, read from the buffer the apply is about
// to overwrite, then alpha times the old direction: 12 B per point and iteration on
// average, instead of the 24 B of a separate pass, with the bits of one x += alpha d per
// iteration.  A finished problem gets everything pending before the apply leaves it.
//
// A wave owns a strip of kApplyRows output rows and 128 loaded columns: lane l holds the
// column pair (j0 - 1 + 2 l, j0 + 2 l) of each of its rows plus the rows above and below,
// so vertical neighbours are in the lane and horizontal ones come from DPP row shifts.
// Neighbouring strips overlap by two columns: a strip writes the 126 columns j0 .. j0 + 125
// (and column 0 when it is the grid's first strip, j0 = 1); strip origins are odd, so every
// pair load is aligned.  Four strips (stacked vertically) make a block.  The stencil sums
// its neighbours in the order up, down, left, right, as the other stencils here do.
#pragma once
#include "gdp_strip.h"

namespace gdp {

constexpr int kApplyRows = 4;    // output rows per strip
constexpr int kApplyCols = 126;  // output columns per strip
constexpr int kApplyBlockRows = (kThreads / 64) * kApplyRows;  // one strip per wave

template <typename T, bool MASK, bool XUPD = false>
__global__ __launch_bounds__(kThreads) void k_apply_s(const KTile* __restrict__ tiles,
                                                      ProblemState* __restrict__ st,
                                                      const double2* __restrict__ zpart,
                                                      double2* __restrict__ apart, int din,
                                                      int jacobi) {
  using TP = typename Pair<T>::type;
  using D2 = typename Pair<double>::type;
  constexpr int R = kApplyRows, FH = R + 2;
  const KTile t = tiles[blockIdx.x];
  const int p = t.problem;
  __shared__ double sh[2 * kThreads / 64];
  __shared__ double2 bc;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int m = t.m, n = t.n, pitch = t.pitch;
  const int r0 = t.i0 + w * R;
  const int gj = t.j0 - 1 + 2 * lane;  // x column (even); y is gj + 1
  const bool xin = gj < n, yin = gj + 1 < n;
  const GDP_GLOBAL T* z = (const GDP_GLOBAL T*)t.z;
  const GDP_GLOBAL double* dold = (const GDP_GLOBAL double*)(din ? t.d1 : t.d0);
  GDP_GLOBAL double* dnew = (GDP_GLOBAL double*)(din ? t.d0 : t.d1);
  GDP_GLOBAL double* qv = (GDP_GLOBAL double*)t.q;

  // Loads first, branch-free: pairs come from clamped addresses inside the padded rows
  // (the pitch is even) and are zeroed outside the grid.  MASK (masked plans, a kernel of
  // their own): mb holds the lane's unknowns, the in-grid points less the fixed points of
  // a masked domain (all of them for a tile without a mask); the others stay at zero.
  // The scalar reduction below overlaps the loads.
  const int jc = xin ? gj : 0;
  TP zv[FH];
  D2 dv[FH];
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    const size_t o = (size_t)min(max(r0 - 1 + a, 0), m - 1) * pitch + jc;
    zv[a] = *(const GDP_GLOBAL TP*)(z + o);
    dv[a] = *(const GDP_GLOBAL D2*)(dold + o);
  }
  // XUPD: x of the output rows, and for a double update the held direction (in dnew's
  // buffer, read before this thread overwrites it), loaded with the rest when x is due
  GDP_GLOBAL double* xv = (GDP_GLOBAL double*)t.x;
  D2 xo[XUPD ? R : 1], dh[XUPD ? R : 1];
  int xdue = 0;  // 0: x is left alone, 1: x += alpha dold, 2: the held direction first
  if constexpr (XUPD) {
    const int xp = st[p].xpend, xd = st[p].xdouble, fin = st[p].done;
    xdue = !xp ? 0 : xd ? 2 : fin ? 1 : 0;
    if (xdue) {
#pragma unroll
      for (int a = 1; a <= R; ++a) {
        const size_t o = (size_t)min(r0 - 1 + a, m - 1) * pitch + jc;
        xo[a - 1] = *(const GDP_GLOBAL D2*)(xv + o);
        if (xdue == 2) dh[a - 1] = *(const GDP_GLOBAL D2*)(dnew + o);
      }
    }
    // the lead records whether this apply holds a single pending update back; no block
    // of this launch reads xhold, and xalpha only for a double update, which is never
    // held
    if (t.lead && threadIdx.x == 0) {
      const int hold = xp && !xd && !fin;
      st[p].xhold = hold;
      if (hold) st[p].xalpha = st[p].alpha;
    }
  }
  uint64_t mb = 0;
  if constexpr (MASK) mb = load_mask_bits<FH>(t.mask, r0 - 1, gj, m, n, true);
  auto unknown = [&](int a, int y) {
    if constexpr (MASK) return mask_bit(mb, 2 * a + y);
    const int gi = r0 - 1 + a;
    return gi >= 0 && gi < m && (y ? yin : xin);
  };
  // the finished-problem test follows the reduction's barrier, which keeps the loads
  // above in front of it (a test before it lets the compiler sink them below the state
  // load)
  const double2 z2 = reduce_slots<kThreads>(
      zpart, jacobi ? SlotRange{t.kbeg, t.kcnt} : SlotRange{t.zbeg, t.zcnt}, sh, &bc);
  const bool outx = xin && (gj >= t.j0 || gj == 0) && gj < t.j0 + kApplyCols;
  const bool outy = yin && gj + 1 < t.j0 + kApplyCols;
  if constexpr (XUPD) {
    // the pending x += alpha d of the previous iteration (solvers.py:151), d = dold, after
    // the held one of the iteration before it
    if (xdue) {
      const double al = st[p].alpha;
      const double ah = xdue == 2 ? st[p].xalpha : 0.0;
#pragma unroll
      for (int a = 1; a <= R; ++a) {
        const int gi = r0 - 1 + a;
        if (gi >= m) break;
        const size_t o = (size_t)gi * pitch + gj;
        double nx = xo[a - 1].x, ny = xo[a - 1].y;
        if (xdue == 2) {
          nx = __builtin_fma(ah, dh[a - 1].x, nx);
          ny = __builtin_fma(ah, dh[a - 1].y, ny);
        }
        nx = __builtin_fma(al, dv[a].x, nx);
        ny = __builtin_fma(al, dv[a].y, ny);
        if (outx && outy) {
          *(GDP_GLOBAL D2*)(xv + o) = D2{nx, ny};
        } else {
          if (outx) xv[o] = nx;
          if (outy) xv[o + 1] = ny;
        }
      }
    }
  }
  if (st[p].done) return;
  double beta;
  if (st[p].iter == 0 || st[p].restart) beta = 0.0;
  else if (st[p].flexible) beta = -st[p].alpha * z2.y / st[p].delta_old;
  else beta = z2.x / st[p].delta_old;
  if (t.lead && threadIdx.x == 0) {
    st[p].delta = z2.x;
    st[p].beta = beta;
  }
  D2 dn[FH];
#pragma unroll
  for (int a = 0; a < FH; ++a) {
    // a new sequence (beta = 0) takes d = z without reading the old direction, which
    // the set-up leaves unwritten
    dn[a].x = unknown(a, 0) ? (beta == 0.0 ? double(zv[a].x) : double(zv[a].x) + beta * dv[a].x)
                            : 0.0;
    dn[a].y = unknown(a, 1) ? (beta == 0.0 ? double(zv[a].y) : double(zv[a].y) + beta * dv[a].y)
                            : 0.0;
  }
  const bool neu = t.bc == kNeumann;
  double acc = 0.0;
#pragma unroll
  for (int a = 1; a <= R; ++a) {
    const int gi = r0 - 1 + a;
    if (gi >= m) break;
    const double lx = dpp_shr1(dn[a].y), ry = dpp_shl1(dn[a].x);
    const double nbx = ((dn[a - 1].x + dn[a + 1].x) + lx) + dn[a].y;
    const double nby = ((dn[a - 1].y + dn[a + 1].y) + dn[a].x) + ry;
    double qx = -(nbx - diag_of<double>(neu, gi, gj, m, n) * dn[a].x) * t.inv_h2;
    double qy = -(nby - diag_of<double>(neu, gi, gj + 1, m, n) * dn[a].y) * t.inv_h2;
    if (MASK && !mask_bit(mb, 2 * a)) qx = 0.0;  // no equation at a fixed point
    if (MASK && !mask_bit(mb, 2 * a + 1)) qy = 0.0;
    const size_t o = (size_t)gi * pitch + gj;
    if (outx && outy) {
      *(GDP_GLOBAL D2*)(dnew + o) = dn[a];
      *(GDP_GLOBAL D2*)(qv + o) = D2{qx, qy};
      acc += dn[a].x * qx + dn[a].y * qy;
    } else {
      if (outx) {
        dnew[o] = dn[a].x;
        qv[o] = qx;
        acc += dn[a].x * qx;
      }
      if (outy) {
        dnew[o + 1] = dn[a].y;
        qv[o + 1] = qy;
        acc += dn[a].y * qy;
      }
    }
  }
  write_partial<kThreads>(acc, 0.0, apart, t.slot, sh);
}

}  // namespace gdp
