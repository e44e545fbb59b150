// Block-level reductions shared by every kernel of libgdpoisson.
//
// Per-problem sums (dot products, norms) are deterministic and use no atomics: each
// block writes its partial with plain stores to its own slot, and the kernels that
// need the total reduce the problem's slots in a fixed order.  A problem's result
// therefore does not depend on what else is in the batch.
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
  int i = lo;
  // four loads in flight at a time, added in slot order as one at a time would be
  for (; i + 4 <= hi; i += 4) {
    const double2 v0 = partial[i], v1 = partial[i + 1], v2 = partial[i + 2], v3 = partial[i + 3];
    sx += v0.x;
    sy += v0.y;
    sx += v1.x;
    sy += v1.y;
    sx += v2.x;
    sy += v2.y;
    sx += v3.x;
    sy += v3.y;
  }
  for (; i < hi; ++i) {
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

}  // namespace gdp
