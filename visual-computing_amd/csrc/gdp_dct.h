// Exact coarsest-level solve by an orthonormal transform in LDS (FP64), shared by the
// two coarse-tail kernels.
#pragma once
#include "gdp_device.h"

namespace gdp {

constexpr int kTailWaves = kTailThreads / 64;

// Exact solve of the coarsest level by an orthonormal transform that diagonalises its
// operator: U = Cm^T ((Cm B Cn^T) ./ lambda) Cn.  Neumann: DCT-II (the solvers.py:224-262
// construction, zero-mean solution: 1/lambda of the constant mode is 0).  Dirichlet:
// DST-I.  Tables live in LDS: Cm (m*m), Cn (n*n), 1/lambda (m*n), and optionally CnT,
// the transpose of Cn (n*n), which keeps the reads of the second pass on consecutive
// addresses across lanes (with Cn itself they are n elements apart, so the lanes of a
// row all hit one LDS bank).  Each of the four passes gives every thread of the block
// output elements of its own (one each for a level of at most nthreads points), with the
// loads of each inner loop issued together.
__device__ void lds_dct_solve(double* w0, double* w1, int m, int n, const double* tab,
                              int nthreads = kTailThreads, const double* CnT = nullptr) {
  const double* Cm = tab;
  const double* Cn = tab + m * m;
  const double* il = tab + m * m + n * n;
  const int N = m * n;
  for (int t = threadIdx.x; t < N; t += nthreads) {  // w1 = Cm w0
    const int k = t / n, j = t - k * n;
    double s = 0.0;
#pragma unroll 8
    for (int i = 0; i < m; ++i) s += Cm[k * m + i] * w0[i * n + j];
    w1[t] = s;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N; t += nthreads) {  // w0 = (w1 Cn^T) ./ lambda
    const int k = t / n, l = t - k * n;
    double s = 0.0;
    if (CnT) {
#pragma unroll 8
      for (int j = 0; j < n; ++j) s += w1[k * n + j] * CnT[j * n + l];
    } else {
#pragma unroll 8
      for (int j = 0; j < n; ++j) s += w1[k * n + j] * Cn[l * n + j];
    }
    w0[t] = s * il[t];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N; t += nthreads) {  // w1 = Cm^T w0
    const int i = t / n, l = t - i * n;
    double s = 0.0;
#pragma unroll 8
    for (int k = 0; k < m; ++k) s += Cm[k * m + i] * w0[k * n + l];
    w1[t] = s;
  }
  __syncthreads();
  for (int t = threadIdx.x; t < N; t += nthreads) {  // w0 = w1 Cn
    const int i = t / n, j = t - i * n;
    double s = 0.0;
#pragma unroll 8
    for (int l = 0; l < n; ++l) s += w1[i * n + l] * Cn[l * n + j];
    w0[t] = s;
  }
  __syncthreads();
}

}  // namespace gdp
