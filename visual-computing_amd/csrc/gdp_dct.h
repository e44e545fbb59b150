// Exact coarsest-level solve by an orthonormal transform in LDS (FP64), shared by the
// two coarse-tail kernels.
#pragma once
#include "gdp_device.h"

namespace gdp {

constexpr int kTailWaves = kTailThreads / 64;

// Exact solve of the coarsest level by an orthonormal transform that diagonalises its
// operator: U = Cm^T ((Cm B Cn^T) ./ lambda) Cn.  Neumann: DCT-II (the solvers.py:224-262
// construction, zero-mean solution: 1/lambda of the constant mode is 0).  Dirichlet:
// DST-I.  Tables live in LDS: Cm (m*m), Cn (n*n), 1/lambda (m*n).  Each pass puts
// output rows on the block's nwaves waves and output columns on lanes.
__device__ void lds_dct_solve(double* w0, double* w1, int m, int n, const double* tab,
                              int nwaves = kTailWaves) {
  const double* Cm = tab;
  const double* Cn = tab + m * m;
  const double* il = tab + m * m + n * n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int k = wave; k < m; k += nwaves)  // w1 = Cm w0
    for (int j = lane; j < n; j += 64) {
      double s = 0.0;
#pragma unroll 8
      for (int i = 0; i < m; ++i) s += Cm[k * m + i] * w0[i * n + j];
      w1[k * n + j] = s;
    }
  __syncthreads();
  for (int k = wave; k < m; k += nwaves)  // w0 = (w1 Cn^T) ./ lambda
    for (int l = lane; l < n; l += 64) {
      double s = 0.0;
#pragma unroll 8
      for (int j = 0; j < n; ++j) s += w1[k * n + j] * Cn[l * n + j];
      w0[k * n + l] = s * il[k * n + l];
    }
  __syncthreads();
  for (int i = wave; i < m; i += nwaves)  // w1 = Cm^T w0
    for (int l = lane; l < n; l += 64) {
      double s = 0.0;
#pragma unroll 8
      for (int k = 0; k < m; ++k) s += Cm[k * m + i] * w0[k * n + l];
      w1[i * n + l] = s;
    }
  __syncthreads();
  for (int i = wave; i < m; i += nwaves)  // w0 = w1 Cn
    for (int j = lane; j < n; j += 64) {
      double s = 0.0;
#pragma unroll 8
      for (int l = 0; l < n; ++l) s += w1[i * n + l] * Cn[l * n + j];
      w0[i * n + j] = s;
    }
  __syncthreads();
}

}  // namespace gdp
