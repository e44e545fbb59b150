// Single-grid operators (FP64, dense row-major planes) behind the unit entry points
// gdp_laplacian / gdp_gradient / gdp_divergence / gdp_fold_boundary / gdp_smooth /
// gdp_restrict / gdp_prolong / gdp_relative_residual.  These are the stand-alone
// forms of the operators that the fused V-cycle kernels apply inside LDS tiles.
#include "gdp_device.h"
#include "gdp_internal.h"

namespace gdp {

namespace {

constexpr int BX = 64, BY = 4;

__device__ inline bool coord(int m, int n, int& i, int& j) {
  j = blockIdx.x * BX + (threadIdx.x & (BX - 1));
  i = blockIdx.y * BY + (threadIdx.x >> 6);
  return i < m && j < n;
}

// operators.py:86-107 (Dirichlet) / div(grad u) / h^2 (Neumann)
__global__ __launch_bounds__(256) void k_laplacian(const double* __restrict__ u,
                                                   double* __restrict__ out, int m, int n,
                                                   double ih2, int bc) {
  int i, j;
  if (!coord(m, n, i, j)) return;
  const size_t base = (size_t)blockIdx.z * m * n;
  const double* p = u + base;
  const size_t o = (size_t)i * n + j;
  const double up = i > 0 ? p[o - n] : 0.0;
  const double dn = i < m - 1 ? p[o + n] : 0.0;
  const double lf = j > 0 ? p[o - 1] : 0.0;
  const double rt = j < n - 1 ? p[o + 1] : 0.0;
  const double diag =
      bc == kNeumann ? double((i > 0) + (i < m - 1) + (j > 0) + (j < n - 1)) : 4.0;
  out[base + o] = ((((up + dn) + lf) + rt) - diag * p[o]) * ih2;
}

// operators.py:30-50
__global__ __launch_bounds__(256) void k_gradient(const double* __restrict__ img,
                                                  double* __restrict__ gx,
                                                  double* __restrict__ gy, int h, int w) {
  int i, j;
  if (!coord(h, w, i, j)) return;
  const size_t base = (size_t)blockIdx.z * h * w;
  const size_t o = base + (size_t)i * w + j;
  gx[o] = j < w - 1 ? img[o + 1] - img[o] : 0.0;
  gy[o] = i < h - 1 ? img[o + w] - img[o] : 0.0;
}

// operators.py:53-83
__global__ __launch_bounds__(256) void k_divergence(const double* __restrict__ gx,
                                                    const double* __restrict__ gy,
                                                    double* __restrict__ out, int h, int w) {
  int i, j;
  if (!coord(h, w, i, j)) return;
  const size_t base = (size_t)blockIdx.z * h * w;
  const size_t o = base + (size_t)i * w + j;
  double dx = 0.0, dy = 0.0;
  if (w > 1) {
    if (j == 0) dx = gx[o];
    else if (j == w - 1) dx = -gx[o - 1];
    else dx = gx[o] - gx[o - 1];
  }
  if (h > 1) {
    if (i == 0) dy = gy[o];
    else if (i == h - 1) dy = -gy[o - w];
    else dy = gy[o] - gy[o - w];
  }
  out[o] = dx + dy;
}

// operators.py:139-169: rows first (top, bottom), then columns (left, right), so a
// corner receives the row term before the column term, as the reference does.
__global__ __launch_bounds__(256) void k_fold(const double* __restrict__ rhs,
                                              const double* __restrict__ bd,
                                              double* __restrict__ out, int h, int w,
                                              double wgt) {
  const int mi = h - 2, ni = w - 2;
  int i, j;
  if (!coord(mi, ni, i, j)) return;
  const size_t ib = (size_t)blockIdx.z * mi * ni;
  const size_t bb = (size_t)blockIdx.z * h * w;
  double v = rhs[ib + (size_t)i * ni + j];
  if (i == 0) v -= wgt * bd[bb + (size_t)0 * w + (j + 1)];
  if (i == mi - 1) v -= wgt * bd[bb + (size_t)(h - 1) * w + (j + 1)];
  if (j == 0) v -= wgt * bd[bb + (size_t)(i + 1) * w + 0];
  if (j == ni - 1) v -= wgt * bd[bb + (size_t)(i + 1) * w + (w - 1)];
  out[ib + (size_t)i * ni + j] = v;
}

// One colour of red-black Gauss-Seidel, in place (multigrid.py:158-165).  Updated
// points only read points of the other colour, so in-place is race free.
__global__ __launch_bounds__(256) void k_halfsweep(double* __restrict__ u,
                                                   const double* __restrict__ b, int m, int n,
                                                   double h2, int bc, int parity) {
  const int half = (n + 1) >> 1;
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= m * half) return;
  const int i = k / half;
  const int j = 2 * (k - i * half) + ((i + parity) & 1);
  if (j >= n) return;
  const size_t o = (size_t)i * n + j;
  const double up = i > 0 ? u[o - n] : 0.0;
  const double dn = i < m - 1 ? u[o + n] : 0.0;
  const double lf = j > 0 ? u[o - 1] : 0.0;
  const double rt = j < n - 1 ? u[o + 1] : 0.0;
  const double val = (((up + dn) + lf) + rt) - h2 * b[o];
  if (bc == kNeumann) {
    const int cnt = (i > 0) + (i < m - 1) + (j > 0) + (j < n - 1);
    u[o] = cnt == 4 ? val * 0.25 : val / double(cnt);
  } else {
    u[o] = val * 0.25;
  }
}

__device__ inline double rweight(bool neu, int I, int k, int f, int c) {
  if (!neu) return k == 1 ? 1.0 : (k == 3 ? 0.0 : 0.5);
  const int i = 2 * I - 1 + k;
  if (i < 0 || i >= f) return 0.0;
  if (k == 0 || k == 3) return 0.25;
  if (k == 1) return I == 0 ? 1.0 : 0.75;
  return I == c - 1 ? 1.0 : 0.75;
}

// multigrid.py:169-174 (Dirichlet taps 2I..2I+2) / cell-centred taps 2I-1..2I+2.
__global__ __launch_bounds__(256) void k_restrict(const double* __restrict__ f,
                                                  double* __restrict__ c, int m, int n,
                                                  int mc, int nc, int bc) {
  int I, J;
  if (!coord(mc, nc, I, J)) return;
  const bool neu = bc == kNeumann;
  const int ntap = neu ? 4 : 3, tap0 = neu ? -1 : 0;
  double acc = 0.0;
  for (int cj = 0; cj < ntap; ++cj) {
    const int j = 2 * J + tap0 + cj;
    const double wc = (j >= 0 && j < n) ? rweight(neu, J, cj, n, nc) : 0.0;
    if (wc == 0.0) continue;
    double col = 0.0;
    for (int ri = 0; ri < ntap; ++ri) {
      const int i = 2 * I + tap0 + ri;
      const double wr = (i >= 0 && i < m) ? rweight(neu, I, ri, m, mc) : 0.0;
      if (wr == 0.0) continue;
      col += wr * f[(size_t)i * n + j];
    }
    acc += wc * col;
  }
  c[(size_t)I * nc + J] = 0.25 * acc;
}

// multigrid.py:177-181 (Dirichlet) / cell-centred bilinear with reflection (Neumann).
__global__ __launch_bounds__(256) void k_prolong(const double* __restrict__ c,
                                                 double* __restrict__ f, int m, int n, int mc,
                                                 int nc, int bc) {
  int i, j;
  if (!coord(m, n, i, j)) return;
  const bool neu = bc == kNeumann;
  int ri[2], cj[2];
  double rw[2], cw[2];
  int nr = 0, ncnt = 0;
  auto taps = [&](int x, int cs, int* idx, double* w, int& cnt) {
    cnt = 0;
    if (!neu) {
      if (x & 1) {
        const int s = (x - 1) >> 1;
        if (s >= 0 && s < cs) { idx[cnt] = s; w[cnt] = 1.0; ++cnt; }
      } else {
        const int s0 = (x >> 1) - 1, s1 = x >> 1;
        if (s0 >= 0 && s0 < cs) { idx[cnt] = s0; w[cnt] = 0.5; ++cnt; }
        if (s1 >= 0 && s1 < cs) { idx[cnt] = s1; w[cnt] = 0.5; ++cnt; }
      }
    } else {
      const int p = x >> 1;
      const int o = (x & 1) ? p + 1 : p - 1;
      if (p >= cs) return;
      if (o < 0 || o >= cs) { idx[0] = p; w[0] = 1.0; cnt = 1; }
      else { idx[0] = p; w[0] = 0.75; idx[1] = o; w[1] = 0.25; cnt = 2; }
    }
  };
  taps(i, mc, ri, rw, nr);
  taps(j, nc, cj, cw, ncnt);
  double acc = 0.0;
  for (int y = 0; y < ncnt; ++y) {
    double col = 0.0;
    for (int x = 0; x < nr; ++x) col += rw[x] * c[(size_t)ri[x] * nc + cj[y]];
    acc += cw[y] * col;
  }
  f[(size_t)i * n + j] = acc;
}

// Per-plane sums of (b - L u)^2 and b^2 (two passes of atomics-free block partials).
__global__ __launch_bounds__(256) void k_resnorm_partial(const double* __restrict__ u,
                                                         const double* __restrict__ b,
                                                         int m, int n, double ih2, int bc,
                                                         double2* __restrict__ partial) {
  const size_t N = (size_t)m * n;
  const size_t base = (size_t)blockIdx.y * N;
  double rr = 0.0, bb = 0.0;
  for (size_t o = (size_t)blockIdx.x * 256 + threadIdx.x; o < N; o += (size_t)gridDim.x * 256) {
    const int i = int(o / n), j = int(o - (size_t)i * n);
    const double* p = u + base;
    const double up = i > 0 ? p[o - n] : 0.0;
    const double dn = i < m - 1 ? p[o + n] : 0.0;
    const double lf = j > 0 ? p[o - 1] : 0.0;
    const double rt = j < n - 1 ? p[o + 1] : 0.0;
    const double diag =
        bc == kNeumann ? double((i > 0) + (i < m - 1) + (j > 0) + (j < n - 1)) : 4.0;
    const double r = b[base + o] - ((((up + dn) + lf) + rt) - diag * p[o]) * ih2;
    rr += r * r;
    bb += b[base + o] * b[base + o];
  }
  __shared__ double sh[8];
  for (int off = 32; off > 0; off >>= 1) {
    rr += __shfl_down(rr, off, 64);
    bb += __shfl_down(bb, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[threadIdx.x >> 6] = rr;
    sh[4 + (threadIdx.x >> 6)] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.y * gridDim.x + blockIdx.x] =
        make_double2(((sh[0] + sh[1]) + sh[2]) + sh[3], ((sh[4] + sh[5]) + sh[6]) + sh[7]);
}

// ------------------------------------------------------------------ compositing
// Guidance of a cloned region (poisson.py:52-92, 261-348): the forward-difference
// gradient (operators.py:30-50) of the source placed at (row, col) inside the target,
// combined inside the placed mask by mode (0 import: source; 1 mixed: per component
// the larger magnitude, source on ties; 2 average), the target's gradient outside.
struct CloneArgs {
  const double* src;      // source planes, sh x sw
  const uint8_t* mask;    // source-space mask, sh x sw
  const double* tgt;      // target planes, h x w
  int sh, sw, row, col, h, w, mode;
};

// Source value and mask at image pixel (i, j), zero outside the placed crop.  Loads are
// branch-free (clamped addresses, selects) so a thread issues them all before it waits.
__device__ inline double src_at(const CloneArgs& a, size_t sp, int i, int j) {
  const int si = i - a.row, sj = j - a.col;
  const bool in = si >= 0 && si < a.sh && sj >= 0 && sj < a.sw;
  const double v =
      a.src[sp + (size_t)min(max(si, 0), a.sh - 1) * a.sw + min(max(sj, 0), a.sw - 1)];
  return in ? v : 0.0;
}

__device__ inline bool mask_at(const CloneArgs& a, int i, int j) {
  const int si = i - a.row, sj = j - a.col;
  const bool in = si >= 0 && si < a.sh && sj >= 0 && sj < a.sw;
  const uint8_t v = a.mask[(size_t)min(max(si, 0), a.sh - 1) * a.sw + min(max(sj, 0), a.sw - 1)];
  return in && v != 0;
}

// Guidance (gx, gy) at image pixel (i, j); every pixel this is called on has its right
// and lower neighbours inside the image.  Both the target's and the source's differences
// are formed and the mask and mode select between them.
__device__ inline void clone_guidance(const CloneArgs& a, size_t sp, size_t tp, int i, int j,
                                      double& gx, double& gy) {
#pragma clang fp contract(off)
  const double* t = a.tgt + tp;
  const double tc = t[(size_t)i * a.w + j];
  const double tgx = t[(size_t)i * a.w + j + 1] - tc;
  const double tgy = t[(size_t)(i + 1) * a.w + j] - tc;
  const bool in = mask_at(a, i, j);
  const double sc = src_at(a, sp, i, j);
  const double sgx = src_at(a, sp, i, j + 1) - sc;
  const double sgy = src_at(a, sp, i + 1, j) - sc;
  double ix, iy;
  if (a.mode == 1) {
    ix = fabs(sgx) >= fabs(tgx) ? sgx : tgx;
    iy = fabs(sgy) >= fabs(tgy) ? sgy : tgy;
  } else if (a.mode == 2) {
    ix = 0.5 * (sgx + tgx);
    iy = 0.5 * (sgy + tgy);
  } else {
    ix = sgx;
    iy = sgy;
  }
  gx = in ? ix : tgx;
  gy = in ? iy : tgy;
}

// Folded right-hand side of the window [top, top+hh) x [left, left+ww) of one plane:
// divergence (operators.py:53-83) of the guidance at the window's interior points, minus
// the window's one-pixel ring of target values (fold_boundary, operators.py:139-169,
// h = 1, in its order: first row, last row, first column, last column); or, with an
// interior mask, the masked system's right-hand side.
__global__ __launch_bounds__(256) void k_clone_rhs(CloneArgs a, int top, int left, int hh,
                                                   int ww, const uint8_t* __restrict__ interior,
                                                   double* __restrict__ out) {
  // no fused multiply-adds: the result is bit-identical to the unfused FP64 formula
#pragma clang fp contract(off)
  const int mi = hh - 2, ni = ww - 2;
  int i, j;
  if (!coord(mi, ni, i, j)) return;
  const size_t sp = (size_t)blockIdx.z * a.sh * a.sw, tp = (size_t)blockIdx.z * a.h * a.w;
  const int gi = top + 1 + i, gj = left + 1 + j;
  double gx, gy, gxl, gyu;
  clone_guidance(a, sp, tp, gi, gj, gx, gy);
  double unused;
  clone_guidance(a, sp, tp, gi, gj - 1, gxl, unused);
  clone_guidance(a, sp, tp, gi - 1, gj, unused, gyu);
  double v = (gx - gxl) + (gy - gyu);
  const double* t = a.tgt + tp;
  if (interior) {
    // masked domain (solve_masked, poisson.py:208-216): no equation off the mask; the
    // neighbours that are not unknowns give their target values, up, down, left, right
    const size_t gp = (size_t)gi * a.w + gj;
    const double tu = t[gp - a.w], td = t[gp + a.w], tl = t[gp - 1], tr = t[gp + 1];
    const bool iu = interior[gp - a.w], id = interior[gp + a.w];
    const bool il = interior[gp - 1], ir = interior[gp + 1];
    v = iu ? v : v - tu;
    v = id ? v : v - td;
    v = il ? v : v - tl;
    v = ir ? v : v - tr;
    v = interior[gp] ? v : 0.0;
  } else {
    if (i == 0) v -= t[(size_t)top * a.w + gj];
    if (i == mi - 1) v -= t[(size_t)(top + hh - 1) * a.w + gj];
    if (j == 0) v -= t[(size_t)gi * a.w + left];
    if (j == ni - 1) v -= t[(size_t)gi * a.w + left + ww - 1];
  }
  out[(size_t)blockIdx.z * mi * ni + (size_t)i * ni + j] = v;
}

// Right-hand side of a masked solve from an explicit field (solve_masked,
// poisson.py:205-216): at the unknowns of the window [top, top+hh) x [left, left+ww)
// interior, divergence(gx, gy) minus the values of the neighbours that are not unknowns
// (up, down, left, right); zero elsewhere.  interior: h x w bytes, never set on the
// image's border rows and columns.
__global__ __launch_bounds__(256) void k_masked_rhs(const double* __restrict__ gx,
                                                    const double* __restrict__ gy,
                                                    const double* __restrict__ img,
                                                    const uint8_t* __restrict__ interior, int h,
                                                    int w, int top, int left, int hh, int ww,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
  const int mi = hh - 2, ni = ww - 2;
  int i, j;
  if (!coord(mi, ni, i, j)) return;
  const size_t pl = (size_t)blockIdx.z * h * w;
  const int gi = top + 1 + i, gj = left + 1 + j;
  const size_t gp = (size_t)gi * w + gj;
  // branch-free (every neighbour is inside the image): all loads issue before any wait
  const double* X = gx + pl;
  const double* Y = gy + pl;
  const double* t = img + pl;
  const double tu = t[gp - w], td = t[gp + w], tl = t[gp - 1], tr = t[gp + 1];
  double v = (X[gp] - X[gp - 1]) + (Y[gp] - Y[gp - w]);
  v = interior[gp - w] ? v : v - tu;
  v = interior[gp + w] ? v : v - td;
  v = interior[gp - 1] ? v : v - tl;
  v = interior[gp + 1] ? v : v - tr;
  out[(size_t)blockIdx.z * mi * ni + (size_t)i * ni + j] = interior[gp] ? v : 0.0;
}

// Element-wise guidance-field combinations (GuidanceField, poisson.py:52-92) over planes of
// hw pixels: 0 select (mask ? a : b), 1 mixed (per component |a| >= |b| ? a : b),
// 2 average, 3 scale (a * factor[pixel], or a * fscale without a factor array).
__global__ __launch_bounds__(256) void k_guidance(const double* __restrict__ ax,
                                                  const double* __restrict__ ay,
                                                  const double* __restrict__ bx,
                                                  const double* __restrict__ by,
                                                  const uint8_t* __restrict__ mask,
                                                  const double* __restrict__ factor,
                                                  double fscale, int mode, size_t hw,
                                                  size_t total, double* __restrict__ ox,
                                                  double* __restrict__ oy) {
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= total) return;
  const size_t px = k % hw;
  double x, y;
  if (mode == 0) {
    const bool in = mask[px] != 0;
    x = in ? ax[k] : bx[k];
    y = in ? ay[k] : by[k];
  } else if (mode == 1) {
    x = fabs(ax[k]) >= fabs(bx[k]) ? ax[k] : bx[k];
    y = fabs(ay[k]) >= fabs(by[k]) ? ay[k] : by[k];
  } else if (mode == 2) {
    x = 0.5 * (ax[k] + bx[k]);
    y = 0.5 * (ay[k] + by[k]);
  } else {
    const double f = factor ? factor[px] : fscale;
    x = ax[k] * f;
    y = ay[k] * f;
  }
  ox[k] = x;
  oy[k] = y;
}

// Gradient attenuation of illumination_change (poisson.py:433-437), per element of the
// field: f = scale * max(|(gx, gy)|, epsilon) ** -beta with scale = alpha ** beta, and
// (gx, gy) * f.  max keeps a NaN magnitude, as numpy.maximum does.
__global__ __launch_bounds__(256) void k_illumination(const double* __restrict__ gx,
                                                      const double* __restrict__ gy, size_t n,
                                                      double scale, double beta,
                                                      double epsilon, double* __restrict__ ox,
                                                      double* __restrict__ oy) {
#pragma clang fp contract(off)
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  const double x = gx[k], y = gy[k];
  const double mag = sqrt(x * x + y * y);
  const double f = scale * pow(mag < epsilon ? epsilon : mag, -beta);
  ox[k] = x * f;
  oy[k] = y * f;
}

dim3 grid2d(int m, int n, int planes) {
  return dim3((n + BX - 1) / BX, (m + BY - 1) / BY, planes);
}

}  // namespace

hipError_t launch_clone_rhs(const double* src, int sh, int sw, int row, int col,
                            const uint8_t* mask, const double* tgt, int h, int w, int planes,
                            int top, int left, int hh, int ww, int mode,
                            const uint8_t* interior, double* out, hipStream_t s) {
  CloneArgs a{src, mask, tgt, sh, sw, row, col, h, w, mode};
  k_clone_rhs<<<grid2d(hh - 2, ww - 2, planes), 256, 0, s>>>(a, top, left, hh, ww, interior,
                                                             out);
  return hipGetLastError();
}

hipError_t launch_masked_rhs(const double* gx, const double* gy, const double* img,
                             const uint8_t* interior, int h, int w, int planes, int top,
                             int left, int hh, int ww, double* out, hipStream_t s) {
  k_masked_rhs<<<grid2d(hh - 2, ww - 2, planes), 256, 0, s>>>(gx, gy, img, interior, h, w, top,
                                                              left, hh, ww, out);
  return hipGetLastError();
}

hipError_t launch_guidance(const double* ax, const double* ay, const double* bx,
                           const double* by, const uint8_t* mask, const double* factor,
                           double fscale, int mode, size_t hw, int planes, double* ox,
                           double* oy, hipStream_t s) {
  const size_t total = hw * size_t(planes);
  if (!total) return hipSuccess;
  k_guidance<<<unsigned((total + 255) / 256), 256, 0, s>>>(ax, ay, bx, by, mask, factor, fscale,
                                                          mode, hw, total, ox, oy);
  return hipGetLastError();
}

hipError_t launch_illumination(const double* gx, const double* gy, size_t n, double scale,
                               double beta, double epsilon, double* ox, double* oy,
                               hipStream_t s) {
  if (!n) return hipSuccess;
  k_illumination<<<unsigned((n + 255) / 256), 256, 0, s>>>(gx, gy, n, scale, beta, epsilon, ox,
                                                           oy);
  return hipGetLastError();
}

hipError_t launch_laplacian(const double* u, double* out, int m, int n, int planes,
                            double spacing, int bc, hipStream_t s) {
  k_laplacian<<<grid2d(m, n, planes), 256, 0, s>>>(u, out, m, n, 1.0 / (spacing * spacing), bc);
  return hipGetLastError();
}

hipError_t launch_gradient(const double* img, double* gx, double* gy, int h, int w, int planes,
                           hipStream_t s) {
  k_gradient<<<grid2d(h, w, planes), 256, 0, s>>>(img, gx, gy, h, w);
  return hipGetLastError();
}

hipError_t launch_divergence(const double* gx, const double* gy, double* out, int h, int w,
                             int planes, hipStream_t s) {
  k_divergence<<<grid2d(h, w, planes), 256, 0, s>>>(gx, gy, out, h, w);
  return hipGetLastError();
}

hipError_t launch_fold(const double* rhs, const double* bd, double* out, int h, int w,
                       int planes, double spacing, hipStream_t s) {
  k_fold<<<grid2d(h - 2, w - 2, planes), 256, 0, s>>>(rhs, bd, out, h, w,
                                                      1.0 / (spacing * spacing));
  return hipGetLastError();
}

hipError_t launch_smooth(double* u, const double* b, int m, int n, double spacing, int sweeps,
                         int bc, int order, hipStream_t s) {
  const int half = (n + 1) >> 1;
  const int blocks = (m * half + 255) / 256;
  const double h2 = spacing * spacing;
  for (int k = 0; k < sweeps; ++k) {
    const int p0 = order == 0 ? 0 : 1;
    k_halfsweep<<<blocks, 256, 0, s>>>(u, b, m, n, h2, bc, p0);
    k_halfsweep<<<blocks, 256, 0, s>>>(u, b, m, n, h2, bc, 1 - p0);
  }
  return hipGetLastError();
}

hipError_t launch_restrict(const double* f, double* c, int m, int n, int mc, int nc, int bc,
                           hipStream_t s) {
  k_restrict<<<grid2d(mc, nc, 1), 256, 0, s>>>(f, c, m, n, mc, nc, bc);
  return hipGetLastError();
}

hipError_t launch_prolong(const double* c, double* f, int m, int n, int mc, int nc, int bc,
                          hipStream_t s) {
  k_prolong<<<grid2d(m, n, 1), 256, 0, s>>>(c, f, m, n, mc, nc, bc);
  return hipGetLastError();
}

hipError_t launch_resnorm_partial(const double* u, const double* b, int m, int n, int planes,
                                  double spacing, int bc, double2* partial, int blocks,
                                  hipStream_t s) {
  k_resnorm_partial<<<dim3(blocks, planes), 256, 0, s>>>(u, b, m, n, 1.0 / (spacing * spacing),
                                                          bc, partial);
  return hipGetLastError();
}

}  // namespace gdp
