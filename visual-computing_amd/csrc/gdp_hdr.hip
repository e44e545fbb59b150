// Gradient-domain tone mapping on the device: the front end and back end around the
// Neumann solve of gradient-domain/src/gradient_domain/hdr.py.
//
//   gdp_tm_log_luminance  hdr.py:165-167  log(max(radiance . (0.2126, 0.7152, 0.0722), tiny))
//   gdp_tm_downsample     hdr.py:73-80    5-tap binomial blur, edge padding, keep every
//                                         other row and column
//   gdp_tm_level_factor   hdr.py:126-133  one level of attenuation_field: central-gradient
//                                         magnitude, its mean, the local factor, times the
//                                         bilinear upsampling (hdr.py:83-96) of the
//                                         coarser level's factor
//   gdp_tm_rhs            hdr.py:174-175  divergence(gx f, gy f) of the forward gradient
//   gdp_tm_compose        hdr.py:180-182  ratios^saturation * exp(c - max c)
//   gdp_percentile        hdr.py:184      numpy.percentile (linear) by radix selection
//   gdp_tm_normalise      hdr.py:185      clip(image / max(peak, 1e-12), 0, 1)
//
// Every image is one FP64 plane, row-major, h x w; radiance is h x w x 3 channels-last.
// Products and sums are evaluated in the reference's order with no fused multiply-add,
// so the blur, the upsampling and the divergence give the reference's values exactly;
// the sums behind a mean or a maximum are block partials reduced in a fixed order, and
// the percentile's histograms count with integer atomics, whose totals do not depend
// on the order: every result is the same from run to run.
#include "gdpoisson.h"
#include "gdp_internal.h"
#include "gdp_reduce.h"

#include <cmath>

#define GDP_HDR_HIP(x)                                                          \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      gdp::set_error("%s failed: %s", #x, hipGetErrorString(e_));               \
      return GDP_EHIP;                                                          \
    }                                                                           \
  } while (0)

namespace gdp {
namespace {

constexpr int kHdrThreads = 256;
constexpr int kMaxPartials = 1024;  // blocks of a grid-stride reduction pass

__device__ inline double luminance(const double* __restrict__ rad, size_t px) {
#pragma clang fp contract(off)
  const double l = (rad[3 * px] * 0.2126 + rad[3 * px + 1] * 0.7152) + rad[3 * px + 2] * 0.0722;
  return fmax(l, 2.2250738585072014e-308);  // numpy.finfo(float).tiny
}

__global__ __launch_bounds__(kHdrThreads) void k_log_luminance(const double* __restrict__ rad,
                                                               size_t n, double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (i < n) out[i] = log(luminance(rad, i));
}

// out (I, J) = sum_a k_a * (sum_b k_b * x(clamp(2I - 2 + a), clamp(2J - 2 + b))): the
// horizontal pass first, each sum from 0 left to right as Python's sum() builds it.
__global__ __launch_bounds__(kHdrThreads) void k_downsample(const double* __restrict__ x, int h,
                                                            int w, double* __restrict__ out,
                                                            int ho, int wo) {
#pragma clang fp contract(off)
  const size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (t >= (size_t)ho * wo) return;
  const int I = int(t / wo), J = int(t - (size_t)I * wo);
  const double k[5] = {1.0 / 16.0, 4.0 / 16.0, 6.0 / 16.0, 4.0 / 16.0, 1.0 / 16.0};
  int cc[5];
#pragma unroll
  for (int b = 0; b < 5; ++b) cc[b] = min(max(2 * J - 2 + b, 0), w - 1);
  double v = 0.0;
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    const double* row = x + (size_t)min(max(2 * I - 2 + a, 0), h - 1) * w;
    double hs = 0.0;
#pragma unroll
    for (int b = 0; b < 5; ++b) hs = hs + k[b] * row[cc[b]];
    v = v + k[a] * hs;
  }
  out[t] = v;
}

// Central-gradient magnitude (edge-replicated, spacing s) into mag, and one partial sum
// per block (grid-stride, fixed order).
__global__ __launch_bounds__(kHdrThreads) void k_magnitude(const double* __restrict__ x, int h,
                                                           int w, double spacing,
                                                           double* __restrict__ mag,
                                                           double* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[kHdrThreads / 64];
  const size_t n = (size_t)h * w;
  const double two_s = 2.0 * spacing;
  double acc = 0.0;
  for (size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x; t < n;
       t += (size_t)gridDim.x * kHdrThreads) {
    const int i = int(t / w), j = int(t - (size_t)i * w);
    const double* row = x + (size_t)i * w;
    const double gx = (row[min(j + 1, w - 1)] - row[max(j - 1, 0)]) / two_s;
    const double gy = (x[(size_t)min(i + 1, h - 1) * w + j] - x[(size_t)max(i - 1, 0) * w + j]) / two_s;
    const double m = sqrt(gx * gx + gy * gy);
    mag[t] = m;
    acc += m;
  }
  const double s = block_sum<kHdrThreads>(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[count] = sum of partial[0 .. count) (one block, fixed order).
__global__ __launch_bounds__(kHdrThreads) void k_sum_partials(double* __restrict__ partial,
                                                              int count) {
  __shared__ double sh[kHdrThreads / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += kHdrThreads) acc += partial[i];
  const double s = block_sum<kHdrThreads>(acc, sh);
  if (threadIdx.x == 0) partial[count] = s;
}

// Bilinear sample of the coarse factor (hc x wc) at fine pixel (i, j) of an h x w grid,
// at numpy.linspace(0, hc - 1, h) x numpy.linspace(0, wc - 1, w) (hdr.py:83-96).
__device__ inline double linspace_at(int i, int num, int top) {
#pragma clang fp contract(off)
  if (num < 2) return 0.0;
  if (i == num - 1) return double(top);
  const double step = double(top) / double(num - 1);
  return step == 0.0 ? 0.0 : double(i) * step;
}

__device__ inline double upsample_at(const double* __restrict__ c, int hc, int wc, int h, int w,
                                     int i, int j) {
#pragma clang fp contract(off)
  const double r = linspace_at(i, h, hc - 1), q = linspace_at(j, w, wc - 1);
  const int r0 = int(floor(r)), c0 = int(floor(q));
  const int r1 = min(r0 + 1, hc - 1), c1 = min(c0 + 1, wc - 1);
  const double rw = r - double(r0), cw = q - double(c0);
  const double top = c[(size_t)r0 * wc + c0] * (1.0 - cw) + c[(size_t)r0 * wc + c1] * cw;
  const double bottom = c[(size_t)r1 * wc + c0] * (1.0 - cw) + c[(size_t)r1 * wc + c1] * cw;
  return top * (1.0 - rw) + bottom * rw;
}

// f = (scale / g) (g / scale)^beta, g = max(mag, eps), scale = alpha max(mean, eps),
// times the upsampled coarse factor when there is one.  In place over mag.
__global__ __launch_bounds__(kHdrThreads) void k_level_factor(double* __restrict__ f, int h, int w,
                                                              const double* __restrict__ total,
                                                              double alpha, double beta,
                                                              double eps,
                                                              const double* __restrict__ coarse,
                                                              int hc, int wc) {
#pragma clang fp contract(off)
  const size_t n = (size_t)h * w;
  const size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (t >= n) return;
  const double mean = *total / double(n);
  const double scale = alpha * fmax(mean, eps);
  const double g = fmax(f[t], eps);
  double local = (scale / g) * pow(g / scale, beta);
  if (coarse) {
    const int i = int(t / w), j = int(t - (size_t)i * w);
    local = upsample_at(coarse, hc, wc, h, w, i, j) * local;
  }
  f[t] = local;
}

// rhs = divergence(gx f, gy f), gx / gy the forward differences of L (last column / row
// zero), divergence its negative adjoint (operators.py:30-83).
__global__ __launch_bounds__(kHdrThreads) void k_tm_rhs(const double* __restrict__ L,
                                                        const double* __restrict__ f, int h, int w,
                                                        double* __restrict__ rhs) {
#pragma clang fp contract(off)
  const size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (t >= (size_t)h * w) return;
  const int i = int(t / w), j = int(t - (size_t)i * w);
  // (gx f) at (i, j) and (i, j - 1), (gy f) at (i, j) and (i - 1, j), each from clamped
  // (in-bounds) addresses so every load issues before any wait; the cases below pick the
  // terms that exist, as the branches of the divergence do
  const int jr = min(j + 1, w - 1), jl = max(j - 1, 0);
  const int id = min(i + 1, h - 1), iu = max(i - 1, 0);
  const double* Li = L + (size_t)i * w;
  const double* fi = f + (size_t)i * w;
  const double Lc = Li[j], Lr = Li[jr], Ll = Li[jl];
  const double Ld = L[(size_t)id * w + j], Lu = L[(size_t)iu * w + j];
  const double fc = fi[j], fl = fi[jl], fu = f[(size_t)iu * w + j];
  const double fxr = (Lr - Lc) * fc, fxl = (Lc - Ll) * fl;  // (gx f)(i, j), (i, j - 1)
  const double fyd = (Ld - Lc) * fc, fyu = (Lc - Lu) * fu;  // (gy f)(i, j), (i - 1, j)
  // a one-column (one-row) image has no horizontal (vertical) differences: that term is
  // zero (operators.py:77-84)
  double dx = 0.0, dy = 0.0;
  if (w > 1) dx = j == 0 ? fxr : (j == w - 1 ? -fxl : fxr - fxl);
  if (h > 1) dy = i == 0 ? fyd : (i == h - 1 ? -fyu : fyd - fyu);
  rhs[t] = dx + dy;
}

// Per-block maxima of c (grid-stride).
__global__ __launch_bounds__(kHdrThreads) void k_max_partial(const double* __restrict__ c,
                                                             size_t n, double* __restrict__ partial) {
  __shared__ double sh[kHdrThreads / 64];
  double m = -INFINITY;
  for (size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x; t < n;
       t += (size_t)gridDim.x * kHdrThreads)
    m = fmax(m, c[t]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sh[0];
    for (int k = 1; k < kHdrThreads / 64; ++k) b = fmax(b, sh[k]);
    partial[blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(64) void k_max_partials(double* __restrict__ partial, int count) {
  double m = -INFINITY;
  for (int i = threadIdx.x; i < count; i += 64) m = fmax(m, partial[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if (threadIdx.x == 0) partial[count] = m;
}

// image(px, ch) = (radiance / luminance)^saturation * exp(c - max c).
__global__ __launch_bounds__(kHdrThreads) void k_tm_compose(const double* __restrict__ rad,
                                                            const double* __restrict__ c, size_t n,
                                                            const double* __restrict__ cmax,
                                                            double saturation,
                                                            double* __restrict__ image) {
#pragma clang fp contract(off)
  const size_t px = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (px >= n) return;
  const double lum = luminance(rad, px);
  const double r = exp(c[px] - *cmax);
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) image[3 * px + ch] = pow(rad[3 * px + ch] / lum, saturation) * r;
}

__global__ __launch_bounds__(kHdrThreads) void k_tm_normalise(double* __restrict__ image, size_t n,
                                                              const double* __restrict__ peak) {
  const size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x;
  if (t >= n) return;
  const double v = image[t] / fmax(*peak, 1e-12);
  image[t] = fmin(fmax(v, 0.0), 1.0);
}

// ------------------------------------------------------------ percentile by radix select
// Two order statistics (ranks lo and hi of the ascending order) are found together, eight
// bits per pass from the top, over keys that order like the doubles: a non-negative
// double's bits with the sign bit set, a negative one's bits inverted.
struct SelectState {
  unsigned long long prefix[2];  // key bits fixed so far
  unsigned long long mask;       // which bits are fixed
  unsigned long long rank[2];    // rank still to find among the keys matching prefix
  unsigned int hist[2][256];
  double out;
};

__device__ inline unsigned long long order_key(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double key_value(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

__global__ __launch_bounds__(256) void k_select_init(SelectState* s, unsigned long long lo,
                                                    unsigned long long hi) {
  if (threadIdx.x == 0) {
    s->prefix[0] = s->prefix[1] = 0;
    s->mask = 0;
    s->rank[0] = lo;
    s->rank[1] = hi;
    s->out = 0.0;
  }
  s->hist[0][threadIdx.x] = 0;
  s->hist[1][threadIdx.x] = 0;
}

__global__ __launch_bounds__(kHdrThreads) void k_select_hist(const double* __restrict__ x, size_t n,
                                                            SelectState* __restrict__ s, int shift) {
  __shared__ unsigned int lh[2][256];
  lh[0][threadIdx.x] = 0;
  lh[1][threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long mask = s->mask, p0 = s->prefix[0], p1 = s->prefix[1];
  for (size_t t = (size_t)blockIdx.x * kHdrThreads + threadIdx.x; t < n;
       t += (size_t)gridDim.x * kHdrThreads) {
    const unsigned long long k = order_key(x[t]);
    const unsigned int d = unsigned((k >> shift) & 255ull);
    if ((k & mask) == p0) atomicAdd(&lh[0][d], 1u);
    if ((k & mask) == p1) atomicAdd(&lh[1][d], 1u);
  }
  __syncthreads();
  if (lh[0][threadIdx.x]) atomicAdd(&s->hist[0][threadIdx.x], lh[0][threadIdx.x]);
  if (lh[1][threadIdx.x]) atomicAdd(&s->hist[1][threadIdx.x], lh[1][threadIdx.x]);
}

// Fix the next digit of both keys, clear the histograms; after the last pass, the
// linear interpolation of numpy's percentile (its _lerp: b - (b - a)(1 - t) for t >= 0.5).
__global__ __launch_bounds__(256) void k_select_pick(SelectState* s, int shift, int last,
                                                    double t, double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ unsigned int h[2][256];
  h[0][threadIdx.x] = s->hist[0][threadIdx.x];
  h[1][threadIdx.x] = s->hist[1][threadIdx.x];
  __syncthreads();
  if (threadIdx.x < 2) {
    const int q = threadIdx.x;
    unsigned long long r = s->rank[q], below = 0;
    int d = 0;
    for (; d < 255; ++d) {
      if (below + h[q][d] > r) break;
      below += h[q][d];
    }
    s->rank[q] = r - below;
    s->prefix[q] |= (unsigned long long)d << shift;
  }
  __syncthreads();
  s->hist[0][threadIdx.x] = 0;
  s->hist[1][threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    s->mask |= 255ull << shift;
    if (last) {
      const double a = key_value(s->prefix[0]), b = key_value(s->prefix[1]);
      const double d = b - a;
      s->out = t >= 0.5 ? b - d * (1.0 - t) : a + d * t;
      *out = s->out;
    }
  }
}

unsigned grid_for(size_t n) { return unsigned((n + kHdrThreads - 1) / kHdrThreads); }
int partial_blocks(size_t n) {
  return int(std::min<size_t>((n + kHdrThreads * 4 - 1) / (kHdrThreads * 4), kMaxPartials));
}

}  // namespace
}  // namespace gdp

using namespace gdp;

extern "C" {

int gdp_tm_work_bytes(int h, int w, size_t* bytes) {
  if (!bytes || h < 1 || w < 1) return (set_error("bad arguments"), GDP_EINVAL);
  *bytes = sizeof(double) * (kMaxPartials + 1);
  return GDP_OK;
}

int gdp_tm_log_luminance(const double* radiance, int h, int w, double* out, void* stream) {
  if (!radiance || !out || h < 1 || w < 1) return (set_error("bad arguments"), GDP_EINVAL);
  const size_t n = (size_t)h * w;
  k_log_luminance<<<grid_for(n), kHdrThreads, 0, (hipStream_t)stream>>>(radiance, n, out);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

int gdp_tm_downsample(const double* in, int h, int w, double* out, void* stream) {
  if (!in || !out || h < 1 || w < 1) return (set_error("bad arguments"), GDP_EINVAL);
  const int ho = (h + 1) / 2, wo = (w + 1) / 2;
  k_downsample<<<grid_for((size_t)ho * wo), kHdrThreads, 0, (hipStream_t)stream>>>(in, h, w, out,
                                                                                  ho, wo);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

int gdp_tm_level_factor(const double* level, int h, int w, double spacing, double alpha,
                        double beta, double epsilon, const double* coarse, int hc, int wc,
                        double* factor, double* work, void* stream) {
  if (!level || !factor || !work || h < 1 || w < 1 || !(spacing > 0.0))
    return (set_error("bad arguments"), GDP_EINVAL);
  if (coarse && (hc < 1 || wc < 1)) return (set_error("bad coarse factor shape"), GDP_EINVAL);
  const hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)h * w;
  const int nb = partial_blocks(n);
  k_magnitude<<<nb, kHdrThreads, 0, s>>>(level, h, w, spacing, factor, work);
  GDP_HDR_HIP(hipGetLastError());
  k_sum_partials<<<1, kHdrThreads, 0, s>>>(work, nb);
  GDP_HDR_HIP(hipGetLastError());
  k_level_factor<<<grid_for(n), kHdrThreads, 0, s>>>(factor, h, w, work + nb, alpha, beta,
                                                     epsilon, coarse, hc, wc);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

int gdp_tm_rhs(const double* log_lum, const double* factor, int h, int w, double* rhs,
               void* stream) {
  if (!log_lum || !factor || !rhs || h < 1 || w < 1) return (set_error("bad arguments"), GDP_EINVAL);
  k_tm_rhs<<<grid_for((size_t)h * w), kHdrThreads, 0, (hipStream_t)stream>>>(log_lum, factor, h,
                                                                              w, rhs);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

int gdp_tm_compose(const double* radiance, const double* compressed, int h, int w,
                   double saturation, double* image, double* work, void* stream) {
  if (!radiance || !compressed || !image || !work || h < 1 || w < 1)
    return (set_error("bad arguments"), GDP_EINVAL);
  const hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)h * w;
  const int nb = partial_blocks(n);
  k_max_partial<<<nb, kHdrThreads, 0, s>>>(compressed, n, work);
  GDP_HDR_HIP(hipGetLastError());
  k_max_partials<<<1, 64, 0, s>>>(work, nb);
  GDP_HDR_HIP(hipGetLastError());
  k_tm_compose<<<grid_for(n), kHdrThreads, 0, s>>>(radiance, compressed, n, work + nb, saturation,
                                                   image);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

int gdp_percentile_work_bytes(size_t* bytes) {
  if (!bytes) return (set_error("null argument"), GDP_EINVAL);
  *bytes = sizeof(SelectState);
  return GDP_OK;
}

int gdp_percentile(const double* x, size_t n, double q, double* out, void* work, void* stream) {
  if (!x || !out || !work || n < 1 || n > 0xffffffffull || !(q >= 0.0 && q <= 100.0))
    return (set_error("bad arguments"), GDP_EINVAL);
  // numpy's linear method: virtual index (n - 1) q', q' = q / 100
  const double qq = q / 100.0;
  const double v = double(n - 1) * qq;
  const double fl = std::floor(v);
  const double top = double(n - 1);
  const unsigned long long lo = (unsigned long long)std::min(std::max(fl, 0.0), top);
  const unsigned long long hi = std::min<unsigned long long>(lo + 1, n - 1);
  const double t = v - fl;
  const hipStream_t s = (hipStream_t)stream;
  SelectState* st = static_cast<SelectState*>(work);
  k_select_init<<<1, 256, 0, s>>>(st, lo, hi);
  GDP_HDR_HIP(hipGetLastError());
  const unsigned nb = (unsigned)std::min<size_t>((n + kHdrThreads * 8 - 1) / (kHdrThreads * 8),
                                                 2048);
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = 56 - 8 * pass;
    k_select_hist<<<nb, kHdrThreads, 0, s>>>(x, n, st, shift);
    GDP_HDR_HIP(hipGetLastError());
    k_select_pick<<<1, 256, 0, s>>>(st, shift, pass == 7 ? 1 : 0, t, out);
    GDP_HDR_HIP(hipGetLastError());
  }
  return GDP_OK;
}

int gdp_tm_normalise(double* image, size_t n, const double* peak, void* stream) {
  if (!image || !peak || n < 1) return (set_error("bad arguments"), GDP_EINVAL);
  k_tm_normalise<<<grid_for(n), kHdrThreads, 0, (hipStream_t)stream>>>(image, n, peak);
  GDP_HDR_HIP(hipGetLastError());
  return GDP_OK;
}

}  // extern "C"
