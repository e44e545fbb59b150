// Exact Neumann solve on the device by the orthonormal 2-D DCT-II, the GPU form of
// gradient-domain/src/gradient_domain/solvers.py:224-262 (solve_neumann):
//
//   u = idctn( dctn(b, II, ortho) / lambda ),  lambda_kl = 2 cos(pi k / m) + 2 cos(pi l / n) - 4,
//   coefficient (0, 0) set to 0 (zero-mean solution; a constant in b is discarded).
//
// Every 1-D transform is one complex FFT of the row reordered as Makhoul's algorithm
// prescribes (even samples forward, odd samples backward), followed by a quarter-sample
// rotation; the inverse (DCT-III) runs the same steps backwards through a conjugated
// forward FFT.  One block transforms one row: the row sits in LDS as complex FP64, and a
// Stockham autosort FFT (radix 4, 2, 3 and 5 stages, twiddles from exact tables) runs
// between two LDS buffers with one barrier per stage.  Each row pass writes its result
// transposed, so four passes make the forward and inverse 2-D transforms in the original
// orientation; the division by lambda is fused into the second pass.  Sides of at most
// 4096 points whose only prime factors are 2, 3 and 5 are supported.
//
// Exact Dirichlet solve (the rectangle operator of operators.sparse_laplacian,
// operators.py:110-136, which solvers._solve_direct factorises with SuperLU,
// solvers.py:85-90): the orthonormal DST-I diagonalises it,
//
//   u = S_m ( (S_m b S_n) ./ lambda ) S_n,  lambda_kl = (2 cos(pi (k+1) / (m+1)) +
//                                                        2 cos(pi (l+1) / (n+1)) - 4) / h^2,
//
// S symmetric and its own inverse.  A DST-I of length L is the imaginary part of the FFT
// of the row's odd extension (length 2 (L + 1)); two rows share one complex FFT as its
// real and imaginary parts.  Sides L with 2 (L + 1) <= 4096 and 2 (L + 1) a product of
// 2, 3 and 5 are supported.
#include "gdpoisson.h"
#include "gdp_internal.h"

#include <cmath>
#include <mutex>
#include <vector>

#define GDP_FFT_HIP(x)                                                          \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      gdp::set_error("%s failed: %s", #x, hipGetErrorString(e_));               \
      return GDP_EHIP;                                                          \
    }                                                                           \
  } while (0)

// Work buffer of gdp_neumann_dct_solve_static: a file-scope device array, one copy per GPU
// the library runs on, held until the process ends.  Its name is a host-side stand-in
// (not a device address): host code only ever passes the per-device address that
// hipGetSymbolAddress returns.
constexpr size_t kStaticWorkDoubles = size_t(1) << 23;  // 64 MiB: m n <= 4,194,304
static __device__ __align__(256) double scratch[kStaticWorkDoubles];

namespace gdp {
namespace {

constexpr int kFftThreads = 256;
constexpr int kFftMaxN = 4096;
constexpr int kFftMaxStages = 16;

struct FftShape {
  int N;
  int stages;
  int radix[kFftMaxStages];
};

// Radices of n, largest stages first (4 where it can, then 2, 3, 5); false if n has
// another prime factor or is out of range.
bool factor(int n, FftShape& f) {
  if (n < 1 || n > kFftMaxN) return false;
  f.N = n;
  f.stages = 0;
  int r = n;
  while (r % 4 == 0) { f.radix[f.stages++] = 4; r /= 4; }
  while (r % 2 == 0) { f.radix[f.stages++] = 2; r /= 2; }
  while (r % 3 == 0) { f.radix[f.stages++] = 3; r /= 3; }
  while (r % 5 == 0) { f.radix[f.stages++] = 5; r /= 5; }
  return r == 1 && f.stages <= kFftMaxStages;
}

// Table of one dimension N (5 N doubles): exp(-2 pi i t / N) for t < N (FFT twiddles),
// exp(-i pi k / (2 N)) for k < N (DCT quarter-sample rotation), 2 cos(pi k / N) (the
// dimension's part of lambda).
size_t table_doubles(int N) { return size_t(5) * N; }

__device__ inline double2 cmul(double2 a, double2 b) {
  return double2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ inline double2 cadd(double2 a, double2 b) { return double2{a.x + b.x, a.y + b.y}; }
__device__ inline double2 csub(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
// a * (-i)
__device__ inline double2 mul_mi(double2 a) { return double2{a.y, -a.x}; }

// In-register forward DFT of R points: y_q = sum_p v_p exp(-2 pi i p q / R).
template <int R>
__device__ inline void dft(double2* v) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const double2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const double2 t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    const double c = -0.5, s = 0.86602540378443864676;  // cos, sin of 2 pi / 3
    const double2 a = v[0], b = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const double2 t = double2{a.x + c * b.x, a.y + c * b.y};
    const double2 u = double2{s * d.y, -s * d.x};  // -i s (v1 - v2)
    v[0] = cadd(a, b);
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  } else {  // R == 5
    const double c1 = 0.30901699437494742410, s1 = 0.95105651629515357212;   // 2 pi / 5
    const double c2 = -0.80901699437494742410, s2 = 0.58778525229247312917;  // 4 pi / 5
    const double2 a = v[0];
    const double2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    const double2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    const double2 t1 = double2{a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y};
    const double2 t2 = double2{a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y};
    // -i (s1 d1 + s2 d2) and -i (s2 d1 - s1 d2)
    const double2 u1 = mul_mi(double2{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y});
    const double2 u2 = mul_mi(double2{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y});
    v[0] = cadd(a, cadd(b1, b2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
}

// One Stockham stage of radix R over N points: cur -> nxt, span Ns before the stage.
template <int R>
__device__ inline void stockham_stage(const double2* cur, double2* nxt, int N, int Ns,
                                      const double2* __restrict__ tw) {
  const int nb = N / R, step = N / (Ns * R);
  for (int j = threadIdx.x; j < nb; j += kFftThreads) {
    const int k = j % Ns;
    double2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      v[q] = cur[j + q * nb];
      if (q) v[q] = cmul(v[q], tw[q * k * step]);
    }
    dft<R>(v);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int q = 0; q < R; ++q) nxt[base + q * Ns] = v[q];
  }
}

// Forward FFT of buf[0 .. N) in place (through buf[N .. 2N)); returns the buffer that
// holds the result.
__device__ double2* fft(double2* buf, const FftShape& f, const double2* __restrict__ tw) {
  double2* cur = buf;
  double2* nxt = buf + f.N;
  int Ns = 1;
  for (int s = 0; s < f.stages; ++s) {
    const int R = f.radix[s];
    if (R == 4) stockham_stage<4>(cur, nxt, f.N, Ns, tw);
    else if (R == 2) stockham_stage<2>(cur, nxt, f.N, Ns, tw);
    else if (R == 3) stockham_stage<3>(cur, nxt, f.N, Ns, tw);
    else stockham_stage<5>(cur, nxt, f.N, Ns, tw);
    __syncthreads();
    double2* t = cur;
    cur = nxt;
    nxt = t;
    Ns *= R;
  }
  return cur;
}

__device__ inline double dct_scale(int k, int N) {
  return k == 0 ? sqrt(1.0 / double(N)) : sqrt(2.0 / double(N));
}

// Forward orthonormal DCT-II of each row of in (rows x N); out (N x rows) = transposed.
// With lam_other (the other dimension's 2 cos table), the coefficient of (row r, index k)
// is divided by lambda = tab_lam[k] + lam_other[r] - 4, and (0, 0) is set to 0.
__global__ __launch_bounds__(kFftThreads) void k_dct2_rows(const double* __restrict__ in,
                                                           double* __restrict__ out, int rows,
                                                           FftShape f,
                                                           const double* __restrict__ tab,
                                                           const double* __restrict__ lam_other) {
  extern __shared__ double2 fbuf[];
  const int N = f.N, r = blockIdx.x;
  const double2* tw = reinterpret_cast<const double2*>(tab);
  const double2* tq = reinterpret_cast<const double2*>(tab + 2 * N);
  const double* lam = tab + 4 * N;
  const double* x = in + size_t(r) * N;
  const int half = (N + 1) / 2;
  for (int i = threadIdx.x; i < N; i += kFftThreads) {
    const int src = i < half ? 2 * i : 2 * (N - 1 - i) + 1;
    fbuf[i] = double2{x[src], 0.0};
  }
  __syncthreads();
  const double2* V = fft(fbuf, f, tw);
  for (int k = threadIdx.x; k < N; k += kFftThreads) {
    const double2 w = tq[k], v = V[k];
    double X = dct_scale(k, N) * (w.x * v.x - w.y * v.y);
    if (lam_other) {
      if (k == 0 && r == 0) X = 0.0;
      else X = X / (lam[k] + lam_other[r] - 4.0);
    }
    out[size_t(k) * rows + r] = X;
  }
}

// Inverse orthonormal DCT (DCT-III) of each row of in (rows x N); out (N x rows).
__global__ __launch_bounds__(kFftThreads) void k_dct3_rows(const double* __restrict__ in,
                                                           double* __restrict__ out, int rows,
                                                           FftShape f,
                                                           const double* __restrict__ tab) {
  extern __shared__ double2 fbuf[];
  const int N = f.N, r = blockIdx.x;
  const double2* tw = reinterpret_cast<const double2*>(tab);
  const double2* tq = reinterpret_cast<const double2*>(tab + 2 * N);
  const double* X = in + size_t(r) * N;
  // V_k = exp(i pi k / 2N) (S_k - i S_{N-k}), S = X / s (S_N = 0); the inverse DFT of V
  // is the conjugate of the forward DFT of conj(V), over N
  for (int k = threadIdx.x; k < N; k += kFftThreads) {
    const double sk = X[k] / dct_scale(k, N);
    const double snk = k == 0 ? 0.0 : X[N - k] / dct_scale(N - k, N);
    const double2 w = double2{tq[k].x, -tq[k].y};  // exp(+i pi k / 2N)
    const double2 v = cmul(w, double2{sk, -snk});
    fbuf[k] = double2{v.x, -v.y};
  }
  __syncthreads();
  const double2* y = fft(fbuf, f, tw);
  const int half = (N + 1) / 2;
  const double inv = 1.0 / double(N);
  for (int i = threadIdx.x; i < N; i += kFftThreads) {
    const int dst = i < half ? 2 * i : 2 * (N - 1 - i) + 1;
    out[size_t(dst) * rows + r] = y[i].x * inv;
  }
}

// Table of one dimension of the DST-I of length L (FFT length N = 2 (L + 1), 4 (L + 1) + L
// doubles): exp(-2 pi i t / N) for t < N, then 2 cos(pi (k + 1) / (L + 1)) for k < L.
size_t dst_table_doubles(int L) { return size_t(4) * (L + 1) + L; }

bool dst_factor(int L, FftShape& f) {
  return L >= 1 && L < kFftMaxN / 2 && factor(2 * (L + 1), f);
}

// Orthonormal DST-I of each row of in (rows x L), out (L x rows) = transposed.  Block b
// transforms rows 2b and 2b + 1 as the real and imaginary parts of one FFT of the odd
// extension z = (0, x_0 .. x_{L-1}, 0, -x_{L-1} .. -x_0): a real row's transform is
// Z_{k+1} = -2i X_k, so X_a[k] = -Im Z_{k+1} / 2 and X_b[k] = Re Z_{k+1} / 2.  With
// lam_other (the other dimension's 2 cos table), the coefficient of (row r, index k) is
// multiplied by h2 / (tab_lam[k] + lam_other[r] - 4), the inverse eigenvalue.
__global__ __launch_bounds__(kFftThreads) void k_dst1_rows(const double* __restrict__ in,
                                                           double* __restrict__ out, int rows,
                                                           int L, FftShape f,
                                                           const double* __restrict__ tab,
                                                           const double* __restrict__ lam_other,
                                                           double h2) {
  extern __shared__ double2 fbuf[];
  const int N = f.N;
  const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
  const bool two = r1 < rows;
  const double2* tw = reinterpret_cast<const double2*>(tab);
  const double* lam = tab + 2 * N;
  const double* xa = in + size_t(r0) * L;
  const double* xb = in + size_t(two ? r1 : r0) * L;
  for (int t = threadIdx.x; t < N; t += kFftThreads) {
    double2 v{0.0, 0.0};
    if (t >= 1 && t <= L) {
      v.x = xa[t - 1];
      v.y = two ? xb[t - 1] : 0.0;
    } else if (t >= L + 2) {
      const int j = N - 1 - t;
      v.x = -xa[j];
      v.y = two ? -xb[j] : 0.0;
    }
    fbuf[t] = v;
  }
  __syncthreads();
  const double2* Z = fft(fbuf, f, tw);
  const double sc = 0.5 * sqrt(2.0 / double(L + 1));
  for (int k = threadIdx.x; k < L; k += kFftThreads) {
    const double2 z = Z[k + 1];
    double a = -sc * z.y, b = sc * z.x;
    if (lam_other) {
      a = a * h2 / (lam[k] + lam_other[r0] - 4.0);
      if (two) b = b * h2 / (lam[k] + lam_other[r1] - 4.0);
    }
    out[size_t(k) * rows + r0] = a;
    if (two) out[size_t(k) * rows + r1] = b;
  }
}

void fill_dst_table(int L, double* t) {
  const double pi = 3.14159265358979323846;
  const int N = 2 * (L + 1);
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * pi * double(k) / double(N);
    t[2 * k] = std::cos(a);
    t[2 * k + 1] = std::sin(a);
  }
  for (int k = 0; k < L; ++k) t[2 * N + k] = 2.0 * std::cos(pi * double(k + 1) / double(L + 1));
}

void fill_table(int N, double* t) {
  const double pi = 3.14159265358979323846;
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * pi * double(k) / double(N);
    t[2 * k] = std::cos(a);
    t[2 * k + 1] = std::sin(a);
    const double q = -pi * double(k) / (2.0 * double(N));
    t[2 * N + 2 * k] = std::cos(q);
    t[2 * N + 2 * k + 1] = std::sin(q);
    t[4 * N + k] = 2.0 * std::cos(pi * double(k) / double(N));
  }
}

size_t lds_bytes(int N) { return size_t(2) * N * sizeof(double2); }

hipError_t dct_lds_setup() {
  static bool lds_set = false;
  if (!lds_set) {
    hipError_t e = hipFuncSetAttribute((const void*)k_dct2_rows,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(lds_bytes(kFftMaxN)));
    if (e == hipSuccess)
      e = hipFuncSetAttribute((const void*)k_dct3_rows,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(lds_bytes(kFftMaxN)));
    if (e != hipSuccess) return e;
    lds_set = true;
  }
  return hipSuccess;
}

// The four row passes of the exact Neumann solve; work holds 2 m n doubles.
hipError_t dct_solve_passes(const double* b, double* u, int m, int n, const FftShape& fm,
                            const FftShape& fn, const double* tables, double* work,
                            hipStream_t s) {
  const double* tm = tables;
  const double* tn = tables + table_doubles(m);
  double* w0 = work;                       // n x m
  double* w1 = work + size_t(m) * n;       // m x n
  // along n: b (m x n) -> w0 (n x m); along m, divided by lambda: w0 -> w1 (m x n)
  k_dct2_rows<<<m, kFftThreads, lds_bytes(n), s>>>(b, w0, m, fn, tn, nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_dct2_rows<<<n, kFftThreads, lds_bytes(m), s>>>(w0, w1, n, fm, tm, tn + 4 * size_t(n));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // inverse along n: w1 (m x n) -> w0 (n x m); along m: w0 -> u (m x n)
  k_dct3_rows<<<m, kFftThreads, lds_bytes(n), s>>>(w1, w0, m, fn, tn);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_dct3_rows<<<n, kFftThreads, lds_bytes(m), s>>>(w0, u, n, fm, tm);
  return hipGetLastError();
}

constexpr int kMaxDevices = 64;

// Device address of this device's scratch array, looked up once per device (with that
// device current) and kept.
hipError_t static_work(double** out) {
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) return e;
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return hipSuccess;
}

}  // namespace
}  // namespace gdp

using namespace gdp;

extern "C" {

int gdp_dct_supported(int m, int n) {
  FftShape a, b;
  return factor(m, a) && factor(n, b) ? 1 : 0;
}

int gdp_dct_tables_bytes(int m, int n, size_t* bytes) {
  FftShape a, b;
  if (!bytes || !factor(m, a) || !factor(n, b)) {
    set_error("gdp_dct: sides must be at most %d with prime factors 2, 3, 5 only (got %d x %d)",
              kFftMaxN, m, n);
    return GDP_ENOTSUP;
  }
  *bytes = sizeof(double) * (table_doubles(m) + table_doubles(n));
  return GDP_OK;
}

int gdp_dct_tables(int m, int n, double* tables, void* stream) {
  size_t bytes = 0;
  if (int rc = gdp_dct_tables_bytes(m, n, &bytes)) return rc;
  if (!tables) return (set_error("null tables"), GDP_EINVAL);
  std::vector<double> h(bytes / sizeof(double));
  fill_table(m, h.data());
  fill_table(n, h.data() + table_doubles(m));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  GDP_FFT_HIP(hipMemcpyAsync(tables, h.data(), bytes, hipMemcpyHostToDevice, s));
  GDP_FFT_HIP(hipStreamSynchronize(s));  // h is released on return
  return GDP_OK;
}

int gdp_dct_work_bytes(int m, int n, size_t* bytes) {
  if (!bytes || m < 1 || n < 1) return (set_error("bad arguments"), GDP_EINVAL);
  *bytes = sizeof(double) * 2 * size_t(m) * n;
  return GDP_OK;
}

int gdp_neumann_dct_solve(const double* b, double* u, int m, int n, const double* tables,
                          double* work, void* stream) {
  FftShape fm, fn;
  if (!b || !u || !tables || !work) return (set_error("null argument"), GDP_EINVAL);
  if (!factor(m, fm) || !factor(n, fn)) {
    set_error("gdp_dct: sides must be at most %d with prime factors 2, 3, 5 only (got %d x %d)",
              kFftMaxN, m, n);
    return GDP_ENOTSUP;
  }
  GDP_FFT_HIP(dct_lds_setup());
  GDP_FFT_HIP(dct_solve_passes(b, u, m, n, fm, fn, tables, work,
                               static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_neumann_dct_solve_static(const double* b, double* u, int m, int n,
                                 const double* tables, void* stream) {
  FftShape fm, fn;
  if (!b || !u || !tables) return (set_error("null argument"), GDP_EINVAL);
  if (!factor(m, fm) || !factor(n, fn)) {
    set_error("gdp_dct: sides must be at most %d with prime factors 2, 3, 5 only (got %d x %d)",
              kFftMaxN, m, n);
    return GDP_ENOTSUP;
  }
  if (2 * size_t(m) * n > kStaticWorkDoubles) {
    set_error("gdp_neumann_dct_solve_static: %d x %d needs %zu bytes of work, more than the "
              "static %zu (m n <= %zu); use gdp_neumann_dct_solve with a work buffer",
              m, n, 2 * size_t(m) * n * sizeof(double), kStaticWorkDoubles * sizeof(double),
              kStaticWorkDoubles / 2);
    return GDP_ENOMEM;
  }
  double* work = nullptr;
  GDP_FFT_HIP(static_work(&work));
  GDP_FFT_HIP(dct_lds_setup());
  GDP_FFT_HIP(dct_solve_passes(b, u, m, n, fm, fn, tables, work,
                               static_cast<hipStream_t>(stream)));
  return GDP_OK;
}

int gdp_dst_supported(int m, int n) {
  FftShape a, b;
  return dst_factor(m, a) && dst_factor(n, b) ? 1 : 0;
}

int gdp_dst_tables_bytes(int m, int n, size_t* bytes) {
  FftShape a, b;
  if (!bytes || !dst_factor(m, a) || !dst_factor(n, b)) {
    set_error("gdp_dst: sides m need 2 (m + 1) <= %d with prime factors 2, 3, 5 only "
              "(got %d x %d)", kFftMaxN, m, n);
    return GDP_ENOTSUP;
  }
  *bytes = sizeof(double) * (dst_table_doubles(m) + dst_table_doubles(n));
  return GDP_OK;
}

int gdp_dst_tables(int m, int n, double* tables, void* stream) {
  size_t bytes = 0;
  if (int rc = gdp_dst_tables_bytes(m, n, &bytes)) return rc;
  if (!tables) return (set_error("null tables"), GDP_EINVAL);
  std::vector<double> h(bytes / sizeof(double));
  fill_dst_table(m, h.data());
  fill_dst_table(n, h.data() + dst_table_doubles(m));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  GDP_FFT_HIP(hipMemcpyAsync(tables, h.data(), bytes, hipMemcpyHostToDevice, s));
  GDP_FFT_HIP(hipStreamSynchronize(s));  // h is released on return
  return GDP_OK;
}

int gdp_dirichlet_dst_solve(const double* b, double* u, int m, int n, double spacing,
                            const double* tables, double* work, void* stream) {
  FftShape fm, fn;
  if (!b || !u || !tables || !work || !(spacing > 0.0))
    return (set_error("null argument or spacing <= 0"), GDP_EINVAL);
  if (!dst_factor(m, fm) || !dst_factor(n, fn)) {
    set_error("gdp_dst: sides m need 2 (m + 1) <= %d with prime factors 2, 3, 5 only "
              "(got %d x %d)", kFftMaxN, m, n);
    return GDP_ENOTSUP;
  }
  static bool lds_set = false;
  if (!lds_set) {
    GDP_FFT_HIP(hipFuncSetAttribute((const void*)k_dst1_rows,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds_bytes(kFftMaxN))));
    lds_set = true;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const double* tm = tables;
  const double* tn = tables + dst_table_doubles(m);
  const double* lam_n = tn + 2 * size_t(fn.N);
  const double h2 = spacing * spacing;
  double* w0 = work;                  // n x m
  double* w1 = work + size_t(m) * n;  // m x n
  // S along n: b (m x n) -> w0 (n x m); S along m and 1 / lambda: w0 -> w1 (m x n)
  k_dst1_rows<<<(m + 1) / 2, kFftThreads, lds_bytes(fn.N), s>>>(b, w0, m, n, fn, tn, nullptr,
                                                                 h2);
  GDP_FFT_HIP(hipGetLastError());
  k_dst1_rows<<<(n + 1) / 2, kFftThreads, lds_bytes(fm.N), s>>>(w0, w1, n, m, fm, tm, lam_n,
                                                                 h2);
  GDP_FFT_HIP(hipGetLastError());
  // S along n: w1 -> w0 (n x m); S along m: w0 -> u (m x n)
  k_dst1_rows<<<(m + 1) / 2, kFftThreads, lds_bytes(fn.N), s>>>(w1, w0, m, n, fn, tn, nullptr,
                                                                 h2);
  GDP_FFT_HIP(hipGetLastError());
  k_dst1_rows<<<(n + 1) / 2, kFftThreads, lds_bytes(fm.N), s>>>(w0, u, n, m, fm, tm, nullptr,
                                                                 h2);
  GDP_FFT_HIP(hipGetLastError());
  return GDP_OK;
}

}  // extern "C"
