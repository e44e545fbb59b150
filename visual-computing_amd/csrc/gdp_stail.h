// The coarse tail in registers: every level of one problem from the first one that is at
// most 128 columns wide and 64 (FP32) or 32 (FP64) rows tall down to the coarsest, in
// one block of 8 waves
// (512 threads, so each wave may hold up to 256 registers: the tail keeps every level's
// iterate and right-hand side in registers across the whole V-cycle).
//
// Each level is laid out like a strip kernel's (gdp_strip.h): wave w holds rows
// [w R, w R + R) of the level (R = kRows0 >> depth, at least 1) and lane l the column
// pair (2l, 2l+1) of each row, in registers.  Vertical neighbours across wave
// boundaries come through a small LDS halo slot, double-buffered so each red-black
// half-sweep costs one barrier; horizontal neighbours come from DPP wave shifts.
// Levels hand data to each other through one LDS buffer of rows 128 wide: the fine
// residual on the way down (each wave reads the fine rows its coarse rows restrict
// from), the coarse iterate on the way up.  The coarsest level is solved in FP64: by
// the exact DCT / DST transform in LDS, or by coarsest_sweeps red-black sweeps.
//
// Arithmetic follows gradient-domain/src/gradient_domain/multigrid.py (v_cycle
// :184-215) exactly as the strip kernels do; the exact coarsest solves follow
// solvers.py:224-262 (Neumann) and its DST-I analogue (Dirichlet).
#pragma once
#include "gdp_dct.h"
#include "gdp_internal.h"
#include "gdp_strip.h"

namespace gdp {

constexpr int kSTailThreads = 512;
constexpr int kSTailWaves = kSTailThreads / 64;  // 8
constexpr int kSTailMaxLevels = 5;
constexpr int kSTailCols = 128;                 // columns of a level (64 lanes x 2)
// Exchange-buffer rows: 132 elements, column j of the level at j + 1, so the ghost
// column -1 sits at 0 and the ghost column n at n + 1; 132 elements keep every row
// start 16-byte aligned for the quad reads of the restriction.
constexpr int kXbRow = 132;
constexpr int kSTailXbBytes = 65536;            // level exchange buffer
constexpr int kSTailXhBytes = 2 * kSTailWaves * 2 * 64 * 16;  // halo slots (FP64 pairs)

template <typename T>
struct STailRows {
  // rows per wave of the first level: 64 rows in all (FP32) or 32 (FP64)
  static constexpr int kRows0 = sizeof(T) == 4 ? 8 : 4;
};

template <int RD0, int D>
struct RowsAt {
  static constexpr int value = (RD0 >> D) > 0 ? (RD0 >> D) : 1;
};

// What the level code reads of a tail level's descriptor.  The sizes of every level are
// loaded at kernel entry in one memory round trip (read level by level from global
// memory, each level's first access would wait for a miss of its own under the
// recursion); the spacing doubles from level to level, so h^2 and 1/h^2 of level D are
// the first level's times 4^D and 4^-D, exactly.
struct STailGeo {
  int m, n;
  double h2, inv_h2;
};

struct STailCtx {
  unsigned char* xb;   // level exchange rows (kXbRow elements of T per row)
  unsigned char* xh;   // halo slots [2][16 waves][2 rows][64 lanes] pairs
  double* tab;         // exact coarsest: Cm, Cn, 1/lambda
  double* tabT;        // exact coarsest: Cn transposed
  double* tabTm;       // exact coarsest: Cm transposed
  double* w0;          // exact coarsest work arrays
  double* w1;
  int gmn[kSTailMaxLevels];  // level sizes m | n << 16 (scalar registers; sides <= 128)
  double h2_0, inv_h2_0;                         // spacing of the first tail level
  int nlev, w, lane, pre, post, csweeps, order, hbuf;
  bool exact;
};

// What the final stores of the kernel use of the first tail level and of the Krylov
// vectors, written to LDS at entry.
struct STailEnd {
  void* u;
  double* kr;
  double* kq;
  int pitch;
};

template <int D>
__device__ __attribute__((always_inline)) STailGeo st_geo(const STailCtx& c) {
  constexpr double s = double(1ull << (2 * D));
  return STailGeo{c.gmn[D] & 0xffff, c.gmn[D] >> 16, c.h2_0 * s, c.inv_h2_0 / s};
}

// Halo exchange: publish this wave's first and last rows, return the row above the
// wave's first row and the row below its last (zero at the ends of the block).
template <typename V2>
__device__ __attribute__((always_inline)) void st_halo(STailCtx& c, const V2& top, const V2& bot, V2& above, V2& below) {
  V2* h = reinterpret_cast<V2*>(c.xh) + c.hbuf * (kSTailWaves * 2 * 64);
  h[(c.w * 2 + 0) * 64 + c.lane] = top;
  h[(c.w * 2 + 1) * 64 + c.lane] = bot;
  __syncthreads();
  above = c.w > 0 ? h[((c.w - 1) * 2 + 1) * 64 + c.lane] : V2{0, 0};
  below = c.w < kSTailWaves - 1 ? h[((c.w + 1) * 2 + 0) * 64 + c.lane] : V2{0, 0};
  c.hbuf ^= 1;
}

// Red-black half-sweep of colour Q over a level's rows held by this wave.  wi holds
// each row's inverse diagonals (zero outside the grid, so rows past the grid stay zero
// without a test).  With an even row count per wave a row's parity is its index's, so
// the choice of x or y is fixed at compile time.
template <typename T, int RD, int Q>
__device__ __attribute__((always_inline)) void st_halfsweep_q(STailCtx& c, typename Pair<T>::type* u,
                                                            const typename Pair<T>::type* hb,
                                                            const typename Pair<T>::type* wi) {
  using V2 = typename Pair<T>::type;
  V2 above, below;
  st_halo(c, u[0], u[RD - 1], above, below);
#pragma unroll
  for (int r = 0; r < RD; ++r) {
    const V2 up = r > 0 ? u[r - 1] : above;
    const V2 dn = r < RD - 1 ? u[r + 1] : below;
    const bool ux = (RD % 2 == 0) ? (((r + Q) & 1) == 0) : (((c.w * RD + r + Q) & 1) == 0);
    if (ux) {
      const T nb = ((up.x + dn.x) + dpp_shr1(u[r].y)) + u[r].y;
      u[r].x = wi[r].x * nb - hb[r].x;
    } else {
      const T nb = ((up.y + dn.y) + u[r].x) + dpp_shl1(u[r].x);
      u[r].y = wi[r].y * nb - hb[r].y;
    }
  }
}

template <typename T, int RD>
__device__ __attribute__((always_inline)) void st_halfsweep(STailCtx& c, typename Pair<T>::type* u,
                                                          const typename Pair<T>::type* hb,
                                                          const typename Pair<T>::type* wi, int q) {
  if (q) st_halfsweep_q<T, RD, 1>(c, u, hb, wi);
  else st_halfsweep_q<T, RD, 0>(c, u, hb, wi);
}

// Inverse diagonals wi of this wave's rows and hb = h^2 b / diag (zero outside the
// grid).
template <typename T, int RD, bool NEU>
__device__ __attribute__((always_inline)) void st_scale_rhs(const STailCtx& c, typename Pair<T>::type* hb,
                                                          typename Pair<T>::type* wi,
                                                          const LaneCols<T>& lc, int m, T h2) {
#pragma unroll
  for (int r = 0; r < RD; ++r) {
    const int gi = c.w * RD + r;
    const int vc = row_count(gi, m, NEU);
    const bool rin = gi < m;
    wi[r].x = (rin && lc.inx) ? lc.wx(vc) : T(0);
    wi[r].y = (rin && lc.iny) ? lc.wy(vc) : T(0);
    hb[r].x = wi[r].x * h2 * hb[r].x;
    hb[r].y = wi[r].y * h2 * hb[r].y;
  }
}

// Ghost columns written with a published level (column -1 and column n):
enum : int {
  kGhostZero = 0,     // zero (Dirichlet iterate; residuals never read them)
  kGhostIter = 1,     // replicate the edge columns (Neumann iterate: the prolongation
                      // reads edge-replicated coarse values)
  kGhostResidual = 2  // the Neumann residual extension: r[-1] = r[0], r[n] = r[n-1] for
                      // even n (zero for odd n)
};

// Publish this wave's rows of a level (n columns) to the exchange buffer with their
// ghost columns, then synchronise.
template <typename T, int RD, int ghost>
__device__ __attribute__((always_inline)) void st_publish(const STailCtx& c, const typename Pair<T>::type* v,
                                                        int m, int n) {
  using V2 = typename Pair<T>::type;
  T* xb = reinterpret_cast<T*>(c.xb);
  const int gjx = 2 * c.lane;
#pragma unroll
  for (int r = 0; r < RD; ++r) {
    const int gi = c.w * RD + r;
    if (gi >= m) continue;
    V2 val = v[r];
    const T sy = dpp_shr1(val.y);  // column 2l - 1
    if constexpr (ghost != kGhostZero) {
      if (gjx == n && (ghost == kGhostIter || (n & 1) == 0)) val.x = sy;
      if (gjx + 1 == n && ghost == kGhostIter) val.y = val.x;
    }
    T* row = xb + gi * kXbRow;
    row[gjx + 1] = val.x;
    row[gjx + 2] = val.y;
    if (c.lane == 0) row[0] = ghost == kGhostZero ? T(0) : val.x;
    if (c.lane == 63) {
      // exchange columns 129 .. 131 (grid columns 128 .. 130), past every lane's pair:
      // the ghost column n of a 128-wide level (its edge value, like the ghosts above)
      // and zeros, which the Neumann restriction of a 127- or 128-wide level reads
      row[129] = (ghost != kGhostZero && n == kSTailCols) ? val.y : T(0);
      row[130] = T(0);
      row[131] = T(0);
    }
  }
  __syncthreads();
}

// tab[i] where ok, zero elsewhere (the address clamped to the table's first element).
__device__ __attribute__((always_inline)) double lds_or_zero(const double* tab, int i, bool ok) {
  const double v = tab[ok ? i : 0];
  return ok ? v : 0.0;
}

// The exact coarsest solve of a level of at most 16 x 16 points by wave 0 on the FP64
// matrix cores, published to the exchange buffer (the caller synchronises).  It computes
// what lds_dct_solve computes, U = Cm^T ((Cm W Cn^T) ./ lambda) Cn with W = h^2 b in w0,
// every matrix zero-padded to 16 x 16.  v_mfma_f64_16x16x4f64 takes A[i][k] from lane
// 16 k + i and B[k][j] from lane 16 k + j, and leaves D[g + 4 v][j] in item v of lane
// 16 g + j; so a result X serves the next product as the left operand X^T with k-block v
// taken from item v, and each stage is X^T M for the stage's table M: W^T Cm^T =
// (Cm W)^T, then (Cm W) Cn^T, then (Cm^T Y)^T for the scaled Y, then Cm^T Y Cn.
template <typename T, bool NEU>
__device__ __attribute__((always_inline)) void st_coarsest_mfma(const STailCtx& c, int m, int n) {
  typedef double D4 __attribute__((ext_vector_type(4)));
  const int g = c.lane >> 4, cc = c.lane & 15;
  const double* Cm = c.tab;
  const double* Cn = c.tab + m * m;
  const double* il = Cn + n * n;
  // every operand of the four stages, loaded before the first product
  double a0[4], b1[4], b2[4], b3[4], b4[4], sc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int K = g + 4 * k;
    a0[k] = lds_or_zero(c.w0, K * n + cc, K < m && cc < n);        // W[K][cc]
    b1[k] = lds_or_zero(c.tabTm, K * m + cc, K < m && cc < m);     // Cm^T[K][cc]
    b2[k] = lds_or_zero(c.tabT, K * n + cc, K < n && cc < n);      // Cn^T[K][cc]
    b3[k] = lds_or_zero(Cm, K * m + cc, K < m && cc < m);          // Cm[K][cc]
    b4[k] = lds_or_zero(Cn, K * n + cc, K < n && cc < n);          // Cn[K][cc]
    sc[k] = lds_or_zero(il, K * n + cc, K < m && cc < n);          // 1/lambda[K][cc]
  }
  D4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) x = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[k], b1[k], x, 0, 0, 0);
  D4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) y = __builtin_amdgcn_mfma_f64_16x16x4f64(x[k], b2[k], y, 0, 0, 0);
#pragma unroll
  for (int k = 0; k < 4; ++k) y[k] *= sc[k];
  x = D4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) x = __builtin_amdgcn_mfma_f64_16x16x4f64(y[k], b3[k], x, 0, 0, 0);
  y = D4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k) y = __builtin_amdgcn_mfma_f64_16x16x4f64(x[k], b4[k], y, 0, 0, 0);
  // publish rows g + 4 v (exchange column j + 1 holds column j, zero past the level)
  // with the ghost columns -1 and n: the edge values for Neumann, zero for Dirichlet.
  // Lane j = n (n < 16) writes ghost n from lane n - 1 of its group; lane 15 writes it
  // when n = 16.
  T* xb = reinterpret_cast<T*>(c.xb);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int row = g + 4 * v;
    const T val = T(y[v]);
    const T left = dpp_shr1(val);
    if (row < m) {
      T* xr = xb + row * kXbRow;
      xr[cc + 1] = cc < n ? val : ((NEU && cc == n) ? left : T(0));
      if (cc == 15 && n == 16) xr[17] = NEU ? val : T(0);
      if (cc == 0) xr[0] = NEU ? val : T(0);
    }
  }
}

// Coarsest level: u from the right-hand side b (plain) of this wave's rows; the
// result is published to the exchange buffer.  EXACT: 1 the transform solve, 0 the FP64
// sweeps, -1 whichever the problem has (c.exact).
template <typename T, int RD, bool NEU, int EXACT>
__device__ __attribute__((always_inline)) void st_coarsest(STailCtx& c, const STailGeo& L, typename Pair<T>::type* b) {
  using V2 = typename Pair<T>::type;
  using D2 = typename Pair<double>::type;
  const int m = L.m, n = L.n;
  V2 u[RD];
  if (EXACT == 1 || (EXACT < 0 && c.exact)) {
    const int gj = 2 * c.lane;
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int gi = c.w * RD + r;
      if (gi < m) {
        if (gj < n) c.w0[gi * n + gj] = double(b[r].x) * L.h2;
        if (gj + 1 < n) c.w0[gi * n + gj + 1] = double(b[r].y) * L.h2;
      }
    }
    __syncthreads();
    if (m <= 16 && n <= 16) {
      if (c.w == 0) st_coarsest_mfma<T, NEU>(c, m, n);
      __syncthreads();
      return;
    }
    lds_dct_solve(c.w0, c.w1, m, n, c.tab, kSTailThreads, c.tabT);
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int gi = c.w * RD + r;
      u[r] = V2{T(0), T(0)};
      if (gi < m) {
        if (gj < n) u[r].x = T(c.w0[gi * n + gj]);
        if (gj + 1 < n) u[r].y = T(c.w0[gi * n + gj + 1]);
      }
    }
  } else {
    // coarsest_sweeps red-black sweeps in FP64 from zero (multigrid.py:206-207)
    LaneCols<double> lc;
    lc.init(0, c.lane, n, NEU);
    D2 ud[RD], hd[RD];
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      ud[r] = D2{0.0, 0.0};
      hd[r] = D2{double(b[r].x), double(b[r].y)};
    }
    D2 wd[RD];
    st_scale_rhs<double, RD, NEU>(c, hd, wd, lc, m, L.h2);
    for (int s = 0; s < c.csweeps; ++s) {
      st_halfsweep_q<double, RD, 0>(c, ud, hd, wd);
      st_halfsweep_q<double, RD, 1>(c, ud, hd, wd);
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) u[r] = V2{T(ud[r].x), T(ud[r].y)};
  }
  st_publish<T, RD, NEU ? kGhostIter : kGhostZero>(c, u, m, n);
}

// Fine residual (scaled by h^2) of this wave's rows to the exchange buffer.
template <typename T, int RD, bool NEU>
__device__ __attribute__((always_inline)) void st_residual(STailCtx& c, const typename Pair<T>::type* u,
                            const typename Pair<T>::type* hb, const LaneCols<T>& lc, int m,
                            int lc_n) {
  using V2 = typename Pair<T>::type;
  V2 above, below;
  st_halo(c, u[0], u[RD - 1], above, below);
  V2 r2[RD];
#pragma unroll
  for (int r = 0; r < RD; ++r) {
    const int gi = c.w * RD + r;
    const int vc = row_count(gi, m, NEU);
    const V2 up = r > 0 ? u[r - 1] : above;
    const V2 dn = r < RD - 1 ? u[r + 1] : below;
    const T nx = ((up.x + dn.x) + dpp_shr1(u[r].y)) + u[r].y;
    const T ny = ((up.y + dn.y) + u[r].x) + dpp_shl1(u[r].x);
    const T rx = (T(vc) + lc.hx) * (hb[r].x + u[r].x) - nx;
    const T ry = (T(vc) + lc.hy) * (hb[r].y + u[r].y) - ny;
    r2[r].x = (gi < m && lc.inx) ? rx : T(0);
    r2[r].y = (gi < m && lc.iny) ? ry : T(0);
  }
  st_publish<T, RD, NEU ? kGhostResidual : kGhostZero>(c, r2, m, lc_n);
}

// Four consecutive elements of an exchange row at a 16-byte aligned position.
template <typename T>
struct Quad {
  T a, b, c, d;
};
__device__ inline Quad<float> ld_quad(const float* p) {
  const float4 q = *reinterpret_cast<const float4*>(p);
  return Quad<float>{q.x, q.y, q.z, q.w};
}
__device__ inline Quad<double> ld_quad(const double* p) {
  const double2 q0 = *reinterpret_cast<const double2*>(p);
  const double2 q1 = *reinterpret_cast<const double2*>(p + 2);
  return Quad<double>{q0.x, q0.y, q1.x, q1.y};
}

// Restriction of the published fine residual (scaled by h^2, with its ghost columns) to
// this wave's coarse rows: the plain right-hand side 0.25 P^T r P / h^2 of the coarse
// level.  Coarse lane l (columns 2l, 2l+1) reads the quad of fine columns 4l-1 .. 4l+2
// (exchange columns 4l .. 4l+3) and the next lane's first two.  Every quad of the
// wave's rows is loaded before any arithmetic, so the LDS latency is paid once.
template <typename T, int RC, bool NEU>
__device__ __attribute__((always_inline)) void st_restrict(const STailCtx& c, const STailGeo& F,
                                                         const STailGeo& C,
                                                         typename Pair<T>::type* cb) {
  constexpr int NT = NEU ? 4 : 3;  // vertical taps
  const T* R = reinterpret_cast<const T*>(c.xb);
  const int m = F.m, mc = C.m, nc = C.n;
  const T s = T(0.25) * T(F.inv_h2);
  const int l = c.lane;
  const bool v0ok = 2 * l < nc, v1ok = 2 * l + 1 < nc;
  Quad<T> q[RC][NT];
  T wr[RC][NT];
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int I = c.w * RC + r;
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      // fine row and weight of tap k (Neumann rows extended like columns)
      int i;
      T wk;
      if constexpr (NEU) {
        i = 2 * I - 1 + k;
        wk = (k == 0 || k == 3) ? T(0.25) : T(0.75);
        if (i == -1) i = 0;
        if (i == m && (m & 1) == 0) i = m - 1;
        if (i >= m || I >= mc) {
          i = 0;
          wk = T(0);
        }
      } else {
        i = I < mc ? 2 * I + k : 0;
        wk = I < mc ? (k == 1 ? T(1) : T(0.5)) : T(0);
      }
      wr[r][k] = wk;
      q[r][k] = ld_quad(R + i * kXbRow + 4 * l);
    }
  }
#pragma unroll
  for (int r = 0; r < RC; ++r) {
    const int I = c.w * RC + r;
    // vertical pass on exchange columns 4l .. 4l+5
    T v[6];
    v[0] = wr[r][0] * q[r][0].a;
    v[1] = wr[r][0] * q[r][0].b;
    v[2] = wr[r][0] * q[r][0].c;
    v[3] = wr[r][0] * q[r][0].d;
#pragma unroll
    for (int k = 1; k < NT; ++k) {
      v[0] += wr[r][k] * q[r][k].a;
      v[1] += wr[r][k] * q[r][k].b;
      v[2] += wr[r][k] * q[r][k].c;
      v[3] += wr[r][k] * q[r][k].d;
    }
    v[4] = dpp_shl1(v[0]);
    v[5] = dpp_shl1(v[1]);
    T v0, v1;
    if constexpr (NEU) {
      // fine columns 4l-1 .. 4l+2 and 4l+1 .. 4l+4
      v0 = T(0.25) * v[0] + T(0.75) * v[1] + T(0.75) * v[2] + T(0.25) * v[3];
      v1 = T(0.25) * v[2] + T(0.75) * v[3] + T(0.75) * v[4] + T(0.25) * v[5];
    } else {
      // fine columns 4l .. 4l+2 and 4l+2 .. 4l+4
      v0 = T(0.5) * v[1] + v[2] + T(0.5) * v[3];
      v1 = T(0.5) * v[3] + v[4] + T(0.5) * v[5];
    }
    cb[r].x = (I < mc && v0ok) ? s * v0 : T(0);
    cb[r].y = (I < mc && v1ok) ? s * v1 : T(0);
  }
}

// u += P e for this wave's rows, e = the published coarse iterate (ghost columns zero
// for Dirichlet, edge-replicated for Neumann).  Fine lane l interpolates from coarse
// columns l-1, l, l+1 (exchange columns l .. l+2).
template <typename T, int RD, bool NEU>
__device__ __attribute__((always_inline)) void st_prolong(const STailCtx& c, const STailGeo& F,
                                                        const STailGeo& C,
                                                        typename Pair<T>::type* u,
                                                        const LaneCols<T>& lc) {
  const T* E = reinterpret_cast<const T*>(c.xb) + c.lane;
  const int m = F.m, mc = C.m;
  constexpr bool neu = NEU;
  // coarse row I (clamped for Neumann, zero outside for Dirichlet), columns l-1 .. l+1
  auto row3 = [&](int I, T* t3) {
    T w = T(1);
    if (neu) {
      I = min(max(I, 0), mc - 1);
    } else if (I < 0 || I >= mc) {
      I = 0;
      w = T(0);
    }
    const T* e = E + I * kXbRow;
    t3[0] = w * e[0];
    t3[1] = w * e[1];
    t3[2] = w * e[2];
  };
  auto interp = [&](int r, const T* ta, const T* tb, T wa, T wb) {
    T t[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = wa * ta[k] + wb * tb[k];
    T cx, cy;
    if (neu) {
      cx = T(0.75) * t[1] + T(0.25) * t[0];
      cy = T(0.75) * t[1] + T(0.25) * t[2];
    } else {
      cx = T(0.5) * (t[0] + t[1]);
      cy = t[1];
    }
    if (lc.inx) u[r].x += cx;
    if (lc.iny) u[r].y += cy;
  };
  if constexpr (RD % 2 == 0) {
    // the wave's fine rows w RD .. w RD + RD - 1 read coarse rows w RD/2 - 1 .. w RD/2 + RD/2
    constexpr int NE = RD / 2 + 2;
    T ev[NE][3];
    const int I0 = c.w * (RD / 2) - 1;
#pragma unroll
    for (int k = 0; k < NE; ++k) row3(I0 + k, ev[k]);
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int gi = c.w * RD + r;
      if (gi >= m) continue;
      if (neu) {
        const int kp = r / 2 + 1, ko = (r & 1) ? kp + 1 : kp - 1;
        interp(r, ev[kp], ev[ko], T(0.75), T(0.25));
      } else if (r & 1) {
        interp(r, ev[(r + 1) / 2], ev[(r + 1) / 2], T(1), T(0));
      } else {
        interp(r, ev[r / 2], ev[r / 2 + 1], T(0.5), T(0.5));
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int gi = c.w * RD + r;
      if (gi >= m) continue;
      T ta[3], tb[3];
      if (neu) {
        const int P = gi >> 1, O = (gi & 1) ? P + 1 : P - 1;
        row3(P, ta);
        row3(O, tb);
        interp(r, ta, tb, T(0.75), T(0.25));
      } else if (gi & 1) {
        row3((gi - 1) >> 1, ta);
        interp(r, ta, ta, T(1), T(0));
      } else {
        row3((gi >> 1) - 1, ta);
        row3(gi >> 1, tb);
        interp(r, ta, tb, T(0.5), T(0.5));
      }
    }
  }
  __syncthreads();  // every wave has read the exchange buffer
}

// The code of a tail kernel for problems with NLEV tail levels (0: any, the count read
// per problem) and the EXACT coarsest method (-1: any).  A kernel fixed to one count and
// method holds one copy of the coarsest solve and no level past it: about a sixth of the
// instructions of the general kernel, which the tail's few CUs fetch cold.
template <typename T, int D, bool NEU, int NLEV, int EXACT>
__device__ __attribute__((always_inline)) void st_level(STailCtx& c, typename Pair<T>::type* b);

template <typename T, int D, bool NEU, int NLEV, int EXACT>
struct STailLevel {
  static __device__ __attribute__((always_inline)) void run(STailCtx& c, typename Pair<T>::type* b) {
    if constexpr (D < (NLEV > 0 ? NLEV : kSTailMaxLevels)) st_level<T, D, NEU, NLEV, EXACT>(c, b);
  }
};

// Level D of the tail: b holds the plain right-hand side of this wave's rows.  On
// return the level's iterate is published in the exchange buffer.  NEU: the problem's
// boundary condition (Neumann), fixed for every level of the tail.
template <typename T, int D, bool NEU, int NLEV, int EXACT>
__device__ __attribute__((always_inline)) void st_level(STailCtx& c, typename Pair<T>::type* b) {
  using V2 = typename Pair<T>::type;
  constexpr int RD = RowsAt<STailRows<T>::kRows0, D>::value;
  constexpr int RC = RowsAt<STailRows<T>::kRows0, D + 1>::value;
  const STailGeo L = st_geo<D>(c);
  if constexpr (NLEV > 0) {
    if constexpr (D == NLEV - 1) {
      st_coarsest<T, RD, NEU, EXACT>(c, L, b);
      return;
    }
  } else if (D == c.nlev - 1) {
    st_coarsest<T, RD, NEU, EXACT>(c, L, b);
    return;
  }
  const STailGeo C = st_geo<(D + 1 < kSTailMaxLevels ? D + 1 : D)>(c);
  const int m = L.m;
  LaneCols<T> lc;
  lc.init(0, c.lane, L.n, NEU);
  V2* hb = b;  // scaled in place
  V2 wi[RD];
  st_scale_rhs<T, RD, NEU>(c, hb, wi, lc, m, T(L.h2));
  // pre-smoothing from zero: the first colour-0 half-sweep is u = -hb on colour 0
  V2 u[RD];
#pragma unroll
  for (int r = 0; r < RD; ++r) {
    const int gi = c.w * RD + r;
    u[r] = V2{T(0), T(0)};
    if (c.pre > 0) {
      if ((gi & 1) == 0) u[r].x = -hb[r].x;
      else u[r].y = -hb[r].y;
    }
  }
  for (int s = 1; s < 2 * c.pre; ++s) st_halfsweep<T, RD>(c, u, hb, wi, s & 1);
  st_residual<T, RD, NEU>(c, u, hb, lc, m, L.n);
  V2 cb[RC];
  st_restrict<T, RC, NEU>(c, L, C, cb);
  STailLevel<T, D + 1, NEU, NLEV, EXACT>::run(c, cb);
  {
    // recompute the lane's column data rather than keep it live across the recursion
    int lane = c.lane;
    asm volatile("" : "+v"(lane));
    lc.init(0, lane, L.n, NEU);
  }
  st_prolong<T, RD, NEU>(c, L, C, u, lc);
  for (int s = 0; s < 2 * c.post; ++s)
    st_halfsweep<T, RD>(c, u, hb, wi, (s + c.order) & 1);
  st_publish<T, RD, NEU ? kGhostIter : kGhostZero>(c, u, m, L.n);
}

// The x half of a PCG update that k_update deferred (solvers.py:151, x += alpha d): run
// by the blocks of a tail launch past its tail blocks, on CUs the tail leaves idle.
// Block xb of nxb walks the Krylov tiles (th x tw points each) in items of one column
// pair per thread and kXItems items at a time, every load issued before the first
// wait; a problem whose xpend is zero is left alone.  Pairs are aligned: tile columns
// start at multiples of tw and the row pitch is even.
constexpr int kXItems = kStailXItems;  // one definition, shared with the host sizing
__device__ __attribute__((always_inline)) void st_xupdate(const KTile* __restrict__ tiles,
                                                        int ntiles,
                                                        const ProblemState* __restrict__ st,
                                                        int th, int tw, int xb, int nxb) {
  using D2 = typename Pair<double>::type;
  const int ppr = tw / 2;                                        // pairs per tile row
  const int cpt = (th * ppr + kSTailThreads - 1) / kSTailThreads;  // items per tile
  const int nitems = ntiles * cpt;
  for (int w0 = xb * kXItems; w0 < nitems; w0 += nxb * kXItems) {
    D2 xv[kXItems], dv[kXItems];
    double al[kXItems];
    GDP_GLOBAL double* xp[kXItems];
    size_t o[kXItems];
    bool ok[kXItems], ok1[kXItems];
#pragma unroll
    for (int u = 0; u < kXItems; ++u) {
      const int w = min(w0 + u, nitems - 1);
      const int ti = w / cpt;
      const KTile& t = tiles[ti];
      const int pend = st[t.problem].xpend;
      const int e = (w - ti * cpt) * kSTailThreads + threadIdx.x;
      const int a = e / ppr;
      const int gi = t.i0 + a, gj = t.j0 + 2 * (e - a * ppr);
      ok[u] = w0 + u < nitems && pend != 0 && a < th && gi < t.m && gj < t.n;
      ok1[u] = ok[u] && gj + 1 < t.n;
      o[u] = ok[u] ? (size_t)gi * t.pitch + gj : 0;
      xp[u] = (GDP_GLOBAL double*)t.x;
      const GDP_GLOBAL double* d = (const GDP_GLOBAL double*)(pend == 2 ? t.d1 : t.d0);
      xv[u] = *(const GDP_GLOBAL D2*)(xp[u] + o[u]);
      dv[u] = *(const GDP_GLOBAL D2*)(d + o[u]);
      al[u] = st[t.problem].alpha;
    }
#pragma unroll
    for (int u = 0; u < kXItems; ++u) {
      if (ok1[u]) {
        *(GDP_GLOBAL D2*)(xp[u] + o[u]) =
            D2{__builtin_fma(al[u], dv[u].x, xv[u].x), __builtin_fma(al[u], dv[u].y, xv[u].y)};
      } else if (ok[u]) {
        xp[u][o[u]] = __builtin_fma(al[u], dv[u].x, xv[u].x);
      }
    }
  }
}

// NLEV, NEU (1 Neumann, 0 Dirichlet, -1 either), EXACT: the kernel's code fixed to
// tail problems of that kind (0 / -1: read per problem), see STailLevel.
template <typename T, bool DOTS, int NLEV = 0, int NEU = -1, int EXACT = -1>
__global__ __launch_bounds__(kSTailThreads) void k_stail(const TailDesc* __restrict__ tails,
                                                        const LevelDesc* __restrict__ levels,
                                                        ProblemState* __restrict__ st,
                                                        const KrylovDesc* __restrict__ kds,
                                                        const double2* __restrict__ upart,
                                                        const SlotRange* __restrict__ rk,
                                                        double2* __restrict__ zpart,
                                                        const SlotRange* __restrict__ rz,
                                                        int check_conv, int pre, int post,
                                                        int coarsest_sweeps, int order,
                                                        int ntail, const KTile* __restrict__ xt,
                                                        int nxt, int xth, int xtw) {
  using V2 = typename Pair<T>::type;
  constexpr int R0 = STailRows<T>::kRows0;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double sh[2 * kSTailThreads / 64];
  __shared__ double2 bcc;
  if (int(blockIdx.x) >= ntail) {  // the deferred x updates
    st_xupdate(xt, nxt, st, xth, xtw, int(blockIdx.x) - ntail, int(gridDim.x) - ntail);
    return;
  }
  // descriptors through scalar loads (uniform addresses of read-only kernel arguments)
  const TailDesc& D = tails[blockIdx.x];
  const int p = D.problem;
  STailCtx c;
  c.xb = smem;
  c.xh = smem + kSTailXbBytes;
  c.tab = reinterpret_cast<double*>(smem + kSTailXbBytes + kSTailXhBytes);
  const LevelDesc* __restrict__ LV = levels + D.first;
  c.nlev = D.nlev;
  constexpr int kLev = NLEV > 0 ? NLEV : kSTailMaxLevels;  // levels the code has
#pragma unroll
  for (int k = 0; k < kLev; ++k) {
    const LevelDesc& Lk = LV[min(k, c.nlev - 1)];
    c.gmn[k] = Lk.m | (Lk.n << 16);
  }
  c.h2_0 = LV[0].h2;
  c.inv_h2_0 = LV[0].inv_h2;
  c.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  c.lane = threadIdx.x & 63;
  c.pre = pre;
  c.post = post;
  c.csweeps = coarsest_sweeps;
  c.order = order;
  c.hbuf = 0;
  const double* __restrict__ dct = D.dct;
  c.exact = dct != nullptr;
  // the finest tail level's layout, and what the final stores use of it and of the
  // Krylov vectors (the dots of a problem whose finest level is in the tail), read at
  // entry and kept in LDS for the end
  const LevelDesc& L0 = LV[0];
  const int m0 = L0.m, n0 = L0.n;
  const bool neu0 = L0.bc == kNeumann;
  const bool dots = DOTS && D.krylov >= 0;
  __shared__ STailEnd se;
  if (threadIdx.x == 0)
    se = STailEnd{L0.u, dots ? kds[D.krylov].r : nullptr, dots ? kds[D.krylov].q : nullptr,
                  L0.pitch};
  const LevelDesc& Lc = LV[c.nlev - 1];
  const int cm = Lc.m, cn = Lc.n;
  const int ntab = c.exact ? cm * cm + cn * cn + cm * cn : 0;
  c.tabT = c.tab + ntab;
  c.tabTm = c.tabT + (c.exact ? cn * cn : 0);
  c.w0 = c.tabTm + (c.exact ? cm * cm : 0);
  c.w1 = c.w0 + cm * cn;

  // right-hand side of the first tail level and the first elements of the transform
  // tables, issued together before the convergence test
  V2 b[R0];
  {
    // branch-free: every row's pair is loaded (from a clamped address inside the padded
    // rows) before any is waited on, and zeroed outside the level
    const GDP_GLOBAL T* g = (const GDP_GLOBAL T*)L0.b;
    const int gj = 2 * c.lane;
    const int jc = gj < n0 ? gj : 0;
    const int pitch0 = L0.pitch;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const int gi = c.w * R0 + r;
      const V2 v = *(const GDP_GLOBAL V2*)(g + (size_t)min(gi, m0 - 1) * pitch0 + jc);
      b[r].x = (gi < m0 && gj < n0) ? v.x : T(0);
      b[r].y = (gi < m0 && gj + 1 < n0) ? v.y : T(0);
    }
  }
  constexpr int kPre = 2;  // table elements per thread loaded up front
  double tv[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int o = threadIdx.x + k * kSTailThreads;
    tv[k] = o < ntab ? dct[o] : 0.0;
  }
  if (st[p].done) return;
  // the convergence test of the update: here for a problem whose finest level is in the
  // tail (check_conv 1, k_update's partials), and for every problem after the fused
  // update-down kernel (check_conv 2, that kernel's partials)
  if ((check_conv == 2 || (check_conv == 1 && D.krylov >= 0)) &&
      check_converged<kSTailThreads>(st, p, upart, rk[p], true, sh, &bcc))
    return;
  if (c.exact) {
    // Cm, Cn, 1/lambda as they are, and Cn and Cm once more transposed
    const int cn0 = cm * cm, cn1 = cm * cm + cn * cn;
    auto put = [&](int o, double v) {
      c.tab[o] = v;
      if (o < cn0) {
        const int row = o / cm;
        c.tabTm[(o - row * cm) * cm + row] = v;
      } else if (o < cn1) {
        const int e = o - cn0, row = e / cn;
        c.tabT[(e - row * cn) * cn + row] = v;
      }
    };
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
      const int o = threadIdx.x + k * kSTailThreads;
      if (o < ntab) put(o, tv[k]);
    }
    for (int o = threadIdx.x + kPre * kSTailThreads; o < ntab; o += kSTailThreads) put(o, dct[o]);
  }
  // (the first barrier inside the level recursion orders the table stores)

  if constexpr (NEU >= 0) {
    st_level<T, 0, NEU == 1, NLEV, EXACT>(c, b);
  } else {
    if (neu0) st_level<T, 0, true, NLEV, EXACT>(c, b);
    else st_level<T, 0, false, NLEV, EXACT>(c, b);
  }

  // the first tail level's iterate, from the exchange buffer, to global memory
  const STailEnd e = se;
  GDP_GLOBAL T* g = (GDP_GLOBAL T*)e.u;
  const GDP_GLOBAL double* kr = (const GDP_GLOBAL double*)e.kr;
  const GDP_GLOBAL double* kq = (const GDP_GLOBAL double*)e.kq;
  const int pitch0 = e.pitch;
  double zr = 0.0, zq = 0.0;
  const int gj = 2 * c.lane;
  // the FP64 residual and q of the rows, loaded together (branch-free, clamped addresses
  // inside the padded rows) before any is used
  using D2 = typename Pair<double>::type;
  D2 kr2[R0], kq2[R0];
  if (dots) {
    const int jc = gj < n0 ? gj : 0;
#pragma unroll
    for (int r = 0; r < R0; ++r) {
      const size_t o = (size_t)min(c.w * R0 + r, m0 - 1) * pitch0 + jc;
      kr2[r] = *(const GDP_GLOBAL D2*)(kr + o);
      kq2[r] = *(const GDP_GLOBAL D2*)(kq + o);
    }
  }
#pragma unroll
  for (int r = 0; r < R0; ++r) {
    const int gi = c.w * R0 + r;
    if (gi >= m0 || gj >= n0) continue;
    const T* zr0 = reinterpret_cast<const T*>(c.xb) + gi * kXbRow + gj + 1;
    const V2 z = V2{zr0[0], zr0[1]};
    const size_t o = (size_t)gi * pitch0 + gj;
    g[o] = z.x;
    if (gj + 1 < n0) g[o + 1] = z.y;
    if (dots) {
      zr += double(z.x) * kr2[r].x;
      zq += double(z.x) * kq2[r].x;
      if (gj + 1 < n0) {
        zr += double(z.y) * kr2[r].y;
        zq += double(z.y) * kq2[r].y;
      }
    }
  }
  if (dots) write_partial<kSTailThreads>(zr, zq, zpart, rz[p].begin, sh);
}

}  // namespace gdp
