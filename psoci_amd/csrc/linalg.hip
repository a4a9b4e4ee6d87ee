// Dense kernels of the block Davidson: tall-skinny products on FP64 MFMA,
// block updates, preconditioning and precision conversion.
//
// Replaces the per-vector GA dot/axpy loops of PsociGAsubspace
// (generateIncrementalSubspace, PsociGAsubspace.C:273-362; computeRnormSingle
// :874-989) and PsociGAbasis (MGS append :1371-1470; collapse :483-579;
// precondition :1860-1920).  Vectors are HBM-resident [n][ld] row-major
// blocks; every product over n is one launch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "linalg.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;
constexpr size_t kMaxLds = 48 * 1024;  // dynamic LDS per block used without opt-in
typedef double double4_t __attribute__((ext_vector_type(4)));
#ifndef PSOCI_NN_FULL
#define PSOCI_NN_FULL 1
#endif
#ifndef PSOCI_TN_F32MFMA
#define PSOCI_TN_F32MFMA 1
#endif
constexpr bool kTnF32Mfma = PSOCI_TN_F32MFMA;  // FP32-column projections: FP32 MFMA per 16-row step
constexpr bool kNnFull = PSOCI_NN_FULL;  // unpredicated loads for whole chunks of the basis update
// Columns of A per lane in the projection: 2 for FP32 columns, so that a
// wavefront reads whole 128-byte row segments (32 floats, two output tiles)
template <typename TA>
constexpr int tn_cols() {
  return sizeof(TA) == 4 ? 2 : 1;
}

// C_partial[chunk][m][b] = sum over rows of chunk of A[row][a0 + i] * B[row][b0 + j]
// for CP 16x16 output tiles per wavefront (CP = tn_cols<TA>()), on
// v_mfma_f64_16x16x4f64.  Lane layout of the instruction (measured on
// gfx950): A operand lane l = (row k = l/16, col i = l%16), B operand lane
// l = (row k = l/16, col j = l%16), accumulator register t of lane l =
// (i = 4*t + l/16, j = l%16).  With CP = 2 lane i carries A columns
// 32 tm + 2i and 32 tm + 2i + 1, one per tile.  TA: the element type of A
// and B (float columns are widened on load).
template <typename TA, int CP = tn_cols<TA>()>
__global__ __launch_bounds__(256) void k_gemm_tn_mfma(int64_t n, int m, int b, const TA* __restrict__ A, int lda,
                                                      const TA* __restrict__ B, int ldb, int64_t rows_per_chunk,
                                                      double* __restrict__ partial) {
  constexpr int NCH = 4 / CP;  // accumulation chains per tile
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int tiles_m = (m + 16 * CP - 1) / (16 * CP), tiles_b = (b + 15) / 16;
  const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
  if (tile >= tiles_m * tiles_b) return;
  const int tm = tile / tiles_b, tb = tile - tm * tiles_b;
  const int64_t chunk = blockIdx.y;
  const int64_t r0 = chunk * rows_per_chunk;
  const int64_t r1 = (r0 + rows_per_chunk < n) ? r0 + rows_per_chunk : n;
  const int ci = tm * 16 * CP + CP * (lane & 15);
  const int cj = tb * 16 + (lane & 15);
  const int kk = lane >> 4;
  const bool vj = cj < b;
  // 4 / CP independent accumulation chains per tile, 16 rows per step: the
  // loads of a step are all issued before its MFMAs
  double4_t acc[CP][NCH];
#pragma unroll
  for (int c = 0; c < CP; ++c)
#pragma unroll
    for (int h = 0; h < NCH; ++h) acc[c][h] = double4_t{0.0, 0.0, 0.0, 0.0};
  // raw loads of one 16-row step; the next step's loads are issued before
  // the current step's MFMAs (software pipelining: each wavefront keeps a
  // step of loads in flight while its MFMAs run)
  TA ra[CP][4], rb[4];
  auto load = [&](int64_t r) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t row = r + 4 * u + kk;
      const bool in = row < r1;
      if constexpr (CP == 2) {
        // the lane's two columns in one 8-byte load (dword alignment is
        // enough for global loads)
        typedef float f2a4 __attribute__((ext_vector_type(2), aligned(4)));
        if (in && ci + 1 < m) {
          const f2a4 v = *reinterpret_cast<const f2a4*>(A + row * lda + ci);
          ra[0][u] = v.x;
          ra[1][u] = v.y;
        } else {
          ra[0][u] = (in && ci < m) ? A[row * lda + ci] : TA(0);
          ra[1][u] = TA(0);
        }
      } else {
        ra[0][u] = (in && ci < m) ? A[row * lda + ci] : TA(0);
      }
      rb[u] = (in && vj) ? B[row * ldb + cj] : TA(0);
    }
  };
  if (r0 < r1) load(r0);
  if constexpr (CP == 2 && kTnF32Mfma) {
    // FP32 columns: each 16-row step summed on v_mfma_f32_16x16x4f32 (the
    // products of FP32 values, 16 terms in FP32), the steps added in FP64
    typedef float float4_t __attribute__((ext_vector_type(4)));
    for (int64_t r = r0; r < r1; r += 16) {
      float fa[2][4], fb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        fa[0][u] = ra[0][u];
        fa[1][u] = ra[1][u];
        fb[u] = rb[u];
      }
      if (r + 16 < r1) load(r + 16);
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        float4_t p = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < 4; ++u) p = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[c][u], fb[u], p, 0, 0, 0);
        double4_t& q = acc[c][0];
        q[0] += (double)p[0];
        q[1] += (double)p[1];
        q[2] += (double)p[2];
        q[3] += (double)p[3];
      }
    }
    // accumulator t of lane l holds output row 4 (l / 16) + t (FP32 layout)
    double* out = partial + chunk * (int64_t)m * b;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = tm * 32 + 2 * (4 * (lane >> 4) + t) + c;
        const int j = tb * 16 + (lane & 15);
        if (i < m && j < b) out[(int64_t)i * b + j] = acc[c][0][t];
      }
    return;
  }
  for (int64_t r = r0; r < r1; r += 16) {
    double a[CP][4], bb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int c = 0; c < CP; ++c) a[c][u] = (double)ra[c][u];
      bb[u] = (double)rb[u];
    }
    if (r + 16 < r1) load(r + 16);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < CP; ++c)
        acc[c][u % NCH] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][u], bb[u], acc[c][u % NCH], 0, 0, 0);
  }
  double* out = partial + chunk * (int64_t)m * b;
#pragma unroll
  for (int c = 0; c < CP; ++c) {
    double4_t sum = acc[c][0];
#pragma unroll
    for (int h = 1; h < NCH; ++h) sum += acc[c][h];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int i = tm * 16 * CP + CP * (4 * t + (lane >> 4)) + c;
      const int j = tb * 16 + (lane & 15);
      if (i < m && j < b) out[(int64_t)i * b + j] = sum[t];
    }
  }
}

// out[e] = sum over chunks of partial[c][e]: one workgroup per output entry,
// strided partial sums then a fixed-order tree (deterministic)
__global__ __launch_bounds__(256) void k_reduce_chunks(int64_t nchunks, int mb, const double* __restrict__ partial,
                                                       double* __restrict__ out) {
  __shared__ double sh[256];
  const int e = blockIdx.x;
  double s = 0.0;
  for (int64_t c = threadIdx.x; c < nchunks; c += 256) s += partial[c * mb + e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = sh[0];
}

// Out[r][c] = alpha Out[r][c] + sum_t A[r][t] Y1[t][c] (+ sum_t B[r][t] Y2[t][c]
// with TWO; the caller passes -Y2 for A Y1 - B Y2) and, with XOUT,
// Out2[r][c] = sum_t B[r][t] Y1[t][c], on v_mfma_f64_16x16x4f64.  One
// wavefront per 16-row tile and NCT 16-column tiles (columns from
// 128 * blockIdx.y).  Per chunk of 16 values of t, lane l (i = l % 16,
// q = l / 16) reads A[r0 + i][t0 + 4q .. t0 + 4q + 3]; in step u it feeds
// A[r0 + i][t0 + 4q + u] and Y[t0 + 4q + u][c0 + i], so k slot q of the
// instruction carries t = t0 + 4q + u for both operands.  Accumulator
// register j of lane l holds row r0 + 4j + q, column c0 + i.  Each A and B
// row is read from HBM once; Y (m x k, tiny) is read through the caches.
// INPLACE: the rows are read from Out itself (A unused) and overwritten by
// the product (k <= 128: one wavefront reads every input value of its rows
// before it writes any output of them).  For the same reason A may hold the
// same rows as Out at other or overlapping columns when k <= 128.
// TA: the element type of A and B, TO of Out (float columns are widened on
// load and rounded on store; the arithmetic is FP64).
template <int NCT, bool TWO, bool XOUT, bool INPLACE, int RT, typename TA, typename TO>
__global__ __launch_bounds__(256) void k_gemm_nn_mfma(int64_t n, int m, int k, const TA* A,
                                                      int lda, const TA* __restrict__ B, int ldb,
                                                      const double* __restrict__ Y1, const double* __restrict__ Y2,
                                                      double alpha, TO* Out, int ldo, double* __restrict__ Out2,
                                                      int ldo2) {
  static_assert(!INPLACE || sizeof(TA) == sizeof(TO), "in place: one element type");
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t r0 = ((int64_t)blockIdx.x * (blockDim.x / kWave) + wave) * 16 * RT;
  if (r0 >= n) return;  // uniform per wavefront
  const int cb = blockIdx.y * 128;
  const int i = lane & 15, q = lane >> 4;
  const TA* a_row[RT];
  const TA* b_row[RT];
  bool rv[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const int64_t row = r0 + 16 * rt + i;
    rv[rt] = row < n;
    const int64_t rs = rv[rt] ? row : 0;
    a_row[rt] = INPLACE ? reinterpret_cast<const TA*>(Out) + rs * (int64_t)ldo : A + rs * (int64_t)lda;
    b_row[rt] = TWO ? B + rs * (int64_t)ldb : nullptr;
  }
  double4_t acc[RT][NCT], acc2[RT][NCT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      acc[rt][ct] = double4_t{0.0, 0.0, 0.0, 0.0};
      acc2[rt][ct] = acc[rt][ct];
    }
  // raw loads of one chunk of 16 values of t; the next chunk's loads are
  // issued after this chunk's Y values and before its MFMAs (the in-order
  // load counter then lets the MFMAs start while the next chunk is in flight)
  TA ra[RT][4], rb[RT][4];
  auto load = [&](int t0) {
    const int tq = t0 + 4 * q;
    if (kNnFull && t0 + 16 <= m) {
      // a whole chunk: unpredicated loads (rows past n read row 0; their
      // products land in output rows that are not written)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ra[rt][u] = a_row[rt][tq + u];
          rb[rt][u] = TWO ? b_row[rt][tq + u] : TA(0);
        }
    } else {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool tv = rv[rt] && tq + u < m;
          ra[rt][u] = tv ? a_row[rt][tq + u] : TA(0);
          rb[rt][u] = (TWO && tv) ? b_row[rt][tq + u] : TA(0);
        }
    }
  };
  if (m > 0) load(0);
  for (int t0 = 0; t0 < m; t0 += 16) {
    const int tq = t0 + 4 * q;
    double a[RT][4], b[RT][4], y1[4][NCT], y2[4][NCT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[rt][u] = (double)ra[rt][u];
        b[rt][u] = (double)rb[rt][u];
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int t = tq + u, c = cb + ct * 16 + i;
        const bool yv = t < m && c < k;
        y1[u][ct] = yv ? Y1[(int64_t)t * k + c] : 0.0;
        y2[u][ct] = (TWO && yv) ? Y2[(int64_t)t * k + c] : 0.0;
      }
    if (t0 + 16 < m) load(t0 + 16);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][u], y1[u][ct], acc[rt][ct], 0, 0, 0);
          if (XOUT) acc2[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[rt][u], y1[u][ct], acc2[rt][ct], 0, 0, 0);
          if (TWO) acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[rt][u], y2[u][ct], acc[rt][ct], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
      const int c = cb + ct * 16 + i;
      if (c >= k) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t rr = r0 + 16 * rt + 4 * j + q;
        if (rr >= n) continue;
        TO* o = Out + rr * ldo + c;
        *o = (TO)((alpha == 0.0) ? acc[rt][ct][j] : __builtin_fma(alpha, (double)*o, acc[rt][ct][j]));
        if (XOUT) Out2[rr * ldo2 + c] = acc2[rt][ct][j];
      }
    }
}

// Olsen's correction of the diagonal preconditioner, per column c (root
// theta_c, residual r, Ritz vector x, f = 1 / (D - theta_c) where
// |D - theta_c| > dmin, else 1): t = f (r - eps x) with
// eps = (x . f r) / (x . f x), which makes t orthogonal to x in the metric
// of the preconditioner.  k_olsen_dots: per block of rows, the partial sums
// of x f r (partial[blk][c]) and x f x (partial[blk][k + c]); threads are
// (row slot, column) with 256 / k row slots, summed in a fixed order.
__global__ __launch_bounds__(256) void k_olsen_dots(int64_t n, int k, const double* __restrict__ D,
                                                    const double* __restrict__ theta, double dmin,
                                                    const double* __restrict__ R, int ldr,
                                                    const double* __restrict__ X, int ldx, int64_t rows_per_blk,
                                                    double* __restrict__ partial) {
  __shared__ double sxr[256], sxx[256];
  const int slots = 256 / k;
  const int t = threadIdx.x;
  const int c = t % k, slot = t / k;
  double axr = 0.0, axx = 0.0;
  if (slot < slots) {
    const double th = theta[c];
    const int64_t lo = (int64_t)blockIdx.x * rows_per_blk;
    const int64_t hi = lo + rows_per_blk < n ? lo + rows_per_blk : n;
    for (int64_t row = lo + slot; row < hi; row += slots) {
      const double den = D[row] - th;
      const double f = fabs(den) > dmin ? 1.0 / den : 1.0;
      const double x = X[row * ldx + c];
      axr += x * (f * R[row * ldr + c]);
      axx += x * (f * x);
    }
  }
  sxr[t] = axr;
  sxx[t] = axx;
  __syncthreads();
  if (t < k) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < slots; ++j) {
      s1 += sxr[j * k + t];
      s2 += sxx[j * k + t];
    }
    partial[(int64_t)blockIdx.x * 2 * k + t] = s1;
    partial[(int64_t)blockIdx.x * 2 * k + k + t] = s2;
  }
}

// R[row][c] <- f (R - eps_c X), eps_c = dots[c] / dots[k + c] (0 when the
// denominator vanishes)
__global__ void k_olsen_apply(int64_t n, int k, const double* __restrict__ D, const double* __restrict__ theta,
                              double dmin, const double* __restrict__ dots, double* __restrict__ R, int ldr,
                              const double* __restrict__ X, int ldx) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  const double den = D[row] - theta[c];
  const double f = fabs(den) > dmin ? 1.0 / den : 1.0;
  const double bx = dots[k + c];
  const double eps = bx != 0.0 ? dots[c] / bx : 0.0;
  double* r = R + row * ldr + c;
  *r = f * (*r - eps * X[row * ldx + c]);
}

// t[row][c] = r[row][c] / (D[row] - theta[c]) when |D - theta| > dmin
// Out[row][j] = R[row][sel[j]] / (D[row] - theta[sel[j]]) where that
// denominator exceeds dmin in magnitude, else R[row][sel[j]]: the diagonal
// preconditioner applied to the selected columns on their way into the
// basis (one pass; the same numbers as preconditioning R and then copying)
template <typename TR, typename TO>
__global__ void k_precondition_select(int64_t n, int ns, const double* __restrict__ D,
                                      const double* __restrict__ theta_sel, const int* __restrict__ sel,
                                      double dmin, const TR* __restrict__ R, int ldr, TO* __restrict__ Out,
                                      int ldo) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / ns;
  const int j = (int)(g - row * ns);
  if (row >= n) return;
  const double r = (double)R[row * ldr + sel[j]];
  const double den = D[row] - theta_sel[j];
  Out[row * ldo + j] = (TO)(fabs(den) > dmin ? r / den : r);
}

__global__ void k_precondition(int64_t n, int k, const double* __restrict__ D, const double* __restrict__ theta,
                               double dmin, double* __restrict__ R, int ldr) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  double den = D[row] - theta[c];
  if (fabs(den) > dmin) R[row * ldr + c] /= den;
}

__global__ void k_copy_cols_f64_to_f32(int64_t n, int k, const double* __restrict__ A, int lda, float* __restrict__ B,
                                       int ldb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  B[row * ldb + c] = (float)A[row * lda + c];
}

template <typename T>
__global__ void k_copy_cols(int64_t n, int k, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  B[row * ldb + c] = A[row * lda + c];
}

template <typename T>
__global__ void k_unit_cols(int64_t n, int k, const int64_t* __restrict__ idx, T* __restrict__ V, int ldv) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  V[row * ldv + c] = (row == idx[c]) ? T(1) : T(0);
}

// The k lowest entries of D (k <= 128), ordered by (value, index), of each
// slice of 4096 consecutive rows: block b writes its candidates to
// val[b * k + r], idx[b * k + r] (index -1 past the slice's size).  The k
// lowest of the whole vector are among the candidates of all slices.
constexpr int kLowSlice = 4096;
constexpr int kLowPer = kLowSlice / 256;

__device__ __forceinline__ bool low_less(double a, int64_t ia, double b, int64_t ib) {
  return a < b || (a == b && ia < ib);
}

__global__ void __launch_bounds__(256) k_lowest_cands(int64_t n, int k, const double* __restrict__ D,
                                                      double* __restrict__ val, int64_t* __restrict__ idx) {
  __shared__ double sv[4];
  __shared__ int64_t si[4];
  const int64_t base = (int64_t)blockIdx.x * kLowSlice;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double v[kLowPer];
  unsigned left = 0;  // bit e: element base + t + 256 e is in range and not yet taken
#pragma unroll
  for (int e = 0; e < kLowPer; ++e) {
    const int64_t row = base + t + 256 * e;
    v[e] = row < n ? D[row] : 0.0;
    if (row < n) left |= 1u << e;
  }
  auto local = [&](double& bv, int64_t& bi) {
    bv = HUGE_VAL;
    bi = INT64_MAX;
#pragma unroll
    for (int e = 0; e < kLowPer; ++e)
      if (((left >> e) & 1) && low_less(v[e], base + t + 256 * e, bv, bi)) {
        bv = v[e];
        bi = base + t + 256 * e;
      }
  };
  double bv;
  int64_t bi;
  local(bv, bi);
  for (int r = 0; r < k; ++r) {
    double mv = bv;
    int64_t mi = bi;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double ov = __shfl_xor(mv, o);
      const int64_t oi = __shfl_xor(mi, o);
      if (low_less(ov, oi, mv, mi)) {
        mv = ov;
        mi = oi;
      }
    }
    if (lane == 0) {
      sv[wv] = mv;
      si[wv] = mi;
    }
    __syncthreads();
    mv = sv[0];
    mi = si[0];
#pragma unroll
    for (int w = 1; w < 4; ++w)
      if (low_less(sv[w], si[w], mv, mi)) {
        mv = sv[w];
        mi = si[w];
      }
    __syncthreads();
    if (t == 0) {
      val[(int64_t)blockIdx.x * k + r] = mv;
      idx[(int64_t)blockIdx.x * k + r] = mi == INT64_MAX ? -1 : mi;
    }
    if (mi == bi && mi != INT64_MAX) {
      left &= ~(1u << (int)((mi - base - t) >> 8));
      local(bv, bi);
    }
  }
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

namespace {

// Wavefronts per workgroup of k_gemm_tn_mfma for a tile count: the tiles
// spread evenly over the fewest workgroups of at most 4 wavefronts.
int tn_waves(int tiles) {
  const int groups = (tiles + 3) / 4;
  return (tiles + groups - 1) / groups;
}

// Workgroups of k_gemm_tn_mfma with w wavefronts that the current device
// holds at once (compute units x resident workgroups per unit), queried once
// per device, width and columns per lane.
int64_t tn_resident(int w, int cp) {
  static int64_t cache[64][5][3] = {};
  int dev = 0;
  PSOCI_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  if (cache[dev][w][cp] == 0) {
    int per_cu = 0, cus = 0;
    if (cp == 2)
      PSOCI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_tn_mfma<float>, kWave * w, 0));
    else
      PSOCI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_gemm_tn_mfma<double>, kWave * w, 0));
    PSOCI_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev][w][cp] = std::max<int64_t>(1, (int64_t)per_cu * cus);
  }
  return cache[dev][w][cp];
}

// Output tiles of the projection's wavefronts (cp columns of A per lane).
int tn_tiles(int m, int b, int cp) { return ((m + 16 * cp - 1) / (16 * cp)) * ((b + 15) / 16); }

// Row chunks of the projection: one chunk per group of workgroups covering
// the output tiles, as many groups as the device holds at once (a single
// round of workgroups, no partly filled second round).
void tn_chunks(int64_t n, int m, int b, int cp, int64_t& rows_per_chunk, int64_t& nchunks) {
  const int tiles = tn_tiles(m, b, cp);
  const int w = tn_waves(tiles);
  const int64_t blocks_per_chunk = (tiles + w - 1) / w;
  int64_t want = std::max<int64_t>(1, tn_resident(w, cp) / blocks_per_chunk);
  want = std::min<int64_t>(want, 65535);
  rows_per_chunk = (n + want - 1) / want;
  rows_per_chunk = std::max<int64_t>(64, (rows_per_chunk + 15) / 16 * 16);
  nchunks = (n + rows_per_chunk - 1) / rows_per_chunk;
}

}  // namespace

size_t Linalg::tn_part_size(int64_t n, int m, int b) {
  // the larger of the FP64 and FP32 kernels' needs
  int64_t rpc, nch1, nch2;
  tn_chunks(n, m, b, 1, rpc, nch1);
  tn_chunks(n, m, b, 2, rpc, nch2);
  return (size_t)std::max(nch1, nch2) * m * b;
}

void Linalg::reserve_tn(int64_t n, int mmax, int bmax) {
  size_t need = 0;
  for (int m = 1; m <= mmax; ++m)
    for (int b = 1; b <= bmax; ++b) need = std::max(need, tn_part_size(n, m, b));
  part.reserve(need);
}

template <typename TA>
void Linalg::gemm_tn(int64_t n, int m, int b, const TA* A, int lda, const TA* B, int ldb, double* out_dev) {
  if (m == 0 || b == 0) return;
  constexpr int cp = tn_cols<TA>();
  int64_t rows_per_chunk, nchunks;
  tn_chunks(n, m, b, cp, rows_per_chunk, nchunks);
  part.reserve((size_t)nchunks * m * b);
  const int tiles = tn_tiles(m, b, cp);
  const int w = tn_waves(tiles);
  dim3 grid(nblk(tiles, w), (unsigned)nchunks);
  hipLaunchKernelGGL(k_gemm_tn_mfma<TA>, grid, dim3(kWave * w), 0, stream, n, m, b, A, lda, B, ldb, rows_per_chunk,
                     part.p);
  PSOCI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)(m * b)), dim3(256), 0, stream, nchunks, m * b, part.p, out_dev);
  PSOCI_HIP(hipGetLastError());
}

template <typename TA>
void Linalg::gemm_tn_host(int64_t n, int m, int b, const TA* A, int lda, const TA* B, int ldb, double* out_host) {
  small.reserve((size_t)m * b);
  gemm_tn(n, m, b, A, lda, B, ldb, small.p);
  PSOCI_HIP(hipMemcpyAsync(out_host, small.p, (size_t)m * b * sizeof(double), hipMemcpyDeviceToHost, stream));
  PSOCI_HIP(hipStreamSynchronize(stream));
}

namespace {

#ifndef PSOCI_NN_RT1
#define PSOCI_NN_RT1 2
#endif
#ifndef PSOCI_NN_RT2
#define PSOCI_NN_RT2 1
#endif
constexpr int kNnRt1 = PSOCI_NN_RT1, kNnRt2 = PSOCI_NN_RT2;

template <bool TWO, bool XOUT, bool INPLACE, typename TA, typename TO>
void launch_nn(hipStream_t st, int64_t n, int m, int k, const TA* A, int lda, const TA* B, int ldb,
               const double* Y1, const double* Y2, double alpha, TO* Out, int ldo, double* Out2, int ldo2) {
  const int nct = std::min(8, (k + 15) / 16);
  // row tiles per wavefront: each Y value read feeds RT tiles of 16 rows
#define PSOCI_NN(NCT, RT)                                                                            \
  hipLaunchKernelGGL((k_gemm_nn_mfma<NCT, TWO, XOUT, INPLACE, RT, TA, TO>),                          \
                     dim3(nblk(n, 64 * RT), (unsigned)((k + 127) / 128)), dim3(256), 0, st, n, m, k, A, lda, B, \
                     ldb, Y1, Y2, alpha, Out, ldo, Out2, ldo2)
  if (nct <= 1) PSOCI_NN(1, kNnRt1);
  else if (nct <= 2) PSOCI_NN(2, kNnRt2);
  else if (nct <= 4) PSOCI_NN(4, 1);
  else PSOCI_NN(8, 1);
#undef PSOCI_NN
  PSOCI_HIP(hipGetLastError());
}

}  // namespace

template <typename TA, typename TO>
void Linalg::gemm_nn(int64_t n, int m, int k, const TA* A, int lda, const double* Y1_host, double alpha, TO* Out,
                     int ldo) {
  if (k == 0 || n == 0) return;
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  launch_nn<false, false, false, TA, TO>(stream, n, m, k, A, lda, nullptr, 0, ysmall.p, nullptr, alpha, Out, ldo,
                                         nullptr, 0);
}

template <typename T>
void Linalg::gemm_nn_rows(int64_t n, int m, int k, const T* A, int lda, const double* Y1_host, T* Out, int ldo) {
  if (k == 0 || n == 0) return;
  PSOCI_CHECK(k <= 128, PSOCI_ERR_ARG, "row-aliased product: at most 128 output columns");
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  launch_nn<false, false, false, T, T>(stream, n, m, k, A, lda, nullptr, 0, ysmall.p, nullptr, 0.0, Out, ldo, nullptr,
                                       0);
}

template <typename T>
void Linalg::gemm_nn_inplace(int64_t n, int m, int k, const double* Y1_host, T* Out, int ldo) {
  if (k == 0 || n == 0) return;
  PSOCI_CHECK(k <= 128 && k <= m, PSOCI_ERR_ARG, "in-place product: at most 128 and at most m output columns");
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  launch_nn<false, false, true, T, T>(stream, n, m, k, nullptr, 0, nullptr, 0, ysmall.p, nullptr, 0.0, Out, ldo,
                                      nullptr, 0);
}

template <typename TA, typename TO>
void Linalg::gemm_nn2(int64_t n, int m, int k, const TA* A, int lda, const double* Y1_host, const TA* B, int ldb,
                      const double* Y2_host, TO* Out, int ldo, double* X, int ldx) {
  if (k == 0 || n == 0) return;
  ysmall.reserve((size_t)2 * m * k);
  yneg.assign(Y2_host, Y2_host + (size_t)m * k);
  for (auto& v : yneg) v = -v;
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  PSOCI_HIP(hipMemcpyAsync(ysmall.p + (size_t)m * k, yneg.data(), (size_t)m * k * sizeof(double),
                           hipMemcpyHostToDevice, stream));
  if (X)
    launch_nn<true, true, false, TA, TO>(stream, n, m, k, A, lda, B, ldb, ysmall.p, ysmall.p + (size_t)m * k, 0.0,
                                         Out, ldo, X, ldx);
  else
    launch_nn<true, false, false, TA, TO>(stream, n, m, k, A, lda, B, ldb, ysmall.p, ysmall.p + (size_t)m * k, 0.0,
                                          Out, ldo, nullptr, 0);
  // yneg is read by the copy above: keep it until the stream has consumed it
  PSOCI_HIP(hipStreamSynchronize(stream));
}

void Linalg::precondition(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R, int ldr) {
  theta.reserve(k);
  PSOCI_HIP(hipMemcpyAsync(theta.p, theta_host, k * sizeof(double), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_precondition, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, D, theta.p, dmin, R, ldr);
  PSOCI_HIP(hipGetLastError());
}

template <typename TR, typename TO>
void Linalg::precondition_select(int64_t n, int k, const double* D, const double* theta_host, double dmin,
                                 const TR* R, int ldr, const int* sel_host, int ns, TO* Out, int ldo) {
  if (ns == 0 || n == 0) return;
  std::vector<double> ts(ns);
  for (int j = 0; j < ns; ++j) ts[j] = theta_host[sel_host[j]];
  (void)k;
  theta.reserve(ns);
  isel.reserve(ns);
  PSOCI_HIP(hipMemcpyAsync(theta.p, ts.data(), ns * sizeof(double), hipMemcpyHostToDevice, stream));
  PSOCI_HIP(hipMemcpyAsync(isel.p, sel_host, ns * sizeof(int), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL((k_precondition_select<TR, TO>), dim3(nblk(n * ns, 256)), dim3(256), 0, stream, n, ns, D, theta.p,
                     isel.p, dmin, R, ldr, Out, ldo);
  PSOCI_HIP(hipGetLastError());
  // ts is read by the copy above: keep it until the stream has consumed it
  PSOCI_HIP(hipStreamSynchronize(stream));
}

void Linalg::precondition_olsen(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R,
                                int ldr, const double* X, int ldx) {
  if (k == 0 || n == 0) return;
  theta.reserve(k);
  PSOCI_HIP(hipMemcpyAsync(theta.p, theta_host, k * sizeof(double), hipMemcpyHostToDevice, stream));
  const int64_t rows_per_blk = 4096;
  const int64_t nb = (n + rows_per_blk - 1) / rows_per_blk;
  part.reserve((size_t)nb * 2 * k);
  small.reserve((size_t)2 * k);
  hipLaunchKernelGGL(k_olsen_dots, dim3((unsigned)nb), dim3(256), 0, stream, n, k, D, theta.p, dmin, R, ldr, X, ldx,
                     rows_per_blk, part.p);
  PSOCI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)(2 * k)), dim3(256), 0, stream, nb, 2 * k, part.p, small.p);
  PSOCI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_olsen_apply, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, D, theta.p, dmin, small.p, R,
                     ldr, X, ldx);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::to_f32(int64_t n, int k, const double* A, int lda, float* B, int ldb) {
  hipLaunchKernelGGL(k_copy_cols_f64_to_f32, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, A, lda, B, ldb);
  PSOCI_HIP(hipGetLastError());
}

template <typename T>
void Linalg::copy_cols(int64_t n, int k, const T* A, int lda, T* B, int ldb) {
  hipLaunchKernelGGL(k_copy_cols<T>, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, A, lda, B, ldb);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::lowest(int64_t n, int k, const double* D, int64_t* idx_host) {
  PSOCI_CHECK(k >= 1 && k <= 128 && k <= n, PSOCI_ERR_ARG, "lowest: k must be in 1..min(128, n)");
  const int64_t nb = (n + kLowSlice - 1) / kLowSlice;
  part.reserve((size_t)nb * k);
  ipart.reserve((size_t)nb * k);
  hipLaunchKernelGGL(k_lowest_cands, dim3((unsigned)nb), dim3(256), 0, stream, n, k, D, part.p, ipart.p);
  PSOCI_HIP(hipGetLastError());
  std::vector<double> cv((size_t)nb * k);
  std::vector<int64_t> ci((size_t)nb * k);
  part.download(cv.data(), cv.size(), stream);
  ipart.download(ci.data(), ci.size(), stream);
  PSOCI_HIP(hipStreamSynchronize(stream));
  std::vector<int64_t> ord;
  ord.reserve(ci.size());
  for (size_t j = 0; j < ci.size(); ++j)
    if (ci[j] >= 0) ord.push_back((int64_t)j);
  PSOCI_CHECK((int64_t)ord.size() >= k, PSOCI_ERR_ARG, "lowest: too few candidates");
  std::partial_sort(ord.begin(), ord.begin() + k, ord.end(), [&](int64_t a, int64_t b) {
    return cv[a] < cv[b] || (cv[a] == cv[b] && ci[a] < ci[b]);
  });
  for (int r = 0; r < k; ++r) idx_host[r] = ci[ord[r]];
}

template <typename T>
void Linalg::unit_cols(int64_t n, int k, const int64_t* idx_host, T* V, int ldv) {
  DevBuf<int64_t> idx;
  idx.upload(idx_host, k, stream);
  hipLaunchKernelGGL(k_unit_cols<T>, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, idx.p, V, ldv);
  PSOCI_HIP(hipGetLastError());
  PSOCI_HIP(hipStreamSynchronize(stream));
}

// the element types the block Davidson uses: FP64 bases, FP32 bases in the
// FP32 phase of a mixed-precision solve
template void Linalg::gemm_tn<double>(int64_t, int, int, const double*, int, const double*, int, double*);
template void Linalg::gemm_tn<float>(int64_t, int, int, const float*, int, const float*, int, double*);
template void Linalg::gemm_tn_host<double>(int64_t, int, int, const double*, int, const double*, int, double*);
template void Linalg::gemm_tn_host<float>(int64_t, int, int, const float*, int, const float*, int, double*);
template void Linalg::gemm_nn<double, double>(int64_t, int, int, const double*, int, const double*, double, double*,
                                              int);
template void Linalg::gemm_nn<float, float>(int64_t, int, int, const float*, int, const double*, double, float*, int);
template void Linalg::gemm_nn<float, double>(int64_t, int, int, const float*, int, const double*, double, double*,
                                             int);
template void Linalg::gemm_nn<double, float>(int64_t, int, int, const double*, int, const double*, double, float*,
                                             int);
template void Linalg::gemm_nn_rows<double>(int64_t, int, int, const double*, int, const double*, double*, int);
template void Linalg::gemm_nn_rows<float>(int64_t, int, int, const float*, int, const double*, float*, int);
template void Linalg::gemm_nn_inplace<double>(int64_t, int, int, const double*, double*, int);
template void Linalg::gemm_nn_inplace<float>(int64_t, int, int, const double*, float*, int);
template void Linalg::gemm_nn2<double, double>(int64_t, int, int, const double*, int, const double*, const double*,
                                               int, const double*, double*, int, double*, int);
template void Linalg::gemm_nn2<float, double>(int64_t, int, int, const float*, int, const double*, const float*, int,
                                              const double*, double*, int, double*, int);
template void Linalg::gemm_nn2<float, float>(int64_t, int, int, const float*, int, const double*, const float*, int,
                                             const double*, float*, int, double*, int);
template void Linalg::precondition_select<double, double>(int64_t, int, const double*, const double*, double,
                                                          const double*, int, const int*, int, double*, int);
template void Linalg::precondition_select<double, float>(int64_t, int, const double*, const double*, double,
                                                         const double*, int, const int*, int, float*, int);
template void Linalg::precondition_select<float, float>(int64_t, int, const double*, const double*, double,
                                                        const float*, int, const int*, int, float*, int);
template void Linalg::copy_cols<double>(int64_t, int, const double*, int, double*, int);
template void Linalg::copy_cols<float>(int64_t, int, const float*, int, float*, int);
template void Linalg::unit_cols<double>(int64_t, int, const int64_t*, double*, int);
template void Linalg::unit_cols<float>(int64_t, int, const int64_t*, float*, int);

}  // namespace psoci
