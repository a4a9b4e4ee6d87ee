// Sigma (H . X for k vectors) from the stored Hamiltonian.
//
// Replaces matrixVectorProductsSplitChunkMultipleJchunks
// (PsociGAhamiltonian.C:6921-7269), which forms y_i += sum_j H_ij x_j and
// y_j += H_ij x_i from the stored lower triangle, one vector at a time, with
// per-row GA gets and an all-reduce of y per chunk.  Here H is resident in
// HBM as the symmetric full CSR (lower part of each row first), X and Y are
// [n][k] row-major (the reference's GA vector layout), and one wavefront
// reduces one row for all k vectors at once: the row's values and columns
// stream from HBM once per launch, X rows are gathered from L2/MALL.
#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "ctx.hpp"
#include "pad.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c);
template <>
__device__ __forceinline__ double fmadd<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// R rows per wavefront (64/R lanes per row: short rows share a wavefront)
// and G lanes per nonzero (each of them multiplies K/G of the K columns, so
// the G lanes of a nonzero read one contiguous piece of its x row).  Lanes
// of a row stride over its nonzeros; partial sums are reduced across the
// nonzero slots of the row with shuffles.  (R, G) = (1, 1) is one wavefront
// per row with whole x rows per lane.  rows (optional): the launch covers
// the nrows rows listed there (one row-length bucket) instead of 0..n-1.
template <typename T, int K, int G, int R>
__global__ __launch_bounds__(256) void k_spmm(int64_t nrows, const int32_t* __restrict__ rows,
                                              const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                              const T* __restrict__ val, const T* __restrict__ X, int ldx,
                                              T* __restrict__ Y, int ldy) {
  constexpr int LR = kWave / R;   // lanes per row
  constexpr int NS = LR / G;      // nonzero slots per row
  constexpr int C = K / G;        // columns per lane
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t slot_r =
      ((int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave)) * R + lane / LR;
  const bool live = slot_r < nrows;
  const int64_t r = live ? (rows ? (int64_t)rows[slot_r] : slot_r) : 0;
  const int lr = lane % LR, slot = lr / G, cc = lr % G;
  T acc[C];
#pragma unroll
  for (int v = 0; v < C; ++v) acc[v] = (T)0;
  if (live) {
    const int64_t lo = rowptr[r], hi = rowptr[r + 1];
    for (int64_t q = lo + slot; q < hi; q += NS) {
      const T a = val[q];
      const T* __restrict__ x = X + (int64_t)col[q] * ldx + cc * C;
#pragma unroll
      for (int v = 0; v < C; ++v) acc[v] = fmadd(a, x[v], acc[v]);
    }
  }
  // reduce over the nonzero slots of the row (lane bits G .. LR-1)
#pragma unroll
  for (int off = G; off < LR; off <<= 1) {
#pragma unroll
    for (int v = 0; v < C; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  }
  if (live && slot == 0) {
    T* y = Y + r * ldy + cc * C;
#pragma unroll
    for (int v = 0; v < C; ++v) y[v] = acc[v];
  }
}

template <typename T, int K, int G, int R>
void launch_gr(Ctx& c, const T* val, const T* X, int ldx, T* Y, int ldy, const int32_t* rows, int64_t nrows) {
  const int64_t rows_per_block = 4 * R;
  if (nrows <= 0) return;
  hipLaunchKernelGGL((k_spmm<T, K, G, R>), dim3((unsigned)((nrows + rows_per_block - 1) / rows_per_block)), dim3(256),
                     0, c.stream, nrows, rows, c.d_rowptr.p, c.d_col.p, val, X, ldx, Y, ldy);
  PSOCI_HIP(hipGetLastError());
}

// one launch of shape (g lanes per nonzero, r rows per wavefront) over the
// listed rows (rows null: all n)
template <typename T, int K>
void launch_shape(Ctx& c, int g0, int r, const T* val, const T* X, int ldx, T* Y, int ldy, const int32_t* rows,
                  int64_t nrows) {
  const int g = K >= g0 ? g0 : 1;
  const int key = g * 16 + r;
  switch (key) {
#define PSOCI_SPMM_CASE(GG, RR)                                                    \
  case GG * 16 + RR:                                                               \
    if constexpr (K >= GG) {                                                       \
      launch_gr<T, K, (K >= GG ? GG : 1), RR>(c, val, X, ldx, Y, ldy, rows, nrows); \
      return;                                                                      \
    }                                                                              \
    break;
    PSOCI_SPMM_CASE(1, 1)
    PSOCI_SPMM_CASE(1, 2)
    PSOCI_SPMM_CASE(1, 4)
    PSOCI_SPMM_CASE(1, 8)
    PSOCI_SPMM_CASE(2, 1)
    PSOCI_SPMM_CASE(2, 2)
    PSOCI_SPMM_CASE(2, 4)
    PSOCI_SPMM_CASE(2, 8)
    PSOCI_SPMM_CASE(4, 1)
    PSOCI_SPMM_CASE(4, 2)
    PSOCI_SPMM_CASE(4, 4)
    PSOCI_SPMM_CASE(4, 8)
#undef PSOCI_SPMM_CASE
    default: break;
  }
  launch_gr<T, K, 1, 1>(c, val, X, ldx, Y, ldy, rows, nrows);
}

// the tuned shape(s) of this precision: one launch over all rows, or one
// per non-empty row-length bucket with that bucket's shape
template <typename T, int K>
void launch_spmm(Ctx& c, const T* val, const T* X, int ldx, T* Y, int ldy) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  if (!c.spmm_bucketed[pi]) {
    launch_shape<T, K>(c, c.spmm_g[pi], c.spmm_r[pi], val, X, ldx, Y, ldy, nullptr, c.sp.nsef_total);
    return;
  }
  for (int b = 0; b < kSpmmBuckets; ++b) {
    const int64_t lo = c.spmm_boff[b], cnt = c.spmm_boff[b + 1] - lo;
    if (cnt > 0)
      launch_shape<T, K>(c, c.spmm_bg[pi][b], c.spmm_br[pi][b], val, X, ldx, Y, ldy, c.d_brows.p + lo, cnt);
  }
}

template <typename T>
void spmm_any_k(Ctx& c, const T* val, int k, const T* X, int ldx, T* Y, int ldy, const Staging& st) {
  pad::blocks<T>(c, c.sp.nsef_total, k, 32, X, ldx, Y, ldy, [&](int K, const T* x, int lx, T* y, int ly) {
    switch (K) {
      case 32: launch_spmm<T, 32>(c, val, x, lx, y, ly); break;
      case 16: launch_spmm<T, 16>(c, val, x, lx, y, ly); break;
      case 8: launch_spmm<T, 8>(c, val, x, lx, y, ly); break;
      case 4: launch_spmm<T, 4>(c, val, x, lx, y, ly); break;
      case 2: launch_spmm<T, 2>(c, val, x, lx, y, ly); break;
      default: launch_spmm<T, 1>(c, val, x, lx, y, ly); break;
    }
  }, false, st);
}

}  // namespace

// Shape autotuner of the stored sigma: for each precision, every (lanes per
// nonzero, rows per wavefront) pair times the 16-column product on the built
// matrix (best of two after a warm-up) and the fastest is kept for all
// widths.  PSOCI_SPMM_SHAPE=<G>x<R> fixes it instead.
// best of two timed 16-column launches (after a warm-up) of f
template <class F>
double time_ms(Ctx& c, hipEvent_t e0, hipEvent_t e1, F&& f) {
  double t = 1e300;
  for (int rep = 0; rep < 3; ++rep) {
    PSOCI_HIP(hipEventRecord(e0, c.stream));
    f();
    PSOCI_HIP(hipEventRecord(e1, c.stream));
    PSOCI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    PSOCI_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0) t = std::min(t, (double)ms);
  }
  return t;
}

// Per precision: the best (lanes per nonzero, rows per wavefront) for all
// rows in one launch, and the best for each row-length bucket on its own;
// the bucketed launches are kept when together they beat the single one by
// 3 % or more (PSOCI_SPMM_BUCKETS=0 / 1 fixes the choice).
template <typename T>
double tune_prec(Ctx& c, const T* val, hipEvent_t e0, hipEvent_t e1) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  const int64_t n = c.sp.nsef_total;
  const int K = 16;
  DevBuf<T> X, Y;
  X.alloc((size_t)n * K);
  Y.alloc((size_t)n * K);
  PSOCI_HIP(hipMemsetAsync(X.p, 0, (size_t)n * K * sizeof(T), c.stream));
  const int gs[] = {1, 2, 4}, rs[] = {1, 2, 4, 8};
  auto best_shape = [&](const int32_t* rows, int64_t nrows, int& bg, int& br) {
    double best = 1e300;
    bg = br = 1;
    for (int g : gs)
      for (int r : rs) {
        const double t = time_ms(c, e0, e1, [&] { launch_shape<T, 16>(c, g, r, val, X.p, K, Y.p, K, rows, nrows); });
        if (t < best) {
          best = t;
          bg = g;
          br = r;
        }
      }
    return best;
  };
  c.spmm_bucketed[pi] = false;
  const double single = best_shape(nullptr, n, c.spmm_g[pi], c.spmm_r[pi]);
  int nonempty = 0;
  for (int b = 0; b < kSpmmBuckets; ++b) {
    const int64_t cnt = c.spmm_boff[b + 1] - c.spmm_boff[b];
    c.spmm_bg[pi][b] = c.spmm_br[pi][b] = 1;
    if (cnt > 0) {
      ++nonempty;
      best_shape(c.d_brows.p + c.spmm_boff[b], cnt, c.spmm_bg[pi][b], c.spmm_br[pi][b]);
    }
  }
  if (nonempty > 1) {
    c.spmm_bucketed[pi] = true;
    const double bucketed = time_ms(c, e0, e1, [&] { launch_spmm<T, 16>(c, val, X.p, K, Y.p, K); });
    c.spmm_bucketed[pi] = bucketed < 0.97 * single;
    c.spmm_bucket_ms[pi] = bucketed;
    // PSOCI_SPMM_BUCKETS=0 / 1 fixes the choice instead
    if (const char* eb = std::getenv("PSOCI_SPMM_BUCKETS")) c.spmm_bucketed[pi] = eb[0] == '1';
  } else {
    c.spmm_bucket_ms[pi] = -1.0;
  }
  return c.spmm_bucketed[pi] ? c.spmm_bucket_ms[pi] : single;
}

// Row-length buckets of the stored matrix (full symmetric rows): lengths
// < 64, 64-255, 256-1023, 1024-4095, >= 4096; the row lists in ascending
// order per bucket.  A property of the matrix alone, so a loaded checkpoint
// gets the same buckets.
void set_buckets(Ctx& c) {
  const int64_t n = c.sp.nsef_total;
  std::vector<int64_t> rp(n + 1);
  c.d_rowptr.download(rp.data(), n + 1, c.stream);
  c.sync();
  std::vector<int32_t> rows;
  rows.reserve(n);
  const int64_t lim[kSpmmBuckets] = {64, 256, 1024, 4096, INT64_MAX};
  c.spmm_boff[0] = 0;
  for (int b = 0; b < kSpmmBuckets; ++b) {
    const int64_t lo = b ? lim[b - 1] : 0;
    for (int64_t r = 0; r < n; ++r) {
      const int64_t len = rp[r + 1] - rp[r];
      if (len >= lo && len < lim[b]) rows.push_back((int32_t)r);
    }
    c.spmm_boff[b + 1] = (int64_t)rows.size();
  }
  c.d_brows.upload(rows, c.stream);
  c.sync();
}

void tune_stored(Ctx& c) {
  set_buckets(c);
  c.spmm_bucketed[0] = c.spmm_bucketed[1] = false;
  const char* env = std::getenv("PSOCI_SPMM_SHAPE");
  if (env) {
    int g = 1, r = 1;
    if (std::sscanf(env, "%dx%d", &g, &r) == 2 && (g == 1 || g == 2 || g == 4) &&
        (r == 1 || r == 2 || r == 4 || r == 8)) {
      c.spmm_g[0] = c.spmm_g[1] = g;
      c.spmm_r[0] = c.spmm_r[1] = r;
    }
    c.spmm_tune_ms[0] = c.spmm_tune_ms[1] = -1.0;
    return;
  }
  hipEvent_t e0, e1;
  PSOCI_HIP(hipEventCreate(&e0));
  PSOCI_HIP(hipEventCreate(&e1));
  c.spmm_tune_ms[1] = tune_prec<float>(c, c.d_val32.p, e0, e1);
  c.spmm_tune_ms[0] = tune_prec<double>(c, c.d_val.p, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void sigma_stored(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy, const Staging& st) {
  PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
  if (prec == PSOCI_FP32)
    spmm_any_k<float>(c, c.d_val32.p, k, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy, st);
  else
    spmm_any_k<double>(c, c.d_val.p, k, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy, st);
}

size_t sigma_staging_bytes(const Ctx& c, int mode, int prec, int k, const void* X, int ldx, const void* Y, int ldy) {
  // the stored sigma has widths 1..32 (powers of two), the direct one also 24
  const bool w24 = mode == PSOCI_SIGMA_DIRECT;
  const int64_t n = c.sp.nsef_total;
  if (mode == PSOCI_SIGMA_DIRECT && direct_engine(c) == 2) return 0;  // reads X and writes Y at their strides
  if (prec == PSOCI_FP32)
    return pad::staging_bytes<float>(n, k, 32, static_cast<const float*>(X), ldx, static_cast<const float*>(Y), ldy,
                                     w24);
  return pad::staging_bytes<double>(n, k, 32, static_cast<const double*>(X), ldx, static_cast<const double*>(Y), ldy,
                                    w24);
}

}  // namespace psoci
