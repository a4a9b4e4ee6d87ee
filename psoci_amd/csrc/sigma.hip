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
// per row with whole x rows per lane.
template <typename T, int K, int G, int R>
__global__ __launch_bounds__(256) void k_spmm(int64_t n, const int64_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ X, int ldx, T* __restrict__ Y, int ldy) {
  constexpr int LR = kWave / R;   // lanes per row
  constexpr int NS = LR / G;      // nonzero slots per row
  constexpr int C = K / G;        // columns per lane
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r = ((int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave)) * R + lane / LR;
  const int lr = lane % LR, slot = lr / G, cc = lr % G;
  T acc[C];
#pragma unroll
  for (int v = 0; v < C; ++v) acc[v] = (T)0;
  if (r < n) {
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
  if (r < n && slot == 0) {
    T* y = Y + r * ldy + cc * C;
#pragma unroll
    for (int v = 0; v < C; ++v) y[v] = acc[v];
  }
}

template <typename T, int K, int G, int R>
void launch_gr(Ctx& c, const T* val, const T* X, int ldx, T* Y, int ldy) {
  const int64_t n = c.sp.nsef_total;
  const int64_t rows_per_block = 4 * R;
  hipLaunchKernelGGL((k_spmm<T, K, G, R>), dim3((unsigned)((n + rows_per_block - 1) / rows_per_block)), dim3(256), 0,
                     c.stream, n, c.d_rowptr.p, c.d_col.p, val, X, ldx, Y, ldy);
  PSOCI_HIP(hipGetLastError());
}

// the tuned shape of this precision, where the width allows it
template <typename T, int K>
void launch_spmm(Ctx& c, const T* val, const T* X, int ldx, T* Y, int ldy) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  const int g = K >= c.spmm_g[pi] ? c.spmm_g[pi] : 1;
  const int key = g * 16 + c.spmm_r[pi];
  switch (key) {
#define PSOCI_SPMM_CASE(GG, RR)                                     \
  case GG * 16 + RR:                                                \
    if constexpr (K >= GG) {                                        \
      launch_gr<T, K, (K >= GG ? GG : 1), RR>(c, val, X, ldx, Y, ldy); \
      return;                                                       \
    }                                                               \
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
  launch_gr<T, K, 1, 1>(c, val, X, ldx, Y, ldy);
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
template <typename T>
double tune_prec(Ctx& c, const T* val, hipEvent_t e0, hipEvent_t e1) {
  const int pi = sizeof(T) == 4 ? 1 : 0;
  const int64_t n = c.sp.nsef_total;
  const int K = 16;
  DevBuf<T> X, Y;
  X.alloc((size_t)n * K);
  Y.alloc((size_t)n * K);
  PSOCI_HIP(hipMemsetAsync(X.p, 0, (size_t)n * K * sizeof(T), c.stream));
  double best = 1e300;
  int bg = 1, br = 1;
  for (int g : {1, 2, 4})
    for (int r : {1, 2, 4, 8}) {
      c.spmm_g[pi] = g;
      c.spmm_r[pi] = r;
      double t = 1e300;
      for (int rep = 0; rep < 3; ++rep) {
        PSOCI_HIP(hipEventRecord(e0, c.stream));
        launch_spmm<T, 16>(c, val, X.p, K, Y.p, K);
        PSOCI_HIP(hipEventRecord(e1, c.stream));
        PSOCI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PSOCI_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) t = std::min(t, (double)ms);
      }
      if (t < best) {
        best = t;
        bg = g;
        br = r;
      }
    }
  c.spmm_g[pi] = bg;
  c.spmm_r[pi] = br;
  return best;
}

void tune_stored(Ctx& c) {
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
  if (prec == PSOCI_FP32)
    return pad::staging_bytes<float>(n, k, 32, static_cast<const float*>(X), ldx, static_cast<const float*>(Y), ldy,
                                     w24);
  return pad::staging_bytes<double>(n, k, 32, static_cast<const double*>(X), ldx, static_cast<const double*>(Y), ldy,
                                    w24);
}

}  // namespace psoci
