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

// one wavefront per row; K vectors; rows-per-block = blockDim / 64
template <typename T, int K>
__global__ __launch_bounds__(256) void k_spmm(int64_t n, const int64_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col, const T* __restrict__ val,
                                              const T* __restrict__ X, int ldx, T* __restrict__ Y, int ldy) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (r >= n) return;
  T acc[K];
#pragma unroll
  for (int v = 0; v < K; ++v) acc[v] = (T)0;
  const int64_t lo = rowptr[r], hi = rowptr[r + 1];
  for (int64_t q = lo + lane; q < hi; q += kWave) {
    const T a = val[q];
    const T* __restrict__ x = X + (int64_t)col[q] * ldx;
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = fmadd(a, x[v], acc[v]);
  }
#pragma unroll
  for (int v = 0; v < K; ++v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc[v] += __shfl_xor(acc[v], off, kWave);
  }
  T* y = Y + r * ldy;
#pragma unroll
  for (int v = 0; v < K; ++v)
    if (lane == v) y[v] = acc[v];
}

template <typename T, int K>
void launch_spmm(Ctx& c, const T* val, const T* X, int ldx, T* Y, int ldy) {
  const int64_t n = c.sp.nsef_total;
  const int wpb = 4;
  hipLaunchKernelGGL((k_spmm<T, K>), dim3((unsigned)((n + wpb - 1) / wpb)), dim3(256), 0, c.stream, n, c.d_rowptr.p,
                     c.d_col.p, val, X, ldx, Y, ldy);
  PSOCI_HIP(hipGetLastError());
}

template <typename T>
void spmm_any_k(Ctx& c, const T* val, int k, const T* X, int ldx, T* Y, int ldy) {
  pad::blocks<T>(c, c.sp.nsef_total, k, 32, X, ldx, Y, ldy, [&](int K, const T* x, int lx, T* y, int ly) {
    switch (K) {
      case 32: launch_spmm<T, 32>(c, val, x, lx, y, ly); break;
      case 16: launch_spmm<T, 16>(c, val, x, lx, y, ly); break;
      case 8: launch_spmm<T, 8>(c, val, x, lx, y, ly); break;
      case 4: launch_spmm<T, 4>(c, val, x, lx, y, ly); break;
      case 2: launch_spmm<T, 2>(c, val, x, lx, y, ly); break;
      default: launch_spmm<T, 1>(c, val, x, lx, y, ly); break;
    }
  });
}

}  // namespace

void sigma_stored(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy) {
  PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
  if (prec == PSOCI_FP32)
    spmm_any_k<float>(c, c.d_val32.p, k, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy);
  else
    spmm_any_k<double>(c, c.d_val.p, k, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy);
}

}  // namespace psoci
