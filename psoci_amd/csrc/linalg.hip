// Dense kernels of the block Davidson: tall-skinny products on FP64 MFMA,
// block updates, preconditioning and precision conversion.
//
// Replaces the per-vector GA dot/axpy loops of PsociGAsubspace
// (generateIncrementalSubspace, PsociGAsubspace.C:273-362; computeRnormSingle
// :874-989) and PsociGAbasis (MGS append :1371-1470; collapse :483-579;
// precondition :1860-1920).  Vectors are HBM-resident [n][ld] row-major
// blocks; every product over n is one launch.
#include <algorithm>

#include "linalg.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;
constexpr size_t kMaxLds = 48 * 1024;  // dynamic LDS per block used without opt-in
typedef double double4_t __attribute__((ext_vector_type(4)));

// C_partial[chunk][m][b] = sum over rows of chunk of A[row][a0 + i] * B[row][b0 + j]
// for one 16x16 output tile per wavefront, on v_mfma_f64_16x16x4f64.
// Lane layout of the instruction (measured on gfx950): A operand lane l =
// (row k = l/16, col i = l%16), B operand lane l = (row k = l/16,
// col j = l%16), accumulator register t of lane l = (i = 4*t + l/16, j = l%16).
__global__ __launch_bounds__(256) void k_gemm_tn_mfma(int64_t n, int m, int b, const double* __restrict__ A, int lda,
                                                      const double* __restrict__ B, int ldb, int64_t rows_per_chunk,
                                                      double* __restrict__ partial) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int tiles_m = (m + 15) / 16, tiles_b = (b + 15) / 16;
  const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
  if (tile >= tiles_m * tiles_b) return;
  const int tm = tile / tiles_b, tb = tile - tm * tiles_b;
  const int64_t chunk = blockIdx.y;
  const int64_t r0 = chunk * rows_per_chunk;
  const int64_t r1 = (r0 + rows_per_chunk < n) ? r0 + rows_per_chunk : n;
  const int ci = tm * 16 + (lane & 15);
  const int cj = tb * 16 + (lane & 15);
  const int kk = lane >> 4;
  const bool vi = ci < m, vj = cj < b;
  // four independent accumulation chains, 16 rows per step: the loads of a
  // step are all issued before its MFMAs
  double4_t acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = acc0, acc2 = acc0, acc3 = acc0;
  for (int64_t r = r0; r < r1; r += 16) {
    double a[4], bb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t row = r + 4 * u + kk;
      const bool in = row < r1;
      a[u] = (in && vi) ? A[row * lda + ci] : 0.0;
      bb[u] = (in && vj) ? B[row * ldb + cj] : 0.0;
    }
    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], bb[0], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[1], bb[1], acc1, 0, 0, 0);
    acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[2], bb[2], acc2, 0, 0, 0);
    acc3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[3], bb[3], acc3, 0, 0, 0);
  }
  const double4_t acc = (acc0 + acc1) + (acc2 + acc3);
  double* out = partial + chunk * (int64_t)m * b;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = tm * 16 + 4 * t + (lane >> 4);
    const int j = tb * 16 + (lane & 15);
    if (i < m && j < b) out[(int64_t)i * b + j] = acc[t];
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

// Out[row][0..k) = alpha * Out + sum_t A[row][t] * Y1[t][c]  (- sum_t B[row][t] * Y2[t][c])
// Y matrices [m][k] staged in LDS; one thread per (row, c).
template <bool TWO, bool LDS>
__global__ __launch_bounds__(256) void k_gemm_nn(int64_t n, int m, int k, const double* __restrict__ A, int lda,
                                                 const double* __restrict__ Y1, const double* __restrict__ B, int ldb,
                                                 const double* __restrict__ Y2, double alpha, double* __restrict__ Out,
                                                 int ldo) {
  extern __shared__ double sy[];
  const double* y1 = Y1;
  const double* y2 = Y2;
  if (LDS) {
    for (int e = threadIdx.x; e < m * k; e += blockDim.x) {
      sy[e] = Y1[e];
      if (TWO) sy[m * k + e] = Y2[e];
    }
    __syncthreads();
    y1 = sy;
    y2 = sy + m * k;
  }
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  const double* a = A + row * lda;
  double s = 0.0;
  for (int t = 0; t < m; ++t) s = __builtin_fma(a[t], y1[t * k + c], s);
  if (TWO) {
    const double* bb = B + row * ldb;
    for (int t = 0; t < m; ++t) s = __builtin_fma(-bb[t], y2[t * k + c], s);
  }
  double* o = Out + row * ldo + c;
  *o = (alpha == 0.0) ? s : __builtin_fma(alpha, *o, s);
}

// t[row][c] = r[row][c] / (D[row] - theta[c]) when |D - theta| > dmin
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

__global__ void k_copy_cols_f32_to_f64(int64_t n, int k, const float* __restrict__ A, int lda, double* __restrict__ B,
                                       int ldb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  B[row * ldb + c] = (double)A[row * lda + c];
}

__global__ void k_copy_cols(int64_t n, int k, const double* __restrict__ A, int lda, double* __restrict__ B, int ldb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  B[row * ldb + c] = A[row * lda + c];
}

__global__ void k_unit_cols(int64_t n, int k, const int64_t* __restrict__ idx, double* __restrict__ V, int ldv) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / k;
  const int c = (int)(g - row * k);
  if (row >= n) return;
  V[row * ldv + c] = (row == idx[c]) ? 1.0 : 0.0;
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void Linalg::gemm_tn(int64_t n, int m, int b, const double* A, int lda, const double* B, int ldb, double* out_dev) {
  if (m == 0 || b == 0) return;
  // about 4096 wavefronts in flight: chunks of a multiple of 16 rows
  const int tiles = ((m + 15) / 16) * ((b + 15) / 16);
  int64_t want = std::max<int64_t>(1, 4096 / tiles);
  want = std::min<int64_t>(want, 65535);
  int64_t rows_per_chunk = (n + want - 1) / want;
  rows_per_chunk = std::max<int64_t>(64, (rows_per_chunk + 15) / 16 * 16);
  int64_t nchunks = (n + rows_per_chunk - 1) / rows_per_chunk;
  part.reserve((size_t)nchunks * m * b);
  dim3 grid(nblk(tiles, 4), (unsigned)nchunks);
  hipLaunchKernelGGL(k_gemm_tn_mfma, grid, dim3(256), 0, stream, n, m, b, A, lda, B, ldb, rows_per_chunk, part.p);
  PSOCI_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)(m * b)), dim3(256), 0, stream, nchunks, m * b, part.p, out_dev);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::gemm_tn_host(int64_t n, int m, int b, const double* A, int lda, const double* B, int ldb, double* out_host) {
  small.reserve((size_t)m * b);
  gemm_tn(n, m, b, A, lda, B, ldb, small.p);
  PSOCI_HIP(hipMemcpyAsync(out_host, small.p, (size_t)m * b * sizeof(double), hipMemcpyDeviceToHost, stream));
  PSOCI_HIP(hipStreamSynchronize(stream));
}

void Linalg::gemm_nn(int64_t n, int m, int k, const double* A, int lda, const double* Y1_host, double alpha, double* Out,
                     int ldo) {
  if (k == 0) return;
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  size_t shm = (size_t)m * k * sizeof(double);
  if (shm <= kMaxLds)
    hipLaunchKernelGGL((k_gemm_nn<false, true>), dim3(nblk(n * k, 256)), dim3(256), shm, stream, n, m, k, A, lda,
                       ysmall.p, nullptr, 0, nullptr, alpha, Out, ldo);
  else
    hipLaunchKernelGGL((k_gemm_nn<false, false>), dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, m, k, A, lda,
                       ysmall.p, nullptr, 0, nullptr, alpha, Out, ldo);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::gemm_nn2(int64_t n, int m, int k, const double* A, int lda, const double* Y1_host, const double* B, int ldb,
                      const double* Y2_host, double* Out, int ldo) {
  ysmall.reserve((size_t)2 * m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  PSOCI_HIP(hipMemcpyAsync(ysmall.p + (size_t)m * k, Y2_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice,
                           stream));
  size_t shm = (size_t)2 * m * k * sizeof(double);
  if (shm <= kMaxLds)
    hipLaunchKernelGGL((k_gemm_nn<true, true>), dim3(nblk(n * k, 256)), dim3(256), shm, stream, n, m, k, A, lda,
                       ysmall.p, B, ldb, ysmall.p + (size_t)m * k, 0.0, Out, ldo);
  else
    hipLaunchKernelGGL((k_gemm_nn<true, false>), dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, m, k, A, lda,
                       ysmall.p, B, ldb, ysmall.p + (size_t)m * k, 0.0, Out, ldo);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::precondition(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R, int ldr) {
  theta.reserve(k);
  PSOCI_HIP(hipMemcpyAsync(theta.p, theta_host, k * sizeof(double), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_precondition, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, D, theta.p, dmin, R, ldr);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::to_f32(int64_t n, int k, const double* A, int lda, float* B, int ldb) {
  hipLaunchKernelGGL(k_copy_cols_f64_to_f32, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, A, lda, B, ldb);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::to_f64(int64_t n, int k, const float* A, int lda, double* B, int ldb) {
  hipLaunchKernelGGL(k_copy_cols_f32_to_f64, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, A, lda, B, ldb);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::copy_cols(int64_t n, int k, const double* A, int lda, double* B, int ldb) {
  hipLaunchKernelGGL(k_copy_cols, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, A, lda, B, ldb);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::unit_cols(int64_t n, int k, const int64_t* idx_host, double* V, int ldv) {
  DevBuf<int64_t> idx;
  idx.upload(idx_host, k, stream);
  hipLaunchKernelGGL(k_unit_cols, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, idx.p, V, ldv);
  PSOCI_HIP(hipGetLastError());
  PSOCI_HIP(hipStreamSynchronize(stream));
}

}  // namespace psoci
