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
// before it writes any output of them).
template <int NCT, bool TWO, bool XOUT, bool INPLACE = false>
__global__ __launch_bounds__(256) void k_gemm_nn_mfma(int64_t n, int m, int k, const double* __restrict__ A,
                                                      int lda, const double* __restrict__ B, int ldb,
                                                      const double* __restrict__ Y1, const double* __restrict__ Y2,
                                                      double alpha, double* Out, int ldo, double* __restrict__ Out2,
                                                      int ldo2) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t r0 = ((int64_t)blockIdx.x * (blockDim.x / kWave) + wave) * 16;
  if (r0 >= n) return;  // uniform per wavefront
  const int cb = blockIdx.y * 128;
  const int i = lane & 15, q = lane >> 4;
  const int64_t row = r0 + i;
  const bool rv = row < n;
  const double* a_row = INPLACE ? Out + (rv ? row : 0) * (int64_t)ldo : A + (rv ? row : 0) * (int64_t)lda;
  const double* __restrict__ b_row = TWO ? B + (rv ? row : 0) * (int64_t)ldb : nullptr;
  double4_t acc[NCT], acc2[NCT];
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
    acc[ct] = double4_t{0.0, 0.0, 0.0, 0.0};
    acc2[ct] = acc[ct];
  }
  for (int t0 = 0; t0 < m; t0 += 16) {
    const int tq = t0 + 4 * q;
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool tv = rv && tq + u < m;
      a[u] = tv ? a_row[tq + u] : 0.0;
      b[u] = (TWO && tv) ? b_row[tq + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = tq + u;
#pragma unroll
      for (int ct = 0; ct < NCT; ++ct) {
        const int c = cb + ct * 16 + i;
        const bool yv = t < m && c < k;
        const double y1 = yv ? Y1[(int64_t)t * k + c] : 0.0;
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], y1, acc[ct], 0, 0, 0);
        if (XOUT) acc2[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[u], y1, acc2[ct], 0, 0, 0);
        if (TWO) {
          const double y2 = yv ? Y2[(int64_t)t * k + c] : 0.0;
          acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[u], y2, acc[ct], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
    const int c = cb + ct * 16 + i;
    if (c >= k) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t rr = r0 + 4 * j + q;
      if (rr >= n) continue;
      double* o = Out + rr * ldo + c;
      *o = (alpha == 0.0) ? acc[ct][j] : __builtin_fma(alpha, *o, acc[ct][j]);
      if (XOUT) Out2[rr * ldo2 + c] = acc2[ct][j];
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
__global__ void k_precondition_select(int64_t n, int ns, const double* __restrict__ D,
                                      const double* __restrict__ theta_sel, const int* __restrict__ sel,
                                      double dmin, const double* __restrict__ R, int ldr, double* __restrict__ Out,
                                      int ldo) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = g / ns;
  const int j = (int)(g - row * ns);
  if (row >= n) return;
  const double r = R[row * ldr + sel[j]];
  const double den = D[row] - theta_sel[j];
  Out[row * ldo + j] = fabs(den) > dmin ? r / den : r;
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

namespace {

// about 4096 wavefronts in flight: chunks of a multiple of 16 rows
void tn_chunks(int64_t n, int m, int b, int64_t& rows_per_chunk, int64_t& nchunks) {
  const int tiles = ((m + 15) / 16) * ((b + 15) / 16);
  int64_t want = std::max<int64_t>(1, 4096 / tiles);
  want = std::min<int64_t>(want, 65535);
  rows_per_chunk = (n + want - 1) / want;
  rows_per_chunk = std::max<int64_t>(64, (rows_per_chunk + 15) / 16 * 16);
  nchunks = (n + rows_per_chunk - 1) / rows_per_chunk;
}

}  // namespace

size_t Linalg::tn_part_size(int64_t n, int m, int b) {
  int64_t rpc, nch;
  tn_chunks(n, m, b, rpc, nch);
  return (size_t)nch * m * b;
}

void Linalg::reserve_tn(int64_t n, int mmax, int bmax) {
  size_t need = 0;
  for (int m = 1; m <= mmax; ++m)
    for (int b = 1; b <= bmax; ++b) need = std::max(need, tn_part_size(n, m, b));
  part.reserve(need);
}

void Linalg::gemm_tn(int64_t n, int m, int b, const double* A, int lda, const double* B, int ldb, double* out_dev) {
  if (m == 0 || b == 0) return;
  int64_t rows_per_chunk, nchunks;
  tn_chunks(n, m, b, rows_per_chunk, nchunks);
  part.reserve((size_t)nchunks * m * b);
  const int tiles = ((m + 15) / 16) * ((b + 15) / 16);
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

namespace {

template <bool TWO, bool XOUT, bool INPLACE = false>
void launch_nn(hipStream_t st, int64_t n, int m, int k, const double* A, int lda, const double* B, int ldb,
               const double* Y1, const double* Y2, double alpha, double* Out, int ldo, double* Out2, int ldo2) {
  const dim3 grid(nblk(n, 64), (unsigned)((k + 127) / 128));
  const int nct = std::min(8, (k + 15) / 16);
#define PSOCI_NN(NCT)                                                                                      \
  hipLaunchKernelGGL((k_gemm_nn_mfma<NCT, TWO, XOUT, INPLACE>), grid, dim3(256), 0, st, n, m, k, A, lda, B, \
                     ldb, Y1, Y2, alpha, Out, ldo, Out2, ldo2)
  if (nct <= 1) PSOCI_NN(1);
  else if (nct <= 2) PSOCI_NN(2);
  else if (nct <= 4) PSOCI_NN(4);
  else PSOCI_NN(8);
#undef PSOCI_NN
  PSOCI_HIP(hipGetLastError());
}

}  // namespace

void Linalg::gemm_nn(int64_t n, int m, int k, const double* A, int lda, const double* Y1_host, double alpha, double* Out,
                     int ldo) {
  if (k == 0 || n == 0) return;
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  launch_nn<false, false>(stream, n, m, k, A, lda, nullptr, 0, ysmall.p, nullptr, alpha, Out, ldo, nullptr, 0);
}

void Linalg::gemm_nn_inplace(int64_t n, int m, int k, const double* Y1_host, double* Out, int ldo) {
  if (k == 0 || n == 0) return;
  PSOCI_CHECK(k <= 128 && k <= m, PSOCI_ERR_ARG, "in-place product: at most 128 and at most m output columns");
  ysmall.reserve((size_t)m * k);
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  launch_nn<false, false, true>(stream, n, m, k, nullptr, 0, nullptr, 0, ysmall.p, nullptr, 0.0, Out, ldo, nullptr,
                                0);
}

void Linalg::gemm_nn2(int64_t n, int m, int k, const double* A, int lda, const double* Y1_host, const double* B, int ldb,
                      const double* Y2_host, double* Out, int ldo, double* X, int ldx) {
  if (k == 0 || n == 0) return;
  ysmall.reserve((size_t)2 * m * k);
  yneg.assign(Y2_host, Y2_host + (size_t)m * k);
  for (auto& v : yneg) v = -v;
  PSOCI_HIP(hipMemcpyAsync(ysmall.p, Y1_host, (size_t)m * k * sizeof(double), hipMemcpyHostToDevice, stream));
  PSOCI_HIP(hipMemcpyAsync(ysmall.p + (size_t)m * k, yneg.data(), (size_t)m * k * sizeof(double),
                           hipMemcpyHostToDevice, stream));
  if (X)
    launch_nn<true, true>(stream, n, m, k, A, lda, B, ldb, ysmall.p, ysmall.p + (size_t)m * k, 0.0, Out, ldo, X, ldx);
  else
    launch_nn<true, false>(stream, n, m, k, A, lda, B, ldb, ysmall.p, ysmall.p + (size_t)m * k, 0.0, Out, ldo,
                           nullptr, 0);
  // yneg is read by the copy above: keep it until the stream has consumed it
  PSOCI_HIP(hipStreamSynchronize(stream));
}

void Linalg::precondition(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R, int ldr) {
  theta.reserve(k);
  PSOCI_HIP(hipMemcpyAsync(theta.p, theta_host, k * sizeof(double), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_precondition, dim3(nblk(n * k, 256)), dim3(256), 0, stream, n, k, D, theta.p, dmin, R, ldr);
  PSOCI_HIP(hipGetLastError());
}

void Linalg::precondition_select(int64_t n, int k, const double* D, const double* theta_host, double dmin,
                                 const double* R, int ldr, const int* sel_host, int ns, double* Out, int ldo) {
  if (ns == 0 || n == 0) return;
  std::vector<double> ts(ns);
  for (int j = 0; j < ns; ++j) ts[j] = theta_host[sel_host[j]];
  (void)k;
  theta.reserve(ns);
  isel.reserve(ns);
  PSOCI_HIP(hipMemcpyAsync(theta.p, ts.data(), ns * sizeof(double), hipMemcpyHostToDevice, stream));
  PSOCI_HIP(hipMemcpyAsync(isel.p, sel_host, ns * sizeof(int), hipMemcpyHostToDevice, stream));
  hipLaunchKernelGGL(k_precondition_select, dim3(nblk(n * ns, 256)), dim3(256), 0, stream, n, ns, D, theta.p,
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
