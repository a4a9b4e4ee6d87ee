// Dense GPU kernels used by the block Davidson (linalg.hip).
#pragma once
#include "common.hpp"

namespace psoci {

struct Linalg {
  hipStream_t stream = nullptr;
  DevBuf<double> part, small, ysmall, theta;
  DevBuf<int> isel;
  DevBuf<int64_t> ipart;
  std::vector<double> yneg;

  // partial-sum buffer of gemm_tn for (n, m, b); reserve_tn: for every
  // m <= mmax, b <= bmax (so captured launches allocate nothing)
  static size_t tn_part_size(int64_t n, int m, int b);
  void reserve_tn(int64_t n, int mmax, int bmax);
  // out (device, m x b row-major) = A[:, 0:m]^T B[:, 0:b] over n rows (FP64
  // MFMA; TA float or double, float columns widened on load)
  template <typename TA>
  void gemm_tn(int64_t n, int m, int b, const TA* A, int lda, const TA* B, int ldb, double* out_dev);
  // same, result copied to host (synchronises)
  template <typename TA>
  void gemm_tn_host(int64_t n, int m, int b, const TA* A, int lda, const TA* B, int ldb, double* out_host);
  // Out[:, 0:k] = alpha * Out + A[:, 0:m] Y1 (Y1 host, m x k row-major;
  // FP64 arithmetic, TA / TO float or double)
  template <typename TA, typename TO>
  void gemm_nn(int64_t n, int m, int k, const TA* A, int lda, const double* Y1_host, double alpha, TO* Out, int ldo);
  // Out[:, 0:k) = Out[:, 0:m) Y (m x k, host), in place (k <= m, k <= 128)
  template <typename T>
  void gemm_nn_inplace(int64_t n, int m, int k, const double* Y1_host, T* Out, int ldo);
  // Out[:, 0:k) = A[:, 0:m) Y where A and Out may be the same rows of one
  // buffer at other or overlapping columns (k <= 128)
  template <typename T>
  void gemm_nn_rows(int64_t n, int m, int k, const T* A, int lda, const double* Y1_host, T* Out, int ldo);
  // Out[:, 0:k) = A Y1 - B Y2 (Y1, Y2 host, m x k); with X set, also
  // X = B Y1 (FP64) in the same pass over A and B (synchronises)
  template <typename TA, typename TO>
  void gemm_nn2(int64_t n, int m, int k, const TA* A, int lda, const double* Y1_host, const TA* B, int ldb,
                const double* Y2_host, TO* Out, int ldo, double* X = nullptr, int ldx = 0);
  // R[:, c] /= (D - theta_c) where |D - theta_c| > dmin
  void precondition(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R, int ldr);
  // Out[:, j) = the diagonal preconditioner applied to R[:, sel[j]] (j < ns)
  template <typename TR, typename TO>
  void precondition_select(int64_t n, int k, const double* D, const double* theta_host, double dmin, const TR* R,
                           int ldr, const int* sel_host, int ns, TO* Out, int ldo);
  // the same with Olsen's correction against the Ritz vectors X:
  // R[:, c] = f (R - eps_c X) with eps_c = (x . f r) / (x . f x)
  void precondition_olsen(int64_t n, int k, const double* D, const double* theta_host, double dmin, double* R, int ldr,
                          const double* X, int ldx);
  void to_f32(int64_t n, int k, const double* A, int lda, float* B, int ldb);
  template <typename T>
  void copy_cols(int64_t n, int k, const T* A, int lda, T* B, int ldb);
  template <typename T>
  void unit_cols(int64_t n, int k, const int64_t* idx_host, T* V, int ldv);
  // idx_host[0..k) = the rows of the k lowest D entries in ascending order,
  // ties by the lower row (k <= 128; synchronises)
  void lowest(int64_t n, int k, const double* D, int64_t* idx_host);
};

// Symmetric eigen-decomposition of a small dense matrix (host, FP64):
// Householder tridiagonalisation + implicit QL.  a is m x m row-major and is
// overwritten by the eigenvectors (column c = eigenvector c); w ascending.

}  // namespace psoci
