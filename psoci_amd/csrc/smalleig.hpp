// Dense symmetric eigensolver for the Davidson subspace (m <= a few hundred):
// Householder reduction to tridiagonal form, then the implicit symmetric QR
// iteration with Wilkinson shifts (Golub & Van Loan, Matrix Computations,
// algorithms 8.3.1 and 8.3.3), eigenvectors accumulated throughout.  Host
// code; O(m^3) with small constants, replacing the reference's LAPACK dsygv
// call (PsociBlasLapack.C:69-92; the basis here is orthonormal, so the
// generalised problem is a standard one).
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace psoci {

// a: m x m row-major symmetric, overwritten by the eigenvectors (column c =
// eigenvector of w[c]); w ascending.  Returns false on non-finite input or
// when the iteration does not converge.
inline bool sym_eig(int m, double* a, double* w) {
  if (m <= 0) return true;
  const size_t M = (size_t)m;
  // qt: the accumulated orthogonal transform, transposed (row r of qt is
  // column r of Q), so that rotations and reflections update contiguous rows
  std::vector<double> t(a, a + M * M), qt(M * M, 0.0), v(M), p(M), u(M);
  for (int i = 0; i < m; ++i) qt[i * M + i] = 1.0;
  for (size_t i = 0; i < M * M; ++i)
    if (!std::isfinite(t[i])) return false;
  // Householder: zero column k below the subdiagonal, k = 0 .. m-3
  for (int k = 0; k + 2 < m; ++k) {
    double sig = 0.0;
    for (int i = k + 1; i < m; ++i) sig += t[i * M + k] * t[i * M + k];
    if (sig == 0.0) continue;
    const double x1 = t[(k + 1) * M + k];
    const double alpha = x1 > 0 ? -std::sqrt(sig) : std::sqrt(sig);
    std::fill(v.begin(), v.end(), 0.0);
    for (int i = k + 1; i < m; ++i) v[i] = t[i * M + k];
    v[k + 1] -= alpha;
    double vv = 0.0;
    for (int i = k + 1; i < m; ++i) vv += v[i] * v[i];
    if (vv == 0.0) continue;
    // t <- H t H with H = I - 2 v v^T / vv: p = 2 t v / vv, q' = p - (v.p / vv) v,
    // t -= v q'^T + q' v^T (rows and columns k .. m-1; v_k = 0)
    for (int i = k; i < m; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < m; ++j) s += t[i * M + j] * v[j];
      p[i] = 2.0 * s / vv;
    }
    double vp = 0.0;
    for (int i = k + 1; i < m; ++i) vp += v[i] * p[i];
    const double kk = vp / vv;
    for (int i = k; i < m; ++i) p[i] -= kk * v[i];
    for (int i = k; i < m; ++i)
      for (int j = k; j < m; ++j) t[i * M + j] -= v[i] * p[j] + p[i] * v[j];
    // Q <- Q H, i.e. qt <- H qt: u = 2 qt^T v / vv, qt -= v u^T
    std::fill(u.begin(), u.end(), 0.0);
    for (int j = k + 1; j < m; ++j) {
      const double vj = v[j];
      const double* row = &qt[j * M];
      for (int i = 0; i < m; ++i) u[i] += vj * row[i];
    }
    for (int j = k + 1; j < m; ++j) {
      const double f = 2.0 * v[j] / vv;
      double* row = &qt[j * M];
      for (int i = 0; i < m; ++i) row[i] -= f * u[i];
    }
  }
  // clean the band: t is symmetric tridiagonal up to rounding
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      if (j < i - 1 || j > i + 1) t[i * M + j] = 0.0;
  for (int i = 0; i + 1 < m; ++i) {
    const double e = 0.5 * (t[(i + 1) * M + i] + t[i * M + i + 1]);
    t[(i + 1) * M + i] = t[i * M + i + 1] = e;
  }
  // rotation on coordinates (k, k+1): t <- G^T t G, q <- q G with
  // G = [[c, s], [-s, c]]
  auto rotate = [&](int k, double c, double s) {
    const int lo = std::max(0, k - 1), hi = std::min(m - 1, k + 2);
    for (int j = lo; j <= hi; ++j) {
      const double a0 = t[k * M + j], a1 = t[(k + 1) * M + j];
      t[k * M + j] = c * a0 - s * a1;
      t[(k + 1) * M + j] = s * a0 + c * a1;
    }
    for (int i = lo; i <= hi; ++i) {
      const double a0 = t[i * M + k], a1 = t[i * M + k + 1];
      t[i * M + k] = c * a0 - s * a1;
      t[i * M + k + 1] = s * a0 + c * a1;
    }
    double* r0 = &qt[k * M];
    double* r1 = &qt[(k + 1) * M];
    for (int i = 0; i < m; ++i) {
      const double a0 = r0[i], a1 = r1[i];
      r0[i] = c * a0 - s * a1;
      r1[i] = s * a0 + c * a1;
    }
  };
  const double eps = 2.220446049250313e-16;
  int end = m - 1, iters = 0;
  while (end > 0) {
    if (++iters > 60 * m) return false;
    auto small = [&](int i) {  // is t[i+1][i] negligible?
      return std::fabs(t[(i + 1) * M + i]) <= eps * (std::fabs(t[i * M + i]) + std::fabs(t[(i + 1) * M + i + 1]));
    };
    if (small(end - 1)) {
      t[end * M + end - 1] = t[(end - 1) * M + end] = 0.0;
      --end;
      continue;
    }
    int start = end - 1;
    while (start > 0 && !small(start - 1)) --start;
    // Wilkinson shift from the trailing 2x2 block
    const double dn = t[end * M + end], dm = t[(end - 1) * M + end - 1], en = t[end * M + end - 1];
    const double dd = 0.5 * (dm - dn);
    const double mu = dn - en * en / (dd + (dd >= 0 ? 1.0 : -1.0) * std::hypot(dd, en));
    double x = t[start * M + start] - mu, z = t[(start + 1) * M + start];
    for (int k = start; k < end; ++k) {
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : -z / r;
      rotate(k, c, s);
      if (k + 1 < end) {
        x = t[(k + 1) * M + k];
        z = t[(k + 2) * M + k];
      }
    }
  }
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  std::sort(ord.begin(), ord.end(), [&](int x, int y) { return t[x * M + x] < t[y * M + y]; });
  for (int c = 0; c < m; ++c) {
    w[c] = t[ord[c] * M + ord[c]];
    if (!std::isfinite(w[c])) return false;
    for (int r = 0; r < m; ++r) a[r * M + c] = qt[ord[c] * M + r];
  }
  return true;
}

}  // namespace psoci
