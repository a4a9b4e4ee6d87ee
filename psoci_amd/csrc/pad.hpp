// Column-blocked sigma launches at the kernel widths instantiated.
//
// The sigma kernels are instantiated for block widths 1, 2, 4, ... maxw
// (and 24 where w24 is set).  A block of nb columns runs as ONE launch at the
// next width through a zero-padded staging copy, rather than as several
// narrower launches: a launch's cost is dominated by what does not scale
// with the width (element generation in the direct sigma, reading H in the
// stored one).  Blocks that are columns of a wider array (a stride other
// than the width, or not 16-byte aligned) are staged too: the kernels gather
// whole x rows, which then share cache lines and take 16-byte loads.
#pragma once
#include "ctx.hpp"

namespace psoci {
namespace pad {
namespace {  // each translation unit keeps its own kernels

template <typename T>
__global__ void k_pad_cols(int64_t n, int nb, int K, const T* __restrict__ X, int ldx, T* __restrict__ Xp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * K) return;
  const int64_t row = t / K;
  const int col = (int)(t - row * K);
  Xp[t] = col < nb ? X[row * ldx + col] : (T)0;
}

template <typename T>
__global__ void k_unpad_cols(int64_t n, int nb, int K, const T* __restrict__ Yp, T* __restrict__ Y, int ldy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nb) return;
  const int64_t row = t / nb;
  const int col = (int)(t - row * nb);
  Y[row * ldy + col] = Yp[row * K + col];
}

// Width of the launch that runs the block of nb columns, and whether the
// block runs in place (no staging copy).
inline int width_of(int nb, bool w24) {
  int K = 1;
  while (K < nb) K <<= 1;
  if (w24 && nb > 16 && nb <= 24) K = 24;
  return K;
}

template <typename T>
bool in_place(int K, int nb, int done, const T* X, int ldx, const T* Y, int ldy) {
  return K == nb && ldx == K && ldy == K && (uintptr_t)(X + done) % 16 == 0 && (uintptr_t)(Y + done) % 16 == 0;
}

// Largest staging area (bytes) any block of the call needs (0: none).
template <typename T>
size_t staging_bytes(int64_t n, int k, int maxw, const T* X, int ldx, const T* Y, int ldy, bool w24 = false) {
  size_t need = 0;
  for (int done = 0; done < k; done += maxw) {
    const int nb = k - done < maxw ? k - done : maxw;
    const int K = width_of(nb, w24);
    if (!in_place(K, nb, done, X, ldx, Y, ldy)) {
      const size_t b = (size_t)2 * (size_t)n * (size_t)K * sizeof(T);
      need = b > need ? b : need;
    }
  }
  return need;
}

// launch(K, x, ldx, y, ldy) runs the kernel of width K on n rows.  The
// staging copies live in the context's buffer (grown on demand) or, when
// st.p is set, in the caller's st.bytes bytes (checked for every block
// before anything is launched; nothing is allocated then).
template <typename T, class Launch>
void blocks(Ctx& c, int64_t n, int k, int maxw, const T* X, int ldx, T* Y, int ldy, Launch&& launch,
            bool w24 = false, const Staging& st = Staging{}) {
  if (st.p)
    PSOCI_CHECK(staging_bytes<T>(n, k, maxw, X, ldx, Y, ldy, w24) <= st.bytes, PSOCI_ERR_ARG,
                "the block needs more staging memory than the scratch holds");
  for (int done = 0; done < k; done += maxw) {
    const int nb = k - done < maxw ? k - done : maxw;
    const int K = width_of(nb, w24);
    if (in_place(K, nb, done, X, ldx, Y, ldy)) {
      launch(K, X + done, ldx, Y + done, ldy);
      continue;
    }
    T* Xp;
    if (st.p) {
      Xp = reinterpret_cast<T*>(st.p);
    } else {
      c.d_pad.reserve((size_t)2 * n * K * sizeof(T));
      Xp = reinterpret_cast<T*>(c.d_pad.p);
    }
    T* Yp = Xp + (size_t)n * K;
    const unsigned g1 = (unsigned)((n * K + 255) / 256), g2 = (unsigned)((n * nb + 255) / 256);
    hipLaunchKernelGGL(k_pad_cols<T>, dim3(g1), dim3(256), 0, c.stream, n, nb, K, X + done, ldx, Xp);
    PSOCI_HIP(hipGetLastError());
    launch(K, Xp, K, Yp, K);
    hipLaunchKernelGGL(k_unpad_cols<T>, dim3(g2), dim3(256), 0, c.stream, n, nb, K, Yp, Y + done, ldy);
    PSOCI_HIP(hipGetLastError());
  }
}

}  // namespace
}  // namespace pad
}  // namespace psoci
