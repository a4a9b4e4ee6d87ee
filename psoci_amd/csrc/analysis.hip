// CI energy analysis on the GPU: the energy contribution of every spatial
// configuration to every root, analyzeEnergyContributions
// (PsociCIAnalysis.C:42-267), first part (:126-142):
//
//   deltae[I][r] = sum over the SEFs i of I, in order, of
//                  -c_ir^2 (H_ii - E_r) / (1 - c_ir^2),
//   reset to E_r by an SEF with c_ir^2 == 1 (the remaining SEFs of I still
//   accumulate after it, as in the reference's loop).
//
// One thread per (configuration, root): consecutive threads take the roots
// of one configuration, so each SEF's k coefficients are one coalesced
// read; the SEFs of a configuration are summed sequentially in the
// reference's order with its operation order (the build has no fused
// multiply-add), so the result is bit-identical to the reference's loop.
// The pass reads C and the diagonal once: n * (8k + 8) bytes.
#include "ctx.hpp"

namespace psoci {

namespace {

__global__ void k_energy_contrib(int64_t ncfg, int k, const int64_t* __restrict__ sefoff,
                                 const int32_t* __restrict__ nsef, const double* __restrict__ C,
                                 const double* __restrict__ diag, const double* __restrict__ evals,
                                 double* __restrict__ deltae) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t I = g / k;
  if (I >= ncfg) return;
  const int r = (int)(g - I * k);
  const double e = evals[r];
  const int64_t i0 = sefoff[I], i1 = i0 + nsef[I];
  double acc = 0.0;
  for (int64_t i = i0; i < i1; ++i) {
    const double v = C[i * k + r];
    const double v2 = v * v;
    if (v2 == 1.0) {
      acc = e;
    } else {
      double t = v2 * (diag[i] - e);
      t = t / (1.0 - v2);
      acc = acc - t;
    }
  }
  deltae[I * k + r] = acc;
}

}  // namespace

void energy_contrib(Ctx& c, int k, const double* dC, const double* d_evals, double* d_deltae) {
  const int64_t ncfg = c.sp.ncfg;
  const int64_t threads = ncfg * k;
  if (threads == 0) return;
  hipLaunchKernelGGL(k_energy_contrib, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c.stream, ncfg, k,
                     c.d_sefoff.p, c.d_nsef.p, dC, c.d_diag.p, d_evals, d_deltae);
  PSOCI_HIP(hipGetLastError());
}

}  // namespace psoci
