// Hamiltonian construction on the GPU: pair screening, configuration-pair
// setup, SEF element generation and deterministic compaction into CSR.
//
// Replaces constructGlobalHamiltonianAltOrbMapDirectChunkJmap with
// packSefsef (PsociGAhamiltonian.C:2530-2857, 4077-4136).  The reference's
// row for SEF i of configuration I lists, for J <= I ascending (one J chunk),
// the elements of block (I, J) row i with |h| >= 1e-8 (and the diagonal),
// columns ascending; here every (row, pair) segment gets a slot, a prefix
// sum over the slots gives each segment's CSR offset, and a fill pass writes
// the elements -- no atomics, so the result is bit-for-bit reproducible.
//
// The context keeps the symmetric full matrix: row r holds its lower part
// (the reference's row, diagonal last) followed by the transposed elements of
// blocks (I, J) with I > J, ascending.
#include <chrono>
#include <hipcub/hipcub.hpp>

#include "ctx.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;
constexpr double kThresh = 1.0e-8;  // MIN_DOUBLE_VALUE (PsociHamiltonian.hpp:37)

// the configuration's determinant patterns come from its class's family
// (12 offsets per family)
__device__ __forceinline__ ConfView conf_view(int64_t c, const uint64_t* dm, const uint64_t* sm, const int8_t* nopen,
                                              const uint8_t* cph, const uint16_t* pat, const int32_t* patoff,
                                              const uint8_t* cls) {
  ConfView v;
  v.d = dm[c];
  v.s = sm[c];
  v.nopen = nopen[c];
  v.ph = cph[c];
  v.pat = pat + patoff[12 * cls[c] + v.nopen];
  return v;
}

// One wavefront per configuration I: ballot over candidate J, ascending.
// upper == 0: J <= I (the reference's J list, compareLocalJSubsetOrbMap,
// PsociGADeterminants.C:2501-2546); upper == 1: J > I.
template <bool FILL>
__global__ void k_screen(int64_t ncfg, const uint64_t* __restrict__ dm, const uint64_t* __restrict__ sm,
                         const int8_t* __restrict__ nopen, int upper, int32_t* __restrict__ cnt,
                         const int64_t* __restrict__ off, int32_t* __restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= ncfg) return;  // uniform per wavefront
  const uint64_t dI = dm[I], sI = sm[I];
  const int nI = nopen[I];
  const int64_t lo = upper ? I + 1 : 0, hi = upper ? ncfg : I + 1;
  int64_t base = FILL ? off[I] : 0;
  int c = 0;
  for (int64_t J0 = lo; J0 < hi; J0 += kWave) {
    const int64_t J = J0 + lane;
    bool ok = false;
    if (J < hi) {
      int dn = nI - nopen[J];
      if (dn <= 4 && dn >= -4) ok = moved(dI, sI, dm[J], sm[J]) <= 2;
    }
    const uint64_t m = __ballot(ok);
    if (FILL && ok) out[base + c + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)J;
    c += __popcll(m);
  }
  if (!FILL && lane == 0) cnt[I] = c;
}

__global__ void k_i32_to_i64(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// pair -> row configuration (wavefront per I)
__global__ void k_expand_rows(int64_t ncfg, const int64_t* __restrict__ off, int32_t* __restrict__ rowcfg) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= ncfg) return;
  for (int64_t t = off[I] + lane; t < off[I + 1]; t += kWave) rowcfg[t] = (int32_t)I;
}

__global__ void k_pair_setup(int64_t npairs, const int32_t* __restrict__ rowcfg, const int32_t* __restrict__ colcfg,
                             const uint64_t* __restrict__ dm, const uint64_t* __restrict__ sm,
                             const int8_t* __restrict__ nopen, const int32_t* __restrict__ jsym, int nbf,
                             const double* __restrict__ one, const double* __restrict__ two,
                             PairInfo* __restrict__ pairs) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int I = rowcfg[p], J = colcfg[p];
  PairInfo P;
  pair_setup(P, I, J, dm[I], sm[I], nopen[I], dm[J], sm[J], nopen[J], nbf, one, two, jsym[I], jsym[J]);
  pairs[p] = P;
}

__global__ void k_block_sizes(int64_t npairs, const PairInfo* __restrict__ pairs, const int32_t* __restrict__ nsef,
                              int64_t* __restrict__ sz) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  sz[p] = (int64_t)nsef[pairs[p].I] * nsef[pairs[p].J];
}

__device__ __forceinline__ int64_t upper_bound_i64(const int64_t* __restrict__ a, int64_t n, int64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One thread per block element: staging[blk_off[p] + i*nsef_J + j].
__global__ void k_block_elems(int64_t total, int64_t npairs, const int64_t* __restrict__ blk_off,
                              const PairInfo* __restrict__ pairs, const uint64_t* __restrict__ dm,
                              const uint64_t* __restrict__ sm, const int8_t* __restrict__ nopen,
                              const uint8_t* __restrict__ cph, const uint16_t* __restrict__ pat,
                              const int32_t* __restrict__ patoff, const uint8_t* __restrict__ cls,
                              const double* __restrict__ cmult, const int32_t* __restrict__ nsef, int frac,
                              const double* __restrict__ two, double* __restrict__ staging) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  int64_t p = upper_bound_i64(blk_off, npairs, e) - 1;
  const PairInfo P = pairs[p];
  double v = 0.0;
  if (P.kind != 0) {
    int64_t loc = e - blk_off[p];
    int nj = nsef[P.J];
    int i = (int)(loc / nj), j = (int)(loc - (int64_t)i * nj);
    ConfView A = conf_view(P.I, dm, sm, nopen, cph, pat, patoff, cls);
    ConfView B = conf_view(P.J, dm, sm, nopen, cph, pat, patoff, cls);
    v = sef_elem(P, A, B, i, j, cmult[P.I], cmult[P.J], two, frac != 0);
  }
  staging[e] = v;
}

struct SlotMap {
  int64_t J;   // configuration of the row
  int j;       // SEF within it
  int t;       // pair slot within the row
  int w;       // slots per row of J
  int nl;      // lower slots per row of J
};

__device__ __forceinline__ SlotMap slot_map(int64_t s, int64_t ncfg, const int64_t* __restrict__ slotbase,
                                            const int32_t* __restrict__ nl, const int32_t* __restrict__ nu) {
  SlotMap m;
  m.J = upper_bound_i64(slotbase, ncfg, s) - 1;
  m.nl = nl[m.J];
  m.w = m.nl + 1 + nu[m.J];
  int64_t rem = s - slotbase[m.J];
  m.j = (int)(rem / m.w);
  m.t = (int)(rem - (int64_t)m.j * m.w);
  return m;
}

__device__ __forceinline__ int64_t find_pair(const int32_t* __restrict__ colcfg, int64_t lo, int64_t hi, int32_t J) {
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (colcfg[mid] < J) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// count (FILL = false) or write (FILL = true) the kept elements of one
// (row, pair) segment of the full matrix
template <bool FILL>
__global__ void k_segments(int64_t nslots, int64_t ncfg, const int64_t* __restrict__ slotbase,
                           const int32_t* __restrict__ nl, const int32_t* __restrict__ nu,
                           const int64_t* __restrict__ offL, const int32_t* __restrict__ colL,
                           const int64_t* __restrict__ offU, const int32_t* __restrict__ rowU,
                           const PairInfo* __restrict__ pairs, const int64_t* __restrict__ blk_off,
                           const int32_t* __restrict__ nsef, const int64_t* __restrict__ sefoff,
                           const double* __restrict__ staging, int64_t* __restrict__ cnt,
                           const int64_t* __restrict__ pos, int32_t* __restrict__ col, double* __restrict__ val,
                           int64_t* __restrict__ rowptr, int32_t* __restrict__ lowcnt) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  SlotMap m = slot_map(s, ncfg, slotbase, nl, nu);
  int64_t at = FILL ? pos[s] : 0;
  int64_t c = 0;
  if (m.t < m.nl) {
    int64_t p = offL[m.J] + m.t;
    const PairInfo& P = pairs[p];
    if (P.kind != 0) {
      int nj = nsef[P.J];
      bool diagblk = (P.J == P.I);
      int jmax = diagblk ? m.j : nj - 1;
      const double* row = staging + blk_off[p] + (int64_t)m.j * nj;
      int64_t cbase = sefoff[P.J];
      for (int q = 0; q <= jmax; ++q) {
        double v = row[q];
        if (fabs(v) >= kThresh || (diagblk && q == m.j)) {
          if (FILL) {
            col[at + c] = (int32_t)(cbase + q);
            val[at + c] = v;
          }
          ++c;
        }
      }
    }
  } else if (m.t == m.nl) {
    // upper triangle of the diagonal block, mirrored from its lower triangle
    // (column j, rows q > j) so the full matrix is bitwise symmetric
    int64_t p = offL[m.J] + m.nl - 1;
    int nj = nsef[m.J];
    const double* blk = staging + blk_off[p] + m.j;
    int64_t cbase = sefoff[m.J];
    for (int q = m.j + 1; q < nj; ++q) {
      double v = blk[(int64_t)q * nj];
      if (fabs(v) >= kThresh) {
        if (FILL) {
          col[at + c] = (int32_t)(cbase + q);
          val[at + c] = v;
        }
        ++c;
      }
    }
  } else {
    int32_t I = rowU[offU[m.J] + (m.t - m.nl - 1)];
    int64_t p = find_pair(colL, offL[I], offL[I + 1], (int32_t)m.J);
    // the predicate is symmetric, so (I, J) is in I's lower list; guard anyway
    const bool found = p < offL[I + 1] && colL[p] == (int32_t)m.J;
    if (found && pairs[p].kind != 0) {
      int ni = nsef[I], nj = nsef[m.J];
      const double* blk = staging + blk_off[p] + m.j;
      int64_t cbase = sefoff[I];
      for (int q = 0; q < ni; ++q) {
        double v = blk[(int64_t)q * nj];
        if (fabs(v) >= kThresh) {
          if (FILL) {
            col[at + c] = (int32_t)(cbase + q);
            val[at + c] = v;
          }
          ++c;
        }
      }
    }
  }
  if (!FILL) cnt[s] = c;
  if (FILL && m.t == 0) {
    int64_t r = sefoff[m.J] + m.j;
    rowptr[r] = at;
    // lower part of this row: the first nl segments
    lowcnt[r] = (int32_t)(pos[s + m.nl] - at);
  }
}

__global__ void k_diag_from_rows(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ lowcnt,
                                 const double* __restrict__ val, double* __restrict__ diag) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  diag[r] = val[rowptr[r] + lowcnt[r] - 1];
}

__global__ void k_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

// diagonal only: computeOnlyDiagElements (PsociGAhamiltonian.C:13171-13250)
__global__ void k_diag_only(int64_t ncfg, const uint64_t* __restrict__ dm, const uint64_t* __restrict__ sm,
                            const int8_t* __restrict__ nopen, const uint8_t* __restrict__ cph,
                            const uint16_t* __restrict__ pat, const int32_t* __restrict__ patoff,
                            const uint8_t* __restrict__ cls, const double* __restrict__ cmult,
                            const int32_t* __restrict__ nsef, const int64_t* __restrict__ sefoff, int nbf, int frac,
                            const double* __restrict__ one, const double* __restrict__ two,
                            double* __restrict__ diag) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= ncfg) return;
  PairInfo P;
  if (lane == 0) pair_setup(P, (int)I, (int)I, dm[I], sm[I], nopen[I], dm[I], sm[I], nopen[I], nbf, one, two);
  P.v0 = __shfl(P.v0, 0, kWave);
  P.kind = 1;
  P.I = P.J = (int)I;
  ConfView A = conf_view(I, dm, sm, nopen, cph, pat, patoff, cls);
  const double mult = cmult[I];
  for (int i = lane; i < nsef[I]; i += kWave) diag[sefoff[I] + i] = sef_elem(P, A, A, i, i, mult, mult, two, frac != 0);
}

__global__ void k_single_block(PairInfo* __restrict__ Pout, int I, int J, const uint64_t* __restrict__ dm,
                               const uint64_t* __restrict__ sm, const int8_t* __restrict__ nopen,
                               const uint8_t* __restrict__ cph, const uint16_t* __restrict__ pat,
                               const int32_t* __restrict__ patoff, const uint8_t* __restrict__ cls,
                               const double* __restrict__ cmult, const int32_t* __restrict__ jsym,
                               const int32_t* __restrict__ nsef, int nbf, int frac, const double* __restrict__ one,
                               const double* __restrict__ two, double* __restrict__ out) {
  __shared__ PairInfo P;
  if (threadIdx.x == 0) {
    pair_setup(P, I, J, dm[I], sm[I], nopen[I], dm[J], sm[J], nopen[J], nbf, one, two, jsym[I], jsym[J]);
    *Pout = P;
  }
  __syncthreads();
  ConfView A = conf_view(I, dm, sm, nopen, cph, pat, patoff, cls);
  ConfView B = conf_view(J, dm, sm, nopen, cph, pat, patoff, cls);
  int ni = nsef[I], nj = nsef[J];
  for (int e = threadIdx.x; e < ni * nj; e += blockDim.x) {
    int i = e / nj, j = e - i * nj;
    out[e] = (P.kind != 0) ? sef_elem(P, A, B, i, j, cmult[I], cmult[J], two, frac != 0) : 0.0;
  }
}

__global__ void k_extract_lower(int64_t n, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ lowcnt,
                                const int64_t* __restrict__ lowptr, const int32_t* __restrict__ col,
                                const double* __restrict__ val, int32_t* __restrict__ lcol, double* __restrict__ lval) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t r = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (r >= n) return;
  int64_t a = rowptr[r], b = lowptr[r];
  for (int q = lane; q < lowcnt[r]; q += kWave) {
    lcol[b + q] = col[a + q];
    lval[b + q] = val[a + q];
  }
}

inline unsigned blocks_for(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s, void*& tmp, size_t& tmp_bytes) {
  if (n <= 0) return;
  size_t need = 0;
  PSOCI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)0 + n, s));
  if (need > tmp_bytes) {
    if (tmp) PSOCI_HIP(hipFree(tmp));
    tmp = nullptr;
    PSOCI_HIP(hipMalloc(&tmp, need));
    tmp_bytes = need;
  }
  PSOCI_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, n, s));
}

void upload_space(Ctx& c) {
  const Space& sp = c.sp;
  c.d_dm.upload(sp.dm, c.stream);
  c.d_sm.upload(sp.sm, c.stream);
  c.d_nopen.upload(sp.nopen, c.stream);
  c.d_nsef.upload(sp.nsef, c.stream);
  c.d_sefoff.upload(sp.sef_off, c.stream);
  c.d_cph.upload(sp.cphase, c.stream);
  c.d_pat.upload(sp.pat, c.stream);
  std::vector<int32_t> po(sp.pat_off, sp.pat_off + 24);
  c.d_patoff.upload(po, c.stream);
  c.d_didx.upload(sp.didx, c.stream);
  c.d_cls.upload(sp.cls, c.stream);
  c.d_cmult.upload(sp.cmult, c.stream);
  c.d_jsym.upload(sp.jsym, c.stream);
  c.sync();
}

namespace {

// pair lists: lower (J <= I per I) and, when wanted, upper (I > J per J)
struct PairLists {
  DevBuf<int32_t> nl, nu;
  DevBuf<int64_t> offL, offU;
  DevBuf<int32_t> colL, rowU, rowcfg;
  DevBuf<PairInfo> info;
  int64_t nL = 0, nU = 0;
};

void screen(Ctx& c, int upper, DevBuf<int32_t>& cnt, DevBuf<int64_t>& off, DevBuf<int32_t>& list, int64_t& total) {
  const int64_t ncfg = c.sp.ncfg;
  cnt.alloc(ncfg);
  off.alloc(ncfg + 1);
  const int wpb = 4;  // wavefronts per 256-thread block
  hipLaunchKernelGGL(k_screen<false>, dim3(blocks_for(ncfg, wpb)), dim3(256), 0, c.stream, ncfg, c.d_dm.p, c.d_sm.p,
                     c.d_nopen.p, upper, cnt.p, nullptr, nullptr);
  PSOCI_HIP(hipGetLastError());
  DevBuf<int64_t> cnt64;
  cnt64.alloc(ncfg + 1);
  PSOCI_HIP(hipMemsetAsync(cnt64.p, 0, (ncfg + 1) * sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_i32_to_i64, dim3(blocks_for(ncfg, 256)), dim3(256), 0, c.stream, ncfg, cnt.p, cnt64.p);
  c.scan(cnt64.p, off.p, ncfg + 1);
  PSOCI_HIP(hipMemcpyAsync(&total, off.p + ncfg, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  list.alloc(total > 0 ? total : 1);
  hipLaunchKernelGGL(k_screen<true>, dim3(blocks_for(ncfg, wpb)), dim3(256), 0, c.stream, ncfg, c.d_dm.p, c.d_sm.p,
                     c.d_nopen.p, upper, nullptr, off.p, list.p);
  PSOCI_HIP(hipGetLastError());
}

void lower_pairs(Ctx& c, PairLists& L) {
  const int64_t ncfg = c.sp.ncfg;
  screen(c, 0, L.nl, L.offL, L.colL, L.nL);
  L.rowcfg.alloc(L.nL);
  hipLaunchKernelGGL(k_expand_rows, dim3(blocks_for(ncfg, 4)), dim3(256), 0, c.stream, ncfg, L.offL.p, L.rowcfg.p);
  PSOCI_HIP(hipGetLastError());
  L.info.alloc(L.nL);
  hipLaunchKernelGGL(k_pair_setup, dim3(blocks_for(L.nL, 256)), dim3(256), 0, c.stream, L.nL, L.rowcfg.p, L.colL.p,
                     c.d_dm.p, c.d_sm.p, c.d_nopen.p, c.d_jsym.p, c.sp.nbf, c.d_one.p, c.d_two.p, L.info.p);
  PSOCI_HIP(hipGetLastError());
}

}  // namespace

void build_h(Ctx& c) {
  const Space& sp = c.sp;
  const int64_t ncfg = sp.ncfg, n = sp.nsef_total;
  const auto t0 = std::chrono::steady_clock::now();
  PairLists L;
  lower_pairs(c, L);
  screen(c, 1, L.nu, L.offU, L.rowU, L.nU);
  const int64_t np = L.nL;
  // block staging
  DevBuf<int64_t> bsz, boff;
  bsz.alloc(np + 1);
  boff.alloc(np + 1);
  PSOCI_HIP(hipMemsetAsync(bsz.p + np, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_block_sizes, dim3(blocks_for(np, 256)), dim3(256), 0, c.stream, np, L.info.p, c.d_nsef.p,
                     bsz.p);
  PSOCI_HIP(hipGetLastError());
  c.scan(bsz.p, boff.p, np + 1);
  int64_t nstage = 0;
  PSOCI_HIP(hipMemcpyAsync(&nstage, boff.p + np, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  DevBuf<double> staging;
  staging.alloc(nstage);
  hipLaunchKernelGGL(k_block_elems, dim3(blocks_for(nstage, 256)), dim3(256), 0, c.stream, nstage, np, boff.p,
                     L.info.p, c.d_dm.p, c.d_sm.p, c.d_nopen.p, c.d_cph.p, c.d_pat.p, c.d_patoff.p, c.d_cls.p,
                     c.d_cmult.p, c.d_nsef.p, sp.frac ? 1 : 0, c.d_two.p, staging.p);
  PSOCI_HIP(hipGetLastError());
  // slots: per row of J, nl(J) lower segments, the diagonal block's upper
  // triangle, then nu(J) transposed segments
  std::vector<int32_t> hnl(ncfg), hnu(ncfg);
  L.nl.download(hnl.data(), ncfg, c.stream);
  L.nu.download(hnu.data(), ncfg, c.stream);
  c.sync();
  std::vector<int64_t> slotbase(ncfg + 1);
  int64_t ns = 0;
  for (int64_t J = 0; J < ncfg; ++J) {
    slotbase[J] = ns;
    ns += (int64_t)sp.nsef[J] * (hnl[J] + 1 + hnu[J]);
  }
  slotbase[ncfg] = ns;
  DevBuf<int64_t> d_slotbase, cnt, pos;
  d_slotbase.upload(slotbase, c.stream);
  cnt.alloc(ns + 1);
  pos.alloc(ns + 1);
  PSOCI_HIP(hipMemsetAsync(cnt.p + ns, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_segments<false>, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, ncfg, d_slotbase.p,
                     L.nl.p, L.nu.p, L.offL.p, L.colL.p, L.offU.p, L.rowU.p, L.info.p, boff.p, c.d_nsef.p,
                     c.d_sefoff.p, staging.p, cnt.p, nullptr, nullptr, nullptr, nullptr, nullptr);
  PSOCI_HIP(hipGetLastError());
  c.scan(cnt.p, pos.p, ns + 1);
  int64_t nnz = 0;
  PSOCI_HIP(hipMemcpyAsync(&nnz, pos.p + ns, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  c.d_rowptr.alloc(n + 1);
  c.d_col.alloc(nnz);
  c.d_val.alloc(nnz);
  c.d_lowcnt.alloc(n);
  PSOCI_HIP(hipMemcpyAsync(c.d_rowptr.p + n, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_segments<true>, dim3(blocks_for(ns, 256)), dim3(256), 0, c.stream, ns, ncfg, d_slotbase.p,
                     L.nl.p, L.nu.p, L.offL.p, L.colL.p, L.offU.p, L.rowU.p, L.info.p, boff.p, c.d_nsef.p,
                     c.d_sefoff.p, staging.p, nullptr, pos.p, c.d_col.p, c.d_val.p, c.d_rowptr.p, c.d_lowcnt.p);
  PSOCI_HIP(hipGetLastError());
  c.d_diag.alloc(n);
  hipLaunchKernelGGL(k_diag_from_rows, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, n, c.d_rowptr.p,
                     c.d_lowcnt.p, c.d_val.p, c.d_diag.p);
  c.d_val32.alloc(nnz);
  hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(nnz, 256)), dim3(256), 0, c.stream, nnz, c.d_val.p, c.d_val32.p);
  PSOCI_HIP(hipGetLastError());
  // the reference's element count: the lower parts
  c.nnz_full = nnz;
  c.nnz_lower = (nnz + n) / 2;
  c.sync();
  c.have_h = true;
  const auto t1 = std::chrono::steady_clock::now();
  tune_stored(c);
  c.have_diag = true;
  c.hbuild_seconds[0] = std::chrono::duration<double>(t1 - t0).count();
  c.hbuild_seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
}

void compute_diag(Ctx& c) {
  const Space& sp = c.sp;
  c.d_diag.alloc(sp.nsef_total);
  hipLaunchKernelGGL(k_diag_only, dim3(blocks_for(sp.ncfg, 4)), dim3(256), 0, c.stream, sp.ncfg, c.d_dm.p, c.d_sm.p,
                     c.d_nopen.p, c.d_cph.p, c.d_pat.p, c.d_patoff.p, c.d_cls.p, c.d_cmult.p, c.d_nsef.p,
                     c.d_sefoff.p, sp.nbf, sp.frac ? 1 : 0, c.d_one.p, c.d_two.p, c.d_diag.p);
  PSOCI_HIP(hipGetLastError());
  c.sync();
  c.have_diag = true;
}

void block_on_device(Ctx& c, int64_t I, int64_t J, double* host_out, int* status) {
  const Space& sp = c.sp;
  int ni = sp.nsef[I], nj = sp.nsef[J];
  DevBuf<double> out;
  DevBuf<PairInfo> P;
  out.alloc((size_t)ni * nj);
  P.alloc(1);
  hipLaunchKernelGGL(k_single_block, dim3(1), dim3(256), 0, c.stream, P.p, (int)I, (int)J, c.d_dm.p, c.d_sm.p,
                     c.d_nopen.p, c.d_cph.p, c.d_pat.p, c.d_patoff.p, c.d_cls.p, c.d_cmult.p, c.d_jsym.p,
                     c.d_nsef.p, sp.nbf, sp.frac ? 1 : 0, c.d_one.p, c.d_two.p, out.p);
  PSOCI_HIP(hipGetLastError());
  PairInfo hp;
  out.download(host_out, (size_t)ni * nj, c.stream);
  P.download(&hp, 1, c.stream);
  c.sync();
  if (status) *status = hp.kind != 0 ? 1 : 0;
}

void get_h_lower(Ctx& c, int64_t* rowptr, int32_t* col, double* val, double* diag) {
  const int64_t n = c.sp.nsef_total;
  std::vector<int32_t> lc(n);
  c.d_lowcnt.download(lc.data(), n, c.stream);
  c.sync();
  std::vector<int64_t> lp(n + 1);
  int64_t acc = 0;
  for (int64_t r = 0; r < n; ++r) {
    lp[r] = acc;
    acc += lc[r];
  }
  lp[n] = acc;
  if (rowptr) std::copy(lp.begin(), lp.end(), rowptr);
  if (col || val) {
    DevBuf<int64_t> d_lp;
    DevBuf<int32_t> lcol;
    DevBuf<double> lval;
    d_lp.upload(lp, c.stream);
    lcol.alloc(acc);
    lval.alloc(acc);
    hipLaunchKernelGGL(k_extract_lower, dim3(blocks_for(n, 4)), dim3(256), 0, c.stream, n, c.d_rowptr.p,
                       c.d_lowcnt.p, d_lp.p, c.d_col.p, c.d_val.p, lcol.p, lval.p);
    PSOCI_HIP(hipGetLastError());
    if (col) lcol.download(col, acc, c.stream);
    if (val) lval.download(val, acc, c.stream);
    c.sync();
  }
  if (diag) {
    c.d_diag.download(diag, n, c.stream);
    c.sync();
  }
}

void install_h(Ctx& c, const std::vector<int64_t>& rowptr, const std::vector<int32_t>& col,
               const std::vector<double>& val, const std::vector<int32_t>& lowcnt, const std::vector<double>& diag) {
  const int64_t n = c.sp.nsef_total, nnz = rowptr[n];
  c.have_h = false;
  c.d_rowptr.upload(rowptr, c.stream);
  c.d_col.upload(col, c.stream);
  c.d_val.upload(val, c.stream);
  c.d_lowcnt.upload(lowcnt, c.stream);
  c.d_diag.upload(diag, c.stream);
  c.d_val32.alloc(nnz);
  if (nnz > 0)
    hipLaunchKernelGGL(k_to_f32, dim3(blocks_for(nnz, 256)), dim3(256), 0, c.stream, nnz, c.d_val.p, c.d_val32.p);
  PSOCI_HIP(hipGetLastError());
  c.nnz_full = nnz;
  c.nnz_lower = (nnz + n) / 2;
  c.sync();
  c.have_h = true;
  c.have_diag = true;
  c.hbuild_seconds[0] = c.hbuild_seconds[1] = -1.0;  // not built here
}

}  // namespace psoci
