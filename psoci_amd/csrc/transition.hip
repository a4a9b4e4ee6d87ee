// Transition Hamiltonian between two configuration lists of one orbital
// basis.
//
// Replaces constructGlobalTransitionHamiltonian2WAY
// (PsociGAhamiltonian.C:2089-2287; driver.Transition.C) with packSefsef
// (:4077-4136).  Rows are the SEFs of the context's list (list 1); for each
// configuration I of list 1 and each J <= I of list 2 (positions in the two
// lists) the reference runs generateSOCIblocks(list1[I], list2[J]) -- its
// diagonal path whenever the indices agree, whatever the two configurations,
// and packs with list 1's SEF entry points for the row and the column
// (driver.Transition.C:579 computes them from list 1 only): columns
// sef_off1[J] + j, the lower triangle j <= i when I == J, and |h| >= 1e-8
// or the (I == J, i == j) element.  Kept alike on purpose, so the output
// matches the reference's bit for bit.  On an equal-index pair of two
// configurations a double excitation apart the reference's same-spin
// branch reads vnt[1] of the diagonal vnt, which exists only with two or
// more open shells: the exchange of the two lowest open shells is used
// then, and the build is refused when the read would fall past the end;
// it is refused too for an equal-index block whose list-2 configuration
// has fewer SEFs, where packSefsef's lower triangle (j <= i) reads past
// the block.
//
// Same pipeline as the stored build, on two lists: a wavefront-ballot
// screen, pair setup, one thread per block element, then a count / scan /
// fill of (row, pair) segments -- deterministic, no atomics but the refusal
// flag.
#include <cmath>

#include "ctx.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;
constexpr double kThresh = 1.0e-8;  // MIN_DOUBLE_VALUE (PsociHamiltonian.hpp:37)

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

struct ListView {
  const uint64_t* dm;
  const uint64_t* sm;
  const int8_t* nopen;
  const uint8_t* cph;
  const uint8_t* cls;
  const double* cmult;
  const int32_t* jsym;
  const int32_t* nsef;
  const int64_t* sefoff;
};

__device__ __forceinline__ ConfView view_of(const ListView& L, int64_t c, const uint16_t* pat,
                                            const int32_t* patoff) {
  ConfView v;
  v.d = L.dm[c];
  v.s = L.sm[c];
  v.nopen = L.nopen[c];
  v.ph = L.cph[c];
  v.pat = pat + patoff[12 * L.cls[c] + v.nopen];
  return v;
}

// per I of list 1: the J <= I of list 2 within two moved electrons and four
// open shells (generateSOCIblocks' early returns), ascending
template <bool FILL>
__global__ void k_tscreen(int64_t n1, int64_t n2, ListView a, ListView b, int32_t* __restrict__ cnt,
                          const int64_t* __restrict__ off, int32_t* __restrict__ out) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= n1) return;  // uniform per wavefront
  const uint64_t dI = a.dm[I], sI = a.sm[I];
  const int nI = a.nopen[I];
  const int64_t hi = (I + 1 < n2) ? I + 1 : n2;
  int64_t base = FILL ? off[I] : 0;
  int c = 0;
  for (int64_t J0 = 0; J0 < hi; J0 += kWave) {
    const int64_t J = J0 + lane;
    bool ok = false;
    if (J < hi) {
      const int dn = nI - b.nopen[J];
      if (dn <= 4 && dn >= -4) ok = moved(dI, sI, b.dm[J], b.sm[J]) <= 2;
    }
    const uint64_t m = __ballot(ok);
    if (FILL && ok) out[base + c + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)J;
    c += __popcll(m);
  }
  if (!FILL && lane == 0) cnt[I] = c;
}

__global__ void k_tcnt64(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

__global__ void k_trows(int64_t n1, const int64_t* __restrict__ off, int32_t* __restrict__ rowcfg) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= n1) return;
  for (int64_t t = off[I] + lane; t < off[I + 1]; t += kWave) rowcfg[t] = (int32_t)I;
}

__global__ void k_tsetup(int64_t np, const int32_t* __restrict__ rowcfg, const int32_t* __restrict__ colcfg,
                         ListView a, ListView b, int nbf, const double* __restrict__ one,
                         const double* __restrict__ two, PairInfo* __restrict__ pairs,
                         int64_t* __restrict__ bsz, int* __restrict__ bad) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int I = rowcfg[p], J = colcfg[p];
  PairInfo P;
  pair_setup(P, I, J, a.dm[I], a.sm[I], a.nopen[I], b.dm[J], b.sm[J], b.nopen[J], nbf, one, two, a.jsym[I],
             b.jsym[J]);
  if (I == J && P.kind == 1) {
    // the diagonal path's vnt[1]: exchange of the two lowest open shells of
    // list1[I] (processDiagExcitation's first pair); NaN marks a read past
    // a one-entry vnt, refused after the element pass
    uint64_t s = a.sm[I];
    if (popc(s) >= 2) {
      const int o0 = ctz(s);
      s &= s - 1;
      const int o1 = ctz(s);
      P.v1 = g2(two, o1, o0, o0, o1);
    } else {
      P.v1 = __builtin_nan("");
    }
  }
  if (I == J && P.kind != 0 && b.nsef[J] < a.nsef[I]) atomicOr(bad, 2);
  pairs[p] = P;
  bsz[p] = (P.kind != 0) ? (int64_t)a.nsef[I] * b.nsef[J] : 0;
}

__device__ __forceinline__ int64_t upper_bound_t(const int64_t* __restrict__ x, int64_t n, int64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (x[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// one thread per block element: staging[boff[p] + i nsef_J + j]
__global__ void k_telems(int64_t total, int64_t np, const int64_t* __restrict__ boff,
                         const PairInfo* __restrict__ pairs, ListView a, ListView b, const uint16_t* __restrict__ pat,
                         const int32_t* __restrict__ patoff, int frac, const double* __restrict__ two,
                         double* __restrict__ staging, int* __restrict__ bad) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int64_t p = upper_bound_t(boff, np, e) - 1;
  const PairInfo P = pairs[p];
  const int64_t loc = e - boff[p];
  const int nj = b.nsef[P.J];
  const int i = (int)(loc / nj), j = (int)(loc - (int64_t)i * nj);
  const ConfView A = view_of(a, P.I, pat, patoff);
  const ConfView B = view_of(b, P.J, pat, patoff);
  const double v = sef_elem(P, A, B, i, j, a.cmult[P.I], b.cmult[P.J], two, frac != 0);
  if (v != v) atomicOr(bad, 1);
  staging[e] = v;
}

// count (FILL = false) or write (FILL = true) the kept elements of one (row
// SEF, pair) segment; slot s = (I, i, t) with t the pair's place in I's list
template <bool FILL>
__global__ void k_tsegments(int64_t nslots, int64_t n1, const int64_t* __restrict__ slotbase,
                            const int64_t* __restrict__ off, const int32_t* __restrict__ colcfg,
                            const PairInfo* __restrict__ pairs, const int64_t* __restrict__ boff, ListView a,
                            ListView b, const double* __restrict__ staging, int64_t* __restrict__ cnt,
                            const int64_t* __restrict__ pos, int32_t* __restrict__ col, double* __restrict__ val,
                            int64_t* __restrict__ rowptr) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int64_t I = upper_bound_t(slotbase, n1, s) - 1;
  const int64_t np_I = off[I + 1] - off[I];
  const int64_t rem = s - slotbase[I];
  const int i = (int)(rem / np_I);
  const int t = (int)(rem - (int64_t)i * np_I);
  const int64_t p = off[I] + t;
  const int32_t J = colcfg[p];
  int64_t at = FILL ? pos[s] : 0;
  int64_t c = 0;
  if (pairs[p].kind != 0) {
    const int nj = b.nsef[J];
    const bool diag = J == I;
    const int jmax = diag ? i : nj - 1;
    const double* row = staging + boff[p] + (int64_t)i * nj;
    const int64_t cbase = a.sefoff[J];  // list 1's entry point of index J
    for (int q = 0; q <= jmax; ++q) {
      const double v = row[q];
      if (fabs(v) >= kThresh || (diag && q == i)) {
        if (FILL) {
          col[at + c] = (int32_t)(cbase + q);
          val[at + c] = v;
        }
        ++c;
      }
    }
  }
  if (!FILL) cnt[s] = c;
  if (FILL && t == 0) rowptr[a.sefoff[I] + i] = at;
}

// rows of a configuration with no pair at all (no slot): empty rows
__global__ void k_tempty(int64_t n1, const int64_t* __restrict__ off, const int64_t* __restrict__ slotbase,
                         const int64_t* __restrict__ pos, const int64_t* __restrict__ sefoff,
                         const int32_t* __restrict__ nsef, int64_t* __restrict__ rowptr) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= n1 || off[I + 1] > off[I]) return;
  for (int i = 0; i < nsef[I]; ++i) rowptr[sefoff[I] + i] = pos[slotbase[I]];
}

}  // namespace

void build_transition(Ctx& c, const Space& s2, int64_t* nnz_out) {
  const Space& s1 = c.sp;
  const int64_t n1 = s1.ncfg, n2 = s2.ncfg, nrow = s1.nsef_total;
  // list 2 on the device, its classes numbered as list 1's pattern families
  std::vector<uint8_t> cls2(n2);
  for (int64_t q = 0; q < n2; ++q) cls2[q] = class_flag(s2.nelec, s2.ksym, s2.jsym[q]) != s1.flag0 ? 1 : 0;
  DevBuf<uint64_t> dm2, sm2;
  DevBuf<int8_t> no2;
  DevBuf<uint8_t> cph2, c2;
  DevBuf<double> cm2;
  DevBuf<int32_t> js2, ns2;
  DevBuf<int64_t> so2;
  dm2.upload(s2.dm, c.stream);
  sm2.upload(s2.sm, c.stream);
  no2.upload(s2.nopen, c.stream);
  cph2.upload(s2.cphase, c.stream);
  c2.upload(cls2, c.stream);
  cm2.upload(s2.cmult, c.stream);
  js2.upload(s2.jsym, c.stream);
  ns2.upload(s2.nsef, c.stream);
  so2.upload(s2.sef_off, c.stream);
  const ListView a{c.d_dm.p, c.d_sm.p, c.d_nopen.p, c.d_cph.p, c.d_cls.p, c.d_cmult.p, c.d_jsym.p, c.d_nsef.p,
                   c.d_sefoff.p};
  const ListView b{dm2.p, sm2.p, no2.p, cph2.p, c2.p, cm2.p, js2.p, ns2.p, so2.p};
  // screen: count, scan, fill
  DevBuf<int32_t> cnt, colcfg, rowcfg;
  DevBuf<int64_t> cnt64, off;
  cnt.alloc(n1);
  cnt64.alloc(n1 + 1);
  off.alloc(n1 + 1);
  hipLaunchKernelGGL(k_tscreen<false>, dim3(nblk(n1, 4)), dim3(256), 0, c.stream, n1, n2, a, b, cnt.p, nullptr,
                     nullptr);
  PSOCI_HIP(hipGetLastError());
  PSOCI_HIP(hipMemsetAsync(cnt64.p + n1, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_tcnt64, dim3(nblk(n1, 256)), dim3(256), 0, c.stream, n1, cnt.p, cnt64.p);
  c.scan(cnt64.p, off.p, n1 + 1);
  int64_t np = 0;
  PSOCI_HIP(hipMemcpyAsync(&np, off.p + n1, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  colcfg.alloc(np);
  rowcfg.alloc(np);
  hipLaunchKernelGGL(k_tscreen<true>, dim3(nblk(n1, 4)), dim3(256), 0, c.stream, n1, n2, a, b, nullptr, off.p,
                     colcfg.p);
  hipLaunchKernelGGL(k_trows, dim3(nblk(n1, 4)), dim3(256), 0, c.stream, n1, off.p, rowcfg.p);
  PSOCI_HIP(hipGetLastError());
  // pair setup and block staging
  DevBuf<PairInfo> pairs;
  DevBuf<int64_t> bsz, boff;
  DevBuf<int> bad;
  bad.alloc(1);
  PSOCI_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), c.stream));
  pairs.alloc(np);
  bsz.alloc(np + 1);
  boff.alloc(np + 1);
  PSOCI_HIP(hipMemsetAsync(bsz.p + np, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_tsetup, dim3(nblk(np, 256)), dim3(256), 0, c.stream, np, rowcfg.p, colcfg.p, a, b, s1.nbf,
                     c.d_one.p, c.d_two.p, pairs.p, bsz.p, bad.p);
  PSOCI_HIP(hipGetLastError());
  c.scan(bsz.p, boff.p, np + 1);
  int64_t nstage = 0;
  PSOCI_HIP(hipMemcpyAsync(&nstage, boff.p + np, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  DevBuf<double> staging;
  staging.alloc(nstage);
  hipLaunchKernelGGL(k_telems, dim3(nblk(nstage, 256)), dim3(256), 0, c.stream, nstage, np, boff.p, pairs.p, a, b,
                     c.d_pat.p, c.d_patoff.p, s1.frac ? 1 : 0, c.d_two.p, staging.p, bad.p);
  PSOCI_HIP(hipGetLastError());
  int hbad = 0;
  bad.download(&hbad, 1, c.stream);
  // slots: nsef1[I] rows x pairs of I, in row order
  std::vector<int64_t> hoff(n1 + 1);
  off.download(hoff.data(), n1 + 1, c.stream);
  c.sync();
  PSOCI_CHECK(!(hbad & 2), PSOCI_ERR_UNSUPPORTED,
              "an equal-index block whose second-list configuration has fewer SEFs than the first: the reference's "
              "packSefsef reads past the block there");
  PSOCI_CHECK(!(hbad & 1), PSOCI_ERR_UNSUPPORTED,
              "an equal-index pair of configurations a double excitation apart, the first with fewer than two open "
              "shells: the reference reads past its one-entry diagonal vnt there");
  std::vector<int64_t> slotbase(n1 + 1);
  int64_t ns = 0;
  for (int64_t I = 0; I < n1; ++I) {
    slotbase[I] = ns;
    ns += (int64_t)s1.nsef[I] * (hoff[I + 1] - hoff[I]);
  }
  slotbase[n1] = ns;
  DevBuf<int64_t> dslot, scnt, spos;
  dslot.upload(slotbase, c.stream);
  scnt.alloc(ns + 1);
  spos.alloc(ns + 1);
  PSOCI_HIP(hipMemsetAsync(scnt.p + ns, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_tsegments<false>, dim3(nblk(ns, 256)), dim3(256), 0, c.stream, ns, n1, dslot.p, off.p,
                     colcfg.p, pairs.p, boff.p, a, b, staging.p, scnt.p, nullptr, nullptr, nullptr, nullptr);
  PSOCI_HIP(hipGetLastError());
  c.scan(scnt.p, spos.p, ns + 1);
  int64_t nnz = 0;
  PSOCI_HIP(hipMemcpyAsync(&nnz, spos.p + ns, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  c.d_trow.alloc(nrow + 1);
  c.d_tcol.alloc(nnz);
  c.d_tval.alloc(nnz);
  PSOCI_HIP(hipMemcpyAsync(c.d_trow.p + nrow, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_tsegments<true>, dim3(nblk(ns, 256)), dim3(256), 0, c.stream, ns, n1, dslot.p, off.p,
                     colcfg.p, pairs.p, boff.p, a, b, staging.p, nullptr, spos.p, c.d_tcol.p, c.d_tval.p,
                     c.d_trow.p);
  hipLaunchKernelGGL(k_tempty, dim3(nblk(n1, 256)), dim3(256), 0, c.stream, n1, off.p, dslot.p, spos.p, c.d_sefoff.p,
                     c.d_nsef.p, c.d_trow.p);
  PSOCI_HIP(hipGetLastError());
  c.sync();
  c.nnz_trans = nnz;
  c.have_trans = true;
  if (nnz_out) *nnz_out = nnz;
}

}  // namespace psoci
