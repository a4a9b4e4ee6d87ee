// Adjoint of the Davidson energies: dE/dtheta = sum_k w_k c_k^T (dH/dtheta) c_k
// for every packed one- and two-electron integral theta -- the generalised
// 1- and 2-RDM in the reference's packed conventions.
//
// No reference kernel computes the 2-RDM; the 1-RDM part corresponds to the
// CI density of PsociNOpopulations::constructDensityMatrix
// (PsociNOpopulations.C:858-989, via transformSEFSEFtoMOGET,
// PsociGAhamiltonian.C:10027-10119): dE/dh_pp = D_pp and dE/dh_pq = 2 D_pq.
//
// Every SEF element is a linear form in the integrals (generateSEFCIG and
// transformCDCt, PsociHamiltonian.C:860-1118, 1350-1454), so its adjoint is
// the same case analysis run backwards: one wavefront per row configuration
// I walks the pairs (I, J <= I) in the reference orientation, lanes take the
// elements of the block, weight W_ij = f * sum_k w_k c_k[i] c_k[j] (f = 2
// off the diagonal block, the element standing for (i, j) and (j, i)),
// and accumulate per pair the weight on vnt[0], vnt[1] and on the exchange
// integrals (e0 o|o f0) of the row configuration's open shells.  The pair's
// sums are then spread over the integrals they are made of, with one
// atomic add per integral per pair.
#include "ctx.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

constexpr int kMaxOpen = 10;  // open shells per configuration (exchange slots)

struct GradAcc {
  double wv0, wv1;
  double xg[kMaxOpen];
};

__device__ __forceinline__ void add_grad(double* p, double v) {
  if (v != 0.0) atomicAdd(p, v);
}

// Adjoint of detval_ao / detval_x: adds the weight dW of one determinant pair
// to the integrals its value is made of.
__device__ __forceinline__ void det_grad(const PairInfo& P, const ConfView& A, const ConfView& B, uint64_t ao,
                                         uint64_t bo, double dW, GradAcc& acc, double* __restrict__ g2buf) {
  if (dW == 0.0) return;
  uint64_t aI = A.d | ao, bI = A.d | (A.s & ~ao);
  uint64_t aJ = B.d | bo, bJ = B.d | (B.s & ~bo);
  if (popc(aI) != popc(aJ)) return;
  uint64_t xa = aJ & ~aI, xb = bJ & ~bI;
  int ma = popc(xa), mb = popc(xb);
  int tot = ma + mb;
  if (tot > 2) return;
  if (tot == 0) {
    // vnt[0] minus the same-spin open-shell exchanges
    acc.wv0 += dW;
    uint64_t si = A.s, seen = 0;
    while (si) {
      int io = ctz(si);
      si &= si - 1;
      uint64_t sj = seen;
      int ia = (int)((aI >> io) & 1);
      while (sj) {
        int jo = ctz(sj);
        sj &= sj - 1;
        if (ia == (int)((aI >> jo) & 1)) add_grad(g2buf + g2_slot(io, jo, jo, io), -dW);
      }
      seen |= 1ull << io;
    }
    return;
  }
  int cnt = A.ph + B.ph;
  if (tot == 1) {
    uint64_t S = ma ? aI : bI;
    cnt += popc(S & between(P.e0, P.f0));
    const double s = (cnt & 1) ? -dW : dW;
    acc.wv0 += s;
    const uint64_t T = S & ~A.d & ~(1ull << P.e0) & ~(1ull << P.f0);
    uint64_t rem = A.s;
#pragma unroll
    for (int t = 0; t < kMaxOpen; ++t) {
      if (rem) {
        int o = ctz(rem);
        rem &= rem - 1;
        if ((T >> o) & 1ull) acc.xg[t] -= s;
      }
    }
  } else if (ma == 1) {
    int iex2 = ctz(xa), jex2 = ctz(xb);
    int iex3 = ctz(aI & ~aJ), jex3 = ctz(bI & ~bJ);
    cnt += popc(aI & between(iex2, iex3)) + popc(bI & between(jex2, jex3));
    const double s = (cnt & 1) ? -dW : dW;
    if (P.kind == 3) {
      if ((iex3 == P.e0) == (iex2 == P.f0)) acc.wv0 += s;
      else acc.wv1 += s;
    } else {
      int c = (iex2 == jex3) ? iex2 : ((iex3 == jex2) ? iex3 : -1);
      if (P.kind == 2 && c >= 0 && ((A.s >> c) & 1ull) && c != P.e0 && c != P.f0) {
        const int slot = popc(A.s & ((1ull << c) - 1ull));
#pragma unroll
        for (int t = 0; t < kMaxOpen; ++t)
          if (t == slot) acc.xg[t] += s;
      } else {
        add_grad(g2buf + g2_slot(jex2, jex3, iex2, iex3), s);
      }
    }
  } else {
    uint64_t SI = ma ? aI : bI, SJ = ma ? aJ : bJ;
    cnt += popc(SJ & between(P.f0, P.f1)) + popc(SI & between(P.e0, P.e1));
    const double s = (cnt & 1) ? -dW : dW;
    acc.wv0 += s;
    acc.wv1 -= s;
  }
}

// Adjoint of sef_elem: the element's four determinant pairs with the
// transformCDCt coefficients.
// With one determinant per SEF (frac) the element is D(aI0, aJ0) itself.
__device__ __forceinline__ void sef_grad(const PairInfo& P, const ConfView& A, const ConfView& B, uint64_t aI0,
                                         uint64_t aI1, uint64_t aJ0, uint64_t aJ1, double mult, bool frac, double W,
                                         GradAcc& acc, double* __restrict__ g2buf) {
  const double r2 = 0.5, r1 = kRthlf;
  if (frac || (A.nopen == 0 && B.nopen == 0)) {
    det_grad(P, A, B, aI0, aJ0, W, acc, g2buf);
  } else if (B.nopen == 0) {
    det_grad(P, A, B, aI0, aJ0, W * r1, acc, g2buf);
    det_grad(P, A, B, aI1, aJ0, W * mult * r1, acc, g2buf);
  } else if (A.nopen == 0) {
    det_grad(P, A, B, aI0, aJ0, W * r1, acc, g2buf);
    det_grad(P, A, B, aI0, aJ1, W * mult * r1, acc, g2buf);
  } else {
    det_grad(P, A, B, aI0, aJ0, W * r2, acc, g2buf);
    det_grad(P, A, B, aI0, aJ1, W * mult * r2, acc, g2buf);
    det_grad(P, A, B, aI1, aJ0, W * mult * r2, acc, g2buf);
    det_grad(P, A, B, aI1, aJ1, W * r2, acc, g2buf);
  }
}

// spread a pair's accumulated weights over its integrals (lane 0 of the wave)
__device__ void spread_pair(const PairInfo& P, const ConfView& A, const ConfView& B, const GradAcc& acc,
                            double* __restrict__ g1, double* __restrict__ g2buf) {
  if (P.kind == 1) {
    // vnt[0] of processDiagExcitation
    const double w = acc.wv0;
    if (w == 0.0) return;
    uint64_t occ = A.d | A.s;
    while (occ) {
      const int ib = ctz(occ);
      occ &= occ - 1;
      const int oi = (int)((A.d >> ib) & 1) * 2 + (int)((A.s >> ib) & 1);
      if (oi == 1) {
        add_grad(g1 + pair_idx(ib, ib), w);
      } else {
        add_grad(g1 + pair_idx(ib, ib), 2.0 * w);
        add_grad(g2buf + g2_slot(ib, ib, ib, ib), w);
      }
      uint64_t lower = (A.d | A.s) & ((1ull << ib) - 1ull);
      while (lower) {
        const int jb = ctz(lower);
        lower &= lower - 1;
        const int oj = (int)((A.d >> jb) & 1) * 2 + (int)((A.s >> jb) & 1);
        const int ninj = oi * oj;
        if (ninj == 1) {
          add_grad(g2buf + g2_slot(ib, ib, jb, jb), w);
        } else {
          const double f = (ninj == 2) ? 1.0 : 2.0;
          add_grad(g2buf + g2_slot(ib, ib, jb, jb), 2.0 * f * w);
          add_grad(g2buf + g2_slot(ib, jb, ib, jb), -f * w);
        }
      }
    }
  } else if (P.kind == 2) {
    const int e = P.e0, f = P.f0;
    const double w = acc.wv0;
    if (w != 0.0) {
      add_grad(g1 + pair_idx(e, f), w);
      add_grad(g2buf + g2_slot(e, f, e, e), -w);
      uint64_t occ = A.d | A.s;
      while (occ) {
        const int i = ctz(occ);
        occ &= occ - 1;
        const bool dbl = (A.d >> i) & 1ull;
        add_grad(g2buf + g2_slot(e, f, i, i), dbl ? 2.0 * w : w);
        if (dbl && ((B.d >> i) & 1ull)) add_grad(g2buf + g2_slot(e, i, i, f), -w);
      }
    }
    uint64_t rem = A.s;
#pragma unroll
    for (int t = 0; t < kMaxOpen; ++t) {
      if (rem) {
        const int o = ctz(rem);
        rem &= rem - 1;
        add_grad(g2buf + g2_slot(e, o, o, f), acc.xg[t]);
      }
    }
  } else if (P.kind == 3) {
    add_grad(g2buf + g2_slot(P.e0, P.f0, P.e1, P.f1), acc.wv0);
    add_grad(g2buf + g2_slot(P.e0, P.f1, P.e1, P.f0), acc.wv1);
  }
}

struct RdmArgs {
  int64_t ncfg;
  const int64_t* recoff;
  const PairRec* rec;
  const ConfRec* conf;
  const uint64_t* detao;
  double mult;
  int frac;          // one determinant per SEF (odd electron count)
  int k;
  const double* C;   // [n][k]
  const double* w;   // [k]
  double* g1;
  double* g2;
};

__device__ __forceinline__ ConfView view_rec(const ConfRec& r) {
  ConfView v;
  v.d = r.d;
  v.s = r.s;
  v.nopen = r.nopen;
  v.ph = r.cph;
  v.pat = nullptr;
  return v;
}

__global__ __launch_bounds__(256) void k_rdm(RdmArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (I >= a.ncfg) return;
  const ConfRec ri = a.conf[I];
  const ConfView A = view_rec(ri);
  const int nI = ri.nsef;
  for (int64_t t = a.recoff[I]; t < a.recoff[I + 1]; ++t) {
    PairInfo P;
    int J;
    bool rowI;
    unpack_rec(a.rec[t], P, J, rowI);
    if (!rowI || P.kind == 0) continue;  // each unordered pair once, as the reference builds it
    const ConfRec rj = a.conf[J];
    const ConfView B = view_rec(rj);
    const int nJ = rj.nsef;
    const double fac = (J == I) ? 1.0 : 2.0;
    GradAcc acc;
    acc.wv0 = acc.wv1 = 0.0;
#pragma unroll
    for (int u = 0; u < kMaxOpen; ++u) acc.xg[u] = 0.0;
    for (int e = lane; e < nI * nJ; e += kWave) {
      const int i = e / nJ, j = e - i * nJ;
      const double* ci = a.C + ((int64_t)ri.sefoff + i) * a.k;
      const double* cj = a.C + ((int64_t)rj.sefoff + j) * a.k;
      double W = 0.0;
      for (int v = 0; v < a.k; ++v) W += (a.w ? a.w[v] : 1.0) * ci[v] * cj[v];
      W *= fac;
      const bool fr = a.frac != 0;
      const uint64_t aI0 = a.detao[ri.detoff + (fr ? i : (A.nopen ? 2 * i : 0))];
      const uint64_t aI1 = (!fr && A.nopen) ? a.detao[ri.detoff + 2 * i + 1] : 0ull;
      const uint64_t aJ0 = a.detao[rj.detoff + (fr ? j : (B.nopen ? 2 * j : 0))];
      const uint64_t aJ1 = (!fr && B.nopen) ? a.detao[rj.detoff + 2 * j + 1] : 0ull;
      sef_grad(P, A, B, aI0, aI1, aJ0, aJ1, a.mult, fr, W, acc, a.g2);
    }
    // reduce the pair's weights over the wavefront, then spread them
    for (int off = kWave / 2; off > 0; off >>= 1) {
      acc.wv0 += __shfl_xor(acc.wv0, off, kWave);
      acc.wv1 += __shfl_xor(acc.wv1, off, kWave);
#pragma unroll
      for (int u = 0; u < kMaxOpen; ++u) acc.xg[u] += __shfl_xor(acc.xg[u], off, kWave);
    }
    if (lane == 0) spread_pair(P, A, B, acc, a.g1, a.g2);
  }
}

__global__ void k_zero(int64_t n, double* __restrict__ p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

}  // namespace

void rdm_grad(Ctx& c, int k, const double* dC, const double* w, double* d_g1, double* d_g2) {
  if (!c.have_pairs) build_pairs(c);
  const int64_t n1 = (int64_t)c.sp.nbf * (c.sp.nbf + 1) / 2, n2 = n1 * (n1 + 1) / 2;
  hipLaunchKernelGGL(k_zero, dim3(nblk(n1, 256)), dim3(256), 0, c.stream, n1, d_g1);
  hipLaunchKernelGGL(k_zero, dim3(nblk(n2, 256)), dim3(256), 0, c.stream, n2, d_g2);
  DevBuf<double> dw;
  if (w) dw.upload(w, k, c.stream);
  RdmArgs a;
  a.ncfg = c.sp.ncfg;
  a.recoff = c.d_adjoff.p;
  a.rec = c.d_rec.p;
  a.conf = c.d_conf.p;
  a.detao = c.d_detao.p;
  a.mult = c.sp.mult;
  a.frac = c.sp.frac ? 1 : 0;
  a.k = k;
  a.C = dC;
  a.w = w ? dw.p : nullptr;
  a.g1 = d_g1;
  a.g2 = d_g2;
  hipLaunchKernelGGL(k_rdm, dim3(nblk(a.ncfg, 4)), dim3(256), 0, c.stream, a);
  PSOCI_HIP(hipGetLastError());
  c.sync();
}

}  // namespace psoci
