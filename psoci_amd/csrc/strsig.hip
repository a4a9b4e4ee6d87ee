// String-driven direct sigma on the GPU (the plan and its index arithmetic
// are in ras.hpp / ras.cpp).
//
// Replaces, for complete RAS spaces (full CI and MR-CISD lists), the direct
// path of the reference (matrixVectorProductsScatterAccumulateLowLevel...
// DirectFetch, PsociGAhamiltonian.C:7549-7823, with matrixVectorUpdate
// :13036-13115), which regenerates every configuration-pair block.  One
// sigma call of K columns is, per chunk of KC <= 16 (FP32) or 8 (FP64)
// columns:
//   k_ras_x2c   SEF vectors -> determinant vectors c[det][KC] (each
//               determinant belongs to one SEF: a scatter without collisions);
//   k_ras_gemm  per block: W[row][p-pair][KC] = sum over q-pairs of
//               G[p-pair][q-pair] D[row][q-pair][KC], D gathered from c
//               through the string addition tables into LDS, the product on
//               MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64);
//   k_ras_y     per SEF and column: y = sum over its determinants of
//               coefficient * (sum over the RI terms of +-W), written once
//               (a gather: deterministic, no atomics).
// HBM layout: c [ndet][KC] (64 B per determinant at KC = 16 FP32), W
// [block rows][nqp][KC] so that a term of k_ras_y reads KC contiguous values.
#include <algorithm>
#include <cstdlib>

#include "ctx.hpp"
#include "ras.hpp"

namespace psoci {

namespace {

constexpr int kPT = 64;   // pair rows (p-pairs) per GEMM tile
constexpr int kMT = 128;  // (row, column) values per GEMM tile
constexpr int kQC = 16;   // q-pairs per LDS chunk

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of the 16x16 output held by accumulator t of lane l
  static __device__ __forceinline__ int row(int t, int l) { return 4 * (l >> 4) + t; }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int t, int l) { return 4 * t + (l >> 4); }
};

template <typename T>
__device__ __forceinline__ T sef_coef(uint8_t f) {
  const T r = (T)0x1.6a09e667f3bcdp-1;
  const T m = (f & 2) ? (T)1 : r;
  return (f & 1) ? -m : m;
}

// c[det][KC] = coefficient * x[sef][col0 + q] (zero for q >= kw)
template <typename T, int KC>
__global__ __launch_bounds__(256) void k_ras_x2c(const RasView* __restrict__ vp, const T* __restrict__ X, int ldx,
                                                 int col0, int kw, T* __restrict__ c) {
  const RasView& v = *vp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = t / KC;
  const int q = (int)(t % KC);
  if (s >= v.nsef) return;
  const RasSef e = v.sef[s];
  const T x = q < kw ? X[s * (int64_t)ldx + col0 + q] : (T)0;
  for (int d = 0; d < e.n; ++d) {
    const RasSectorDev& S = v.sec[e.sec[d]];
    const int ca = ras_cls_of(v, S.na, e.a[d]), cb = ras_cls_of(v, S.nb, e.b[d]);
    const int64_t det =
        ras_det(v, e.sec[d], ca, e.a[d] - v.lev[S.na].cls_off[ca], cb, e.b[d] - v.lev[S.nb].cls_off[cb]);
    if (det >= 0 && det < v.ndet) c[det * KC + q] = sef_coef<T>(e.flag[d]) * x;
  }
}

template <typename T, int N>
struct VecN;
template <>
struct VecN<float, 4> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct VecN<double, 2> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// One tile of W = D G^T for block bi: kPT p-pairs x kMT (row, column)
// values.  Four wavefronts, each 64 p-pairs x 32 values (4 x 2 MFMA tiles).
// Per chunk of kQC q-pairs: G[p][q] rows into LDS, the D entries gathered
// from c through the addition tables into LDS, then 4 MFMA steps.  Tiles
// are dealt so that XCD x runs the contiguous tiles [x per, (x + 1) per):
// the p-tiles of one row tile, which gather the same c rows, share an L2.
template <typename T, int KC, int KIND>
__global__ __launch_bounds__(256) void k_ras_gemm(const RasView* __restrict__ vp, int bi, const T* __restrict__ G,
                                                  const T* __restrict__ c, T* __restrict__ W, int64_t ntiles,
                                                  int npt) {
  typedef Mfma<T> M;
  typedef typename M::v4 v4;
  constexpr int RT = kMT / KC;                  // block rows per tile
  constexpr int VPT = KC < 8 ? KC : 8;          // consecutive columns one thread gathers
  constexpr int CPT = 8 / VPT;                  // (row, pair) entries per thread and chunk
  constexpr int TPC = KC > 8 ? KC / 8 : 1;      // threads per entry
  __shared__ T Gs[kPT][kQC + 1];
  __shared__ T Ds[kQC][kMT + 4];
  __shared__ int64_t s_ra[RT], s_rb[RT];
  __shared__ int64_t s_boff[kRasMaxCls][kRasMaxCls];
  __shared__ int64_t s_cnt[kRasMaxCls], s_off1[kRasMaxCls];
  const RasView& v = *vp;
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;  // uniform per workgroup
  const int64_t mt = tile / npt;
  const int pt = (int)(tile % npt);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const RasBlockDev& B = v.blk[bi];
  const int64_t nrows = B.nrows, nw = B.nw, u0 = B.u0, w0 = B.w0;
  const int nqp = B.nqp, nbf = v.nbf;
  const RasSectorDev& S = v.sec[B.sector];
  const int64_t det0 = S.det0, ndet = v.ndet;
  // second-addition level of a same-spin block: l2 + 1
  const int l2 = KIND == 1 ? B.lu : B.lw;
  const int64_t add1_off = KIND == 0 ? 0 : v.lev[l2 + 1].add_off;
  for (int i = tid; i < kRasMaxCls * kRasMaxCls; i += 256) s_boff[i / kRasMaxCls][i % kRasMaxCls] = S.boff[i / kRasMaxCls][i % kRasMaxCls];
  if (tid < kRasMaxCls) {
    s_cnt[tid] = v.lev[S.nb].cls_off[tid + 1] - v.lev[S.nb].cls_off[tid];
    s_off1[tid] = KIND == 0 ? 0 : v.lev[l2 + 1].cls_off[tid];
  }
  for (int rl = tid; rl < RT; rl += 256) {
    const int64_t row = mt * RT + rl;
    int64_t ra = -1, rb = -1;
    if (row < nrows) {
      const int64_t u = u0 + row / nw, w = w0 + row % nw;
      if (KIND == 0) {
        ra = v.lev[B.lu].add_off + u * nbf;
        rb = v.lev[B.lw].add_off + w * nbf;
      } else if (KIND == 1) {
        ra = v.lev[B.lu].add_off + u * nbf;
        const int cb = ras_cls_of(v, S.nb, w);
        rb = ((int64_t)cb << 40) | (w - v.lev[S.nb].cls_off[cb]);
      } else {
        const int ca = ras_cls_of(v, S.na, u);
        ra = ((int64_t)ca << 40) | (u - v.lev[S.na].cls_off[ca]);
        rb = v.lev[B.lw].add_off + w * nbf;
      }
    }
    s_ra[rl] = ra;
    s_rb[rl] = rb;
  }
  const T* __restrict__ Gb = G + B.goff;
  const int16_t* __restrict__ ql = v.qlist + B.qoff;
  const uint32_t* __restrict__ add = v.add;
  v4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};
  __syncthreads();
  for (int q0 = 0; q0 < nqp; q0 += kQC) {
    {
      const int p = tid >> 2, jq = (tid & 3) * 4;
      const int pg = pt * kPT + p;
#pragma unroll
      for (int i = 0; i < 4; ++i) Gs[p][jq + i] = pg < nqp ? Gb[(int64_t)pg * nqp + q0 + jq + i] : (T)0;
    }
#pragma unroll
    for (int ci = 0; ci < CPT; ++ci) {
      const int ent = (tid / TPC) * CPT + ci;  // (row, pair) entry: rl * kQC + jq
      const int v0 = (tid % TPC) * VPT;
      const int rl = ent / kQC, jq = ent % kQC;
      T xv[VPT];
#pragma unroll
      for (int i = 0; i < VPT; ++i) xv[i] = (T)0;
      const int64_t ra = s_ra[rl], rb = s_rb[rl];
      const int16_t qs = ql[q0 + jq];
      if (ra >= 0 && qs >= 0) {
        const int q = qs & 0xff, s = (qs >> 8) & 0xff;
        int64_t det = -1;
        int neg = 0;
        if (KIND == 0) {
          const uint32_t ea = add[ra + q], eb = add[rb + s];
          if (ea != kRasNone && eb != kRasNone) {
            neg = ras_neg(ea) ^ ras_neg(eb);
            const int ca = ras_cls(ea), cb = ras_cls(eb);
            const int64_t bo = s_boff[ca][cb];
            det = bo < 0 ? -1 : det0 + bo + ras_rank(ea) * s_cnt[cb] + ras_rank(eb);
          }
        } else {
          const int64_t base = KIND == 1 ? ra : rb;
          const uint32_t e1 = add[base + s];
          if (e1 != kRasNone) {
            const int64_t g1 = s_off1[ras_cls(e1)] + ras_rank(e1);
            const uint32_t e2 = add[add1_off + g1 * nbf + q];
            if (e2 != kRasNone) {
              neg = ras_neg(e1) ^ ras_neg(e2);
              if (KIND == 1) {
                const int cb = (int)(rb >> 40);
                const int64_t bo = s_boff[ras_cls(e2)][cb];
                det = bo < 0 ? -1 : det0 + bo + ras_rank(e2) * s_cnt[cb] + (rb & 0xffffffffffll);
              } else {
                const int ca = (int)(ra >> 40), cb = ras_cls(e2);
                const int64_t bo = s_boff[ca][cb];
                det = bo < 0 ? -1 : det0 + bo + (ra & 0xffffffffffll) * s_cnt[cb] + ras_rank(e2);
              }
            }
          }
        }
        if (det >= 0 && det < ndet) {
          const T* __restrict__ src = c + det * KC + v0;
          if constexpr (VPT * sizeof(T) % 16 == 0) {
            typedef typename VecN<T, 16 / sizeof(T)>::type V;
            constexpr int W16 = 16 / sizeof(T);
#pragma unroll
            for (int i = 0; i < VPT / W16; ++i) {
              const V x = reinterpret_cast<const V*>(src)[i];
#pragma unroll
              for (int w = 0; w < W16; ++w) xv[i * W16 + w] = neg ? -x[w] : x[w];
            }
          } else {
#pragma unroll
            for (int i = 0; i < VPT; ++i) xv[i] = neg ? -src[i] : src[i];
          }
        }
      }
#pragma unroll
      for (int i = 0; i < VPT; ++i) Ds[jq][rl * KC + v0 + i] = xv[i];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kQC / 4; ++kk) {
      T a[4], b[2];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = Gs[16 * i + (lane & 15)][4 * kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Ds[4 * kk + (lane >> 4)][32 * wave + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::op(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  const int64_t wb = B.wbase;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int m = 32 * wave + 16 * j + (lane & 15);
      const int64_t row = mt * RT + m / KC;
      const int vq = m % KC;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int p = pt * kPT + 16 * i + M::row(t, lane);
        if (row < nrows && p < nqp) W[(wb + row * nqp + p) * KC + vq] = acc[i][j][t];
      }
    }
}

// y[sef][col0 + q] = sum over the SEF's determinants of coefficient (x 2 for
// an Ms > 0 determinant, whose mirror is not stored) times sum of +-W terms
template <typename T, int KC>
__global__ __launch_bounds__(256) void k_ras_y(const RasView* __restrict__ vp, const T* __restrict__ W,
                                               T* __restrict__ Y, int ldy, int col0, int kw) {
  const RasView& v = *vp;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = t / KC;
  const int q = (int)(t % KC);
  if (s >= v.nsef) return;
  const RasSef e = v.sef[s];
  T y = (T)0;
  for (int d = 0; d < e.n; ++d) {
    T sig = (T)0;
    ras_terms(v, e.sec[d], e.a[d], e.b[d], [&](int bi, int64_t off, int neg) {
      if (bi < 0 || off < 0 || off >= v.wunit) return;
      const T w = W[off * KC + q];
      sig += neg ? -w : w;
    });
    const T cf = sef_coef<T>(e.flag[d]) * (e.sec[d] > 0 ? (T)2 : (T)1);
    y += cf * sig;
  }
  if (q < kw) Y[s * (int64_t)ldy + col0 + q] = y;
}

// Flags (*flag = 1) a configuration pair one electron apart whose
// single-excitation sum vnt[0] is exactly zero in either orientation while
// an exchange integral (e o|o f) its elements could use is not: the
// reference skips such a block (processSingleExcitation returns 0,
// PsociHamiltonian.C:190-261), which the determinant Hamiltonian does not.
__global__ __launch_bounds__(256) void k_ras_zero_singles(int64_t ncfg, const uint64_t* __restrict__ dm,
                                                          const uint64_t* __restrict__ sm, uint64_t emask, int m,
                                                          int nbf, const double* __restrict__ one,
                                                          const double* __restrict__ two, int* __restrict__ flag) {
  const int64_t ci = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= ncfg) return;
  const uint64_t dI = dm[ci], sI = sm[ci];
  const int nI = popc(sI);
  for (uint64_t ex = dI | sI; ex; ex &= ex - 1) {
    const int e = ctz(ex);
    const uint64_t be = 1ull << e;
    for (int f = 0; f < nbf; ++f) {
      const uint64_t bf = 1ull << f;
      if (f == e || (dI & bf)) continue;
      uint64_t dJ = dI, sJ = sI;
      if (dJ & be) {
        dJ &= ~be;
        sJ |= be;
      } else {
        sJ &= ~be;
      }
      if (sJ & bf) {
        sJ &= ~bf;
        dJ |= bf;
      } else {
        sJ |= bf;
      }
      if (2 * popc(dJ & emask) + popc(sJ & emask) > m) continue;
      const int nJ = popc(sJ);
      PairInfo P1, P2;
      pair_setup(P1, 1, 0, dI, sI, nI, dJ, sJ, nJ, nbf, one, two);
      pair_setup(P2, 1, 0, dJ, sJ, nJ, dI, sI, nI, nbf, one, two);
      if (P1.kind != 0 && P2.kind != 0) continue;
      for (uint64_t o = (dI | sI | dJ | sJ) & ~be & ~bf; o; o &= o - 1) {
        const int c = ctz(o);
        if (g2(two, e, c, c, f) != 0.0) {
          atomicOr(flag, 1);
          return;
        }
      }
    }
  }
}

template <typename T, int KC>
void ras_chunk(Ctx& c, const T* X, int ldx, T* Y, int ldy, int col0, int kw) {
  const RasPlan& P = c.ras;
  const RasView* vp = c.d_rasview.p;
  T* cv = reinterpret_cast<T*>(c.d_rasc.p);
  T* W = reinterpret_cast<T*>(c.d_rasw.p);
  const T* G = sizeof(T) == 4 ? reinterpret_cast<const T*>(c.d_rasg32.p) : reinterpret_cast<const T*>(c.d_rasg64.p);
  const int64_t nsef = (int64_t)P.sef.size();
  hipLaunchKernelGGL((k_ras_x2c<T, KC>), dim3(nblk(nsef * KC, 256)), dim3(256), 0, c.stream, vp, X, ldx, col0, kw,
                     cv);
  PSOCI_HIP(hipGetLastError());
  for (int bi = 0; bi < (int)P.blk.size(); ++bi) {
    const RasBlockDev& B = P.blk[bi];
    const int npt = (B.nqp + kPT - 1) / kPT;
    const int64_t nmt = (B.nrows * KC + kMT - 1) / kMT;
    const int64_t ntiles = nmt * npt;
    const unsigned grid = (unsigned)(8 * ((ntiles + 7) / 8));
    if (B.kind == 0)
      hipLaunchKernelGGL((k_ras_gemm<T, KC, 0>), dim3(grid), dim3(256), 0, c.stream, vp, bi, G, cv, W, ntiles, npt);
    else if (B.kind == 1)
      hipLaunchKernelGGL((k_ras_gemm<T, KC, 1>), dim3(grid), dim3(256), 0, c.stream, vp, bi, G, cv, W, ntiles, npt);
    else
      hipLaunchKernelGGL((k_ras_gemm<T, KC, 2>), dim3(grid), dim3(256), 0, c.stream, vp, bi, G, cv, W, ntiles, npt);
    PSOCI_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((k_ras_y<T, KC>), dim3(nblk(nsef * KC, 256)), dim3(256), 0, c.stream, vp, W, Y, ldy, col0, kw);
  PSOCI_HIP(hipGetLastError());
}

template <typename T>
void ras_any_k(Ctx& c, int k, const T* X, int ldx, T* Y, int ldy) {
  const int kmax = sizeof(T) == 4 ? 16 : 8;
  for (int c0 = 0; c0 < k; c0 += kmax) {
    const int kw = std::min(kmax, k - c0);
    if (kw > 8) ras_chunk<T, 16>(c, X, ldx, Y, ldy, c0, kw);
    else if (kw > 4) ras_chunk<T, 8>(c, X, ldx, Y, ldy, c0, kw);
    else if (kw > 2) ras_chunk<T, 4>(c, X, ldx, Y, ldy, c0, kw);
    else if (kw > 1) ras_chunk<T, 2>(c, X, ldx, Y, ldy, c0, kw);
    else ras_chunk<T, 1>(c, X, ldx, Y, ldy, c0, kw);
  }
}

}  // namespace

// Builds the string plan and its device tables; false (c.ras_why says why)
// when the space does not qualify.
bool build_ras(Ctx& c) {
  const Space& sp = c.sp;
  c.have_ras = false;
  std::vector<double> one(c.d_one.n), two(c.d_two.n);
  c.d_one.download(one.data(), one.size(), c.stream);
  c.d_two.download(two.data(), two.size(), c.stream);
  c.sync();
  ras_plan(c.ras, sp, one.data(), two.data());
  if (!c.ras.ok) {
    c.ras_why = c.ras.why;
    return false;
  }
  const RasPlan& P = c.ras;
  {
    DevBuf<int> flag;
    flag.alloc(1);
    PSOCI_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), c.stream));
    hipLaunchKernelGGL(k_ras_zero_singles, dim3(nblk(sp.ncfg, 256)), dim3(256), 0, c.stream, sp.ncfg, c.d_dm.p,
                       c.d_sm.p, P.emask, P.m, sp.nbf, c.d_one.p, c.d_two.p, flag.p);
    PSOCI_HIP(hipGetLastError());
    int f = 0;
    flag.download(&f, 1, c.stream);
    c.sync();
    if (f) {
      c.ras_why = "a configuration pair one electron apart has vnt[0] exactly zero with nonzero exchange terms";
      c.ras = RasPlan();
      return false;
    }
  }
  c.d_rasmask.upload(P.mask, c.stream);
  c.d_rascls.upload(P.cls, c.stream);
  c.d_rasadd.upload(P.add, c.stream);
  c.d_rasrem.upload(P.rem, c.stream);
  c.d_rasblk.upload(P.blk, c.stream);
  c.d_rasq.upload(P.qlist, c.stream);
  c.d_raspidx.upload(P.pidx, c.stream);
  c.d_rassef.upload(P.sef, c.stream);
  c.d_rasg64.upload(P.G, c.stream);
  std::vector<float> g32(P.G.begin(), P.G.end());
  c.d_rasg32.upload(g32, c.stream);
  RasView v = ras_host_view(P);
  v.mask = c.d_rasmask.p;
  v.cls = c.d_rascls.p;
  v.add = c.d_rasadd.p;
  v.rem = c.d_rasrem.p;
  v.blk = c.d_rasblk.p;
  v.qlist = c.d_rasq.p;
  v.pidx = c.d_raspidx.p;
  v.sef = c.d_rassef.p;
  c.d_rasview.upload(&v, 1, c.stream);
  // determinant vectors and W: 16 FP32 or 8 FP64 columns per chunk
  c.d_rasc.alloc((size_t)P.ndet * 16 * sizeof(float));
  c.d_rasw.alloc((size_t)P.wunit * 16 * sizeof(float));
  c.sync();
  c.have_ras = true;
  return true;
}

void build_direct(Ctx& c) {
  if (c.direct_engine_req != 1 && !c.have_ras && !c.ras_tried) {
    c.ras_tried = true;
    if (build_ras(c)) return;
  }
  PSOCI_CHECK(c.direct_engine_req != 2 || c.have_ras, PSOCI_ERR_UNSUPPORTED,
              "the string-driven direct sigma does not apply to this space: " + c.ras_why);
  if (!direct_engine(c)) build_pairs(c);
}

void sigma_ras(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy) {
  PSOCI_CHECK(c.have_ras, PSOCI_ERR_STATE, "the string-driven direct sigma is not set up");
  if (prec == PSOCI_FP32)
    ras_any_k<float>(c, k, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy);
  else
    ras_any_k<double>(c, k, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy);
}

}  // namespace psoci
