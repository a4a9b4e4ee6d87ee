// Direct sigma: Y = H X with the Hamiltonian regenerated on the fly.
//
// Replaces the direct path of the reference
// (matrixVectorProductsScatterAccumulateLowLevelAccumulateDirectFetch,
// PsociGAhamiltonian.C:7549-7823, with matrixVectorUpdate :13036-13115),
// which per configuration I screens every J <= I (compareAllOrbMap...),
// rebuilds the block with generateSOCIblocks and scatter-accumulates both
// y_I and y_J into a global array.
//
// Layout on the GPU:
//  * once per space and integrals, every connected ordered pair (I, J) gets
//    a 32-byte record: J, the excitation orbitals and block kind of the
//    reference orientation (row = the later configuration), and its
//    configuration-level sums vnt[0], vnt[1] (pair_setup, computed once and
//    reused by every sigma of a solve);
//  * a three-index table XT[e][f][o] = (e o|o f) (nbf^3 doubles, L2 resident)
//    supplies the single-excitation exchange integrals of any pair;
//  * one wavefront owns configuration I and walks its records, so y_I is
//    reduced in registers and written once: no atomics, a deterministic
//    result, X rows only gathered.  Lanes are (record slot, SEF i of I):
//    64 / nsef_I records are in flight per step, and each lane takes
//    determinant 2i of I and visits only the determinants of J that it
//    connects to (detdrive.hpp), so no lane tests SEF pairs whose elements
//    vanish (about 94% of the structural block entries of an MR-CISD space).
// Elements are the reference's determinant-pair values (detval_ao/detval_x,
// reference orientation) combined into SEF elements as D(2i,2j) +
// mult D(2i,2j+1), equal to transformCDCt's to rounding.
#include <algorithm>
#include <cstdlib>

#include "ctx.hpp"
#include "detdrive.hpp"
#include "pad.hpp"

namespace psoci {

namespace {

constexpr int kWave = 64;
constexpr int kTile = 1024;    // configurations per LDS tile in the screen
constexpr int kIPerWave = 16;  // row configurations per wavefront in the screen

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// alpha open-shell mask of every determinant: deposit(pattern, singly mask)
__global__ void k_detao(int64_t ncfg, const int64_t* __restrict__ detoff, const uint64_t* __restrict__ sm,
                        const int8_t* __restrict__ nopen, const uint16_t* __restrict__ pat,
                        const int32_t* __restrict__ patoff, uint64_t* __restrict__ detao) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t c = (int64_t)blockIdx.x * (blockDim.x / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  if (c >= ncfg) return;
  const int s = nopen[c];
  const int nd = 1 << (s > 0 ? s - 1 : 0);
  const uint16_t* p = pat + patoff[s];
  const uint64_t mask = sm[c];
  for (int d = lane; d < nd; d += kWave) detao[detoff[c] + d] = deposit(p[d], mask);
}

__global__ void k_confrec(int64_t ncfg, const uint64_t* __restrict__ dm, const uint64_t* __restrict__ sm,
                          const int8_t* __restrict__ nopen, const uint8_t* __restrict__ cph,
                          const int32_t* __restrict__ nsef, const int64_t* __restrict__ sefoff,
                          const int64_t* __restrict__ detoff, ConfRec* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncfg) return;
  ConfRec r;
  r.d = dm[c];
  r.s = sm[c];
  r.sefoff = (int32_t)sefoff[c];
  r.detoff = (int32_t)detoff[c];
  r.nopen = nopen[c];
  r.cph = cph[c];
  r.nsef = (int16_t)nsef[c];
  r.pad = 0;
  out[c] = r;
}

// XT[(e * nbf + f) * nbf + o] = (e o|o f)
__global__ void k_xt(int nbf, const double* __restrict__ two, double* __restrict__ xt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n3 = (int64_t)nbf * nbf * nbf;
  if (t >= n3) return;
  const int o = (int)(t % nbf), f = (int)((t / nbf) % nbf), e = (int)(t / ((int64_t)nbf * nbf));
  xt[t] = g2(two, e, o, o, f);
}

// Connected configurations of every I (all J, ascending), screened through
// LDS tiles of the configuration words; each wavefront handles kIPerWave rows.
template <bool FILL>
__global__ __launch_bounds__(256) void k_adj(int64_t ncfg, const uint64_t* __restrict__ dm,
                                             const uint64_t* __restrict__ sm, const int8_t* __restrict__ nopen,
                                             int32_t* __restrict__ cnt, const int64_t* __restrict__ off,
                                             int32_t* __restrict__ out) {
  __shared__ uint64_t tdm[kTile], tsm[kTile];
  __shared__ int8_t tno[kTile];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t I0 = ((int64_t)blockIdx.x * (blockDim.x / kWave) + wave) * kIPerWave;
  uint64_t dI[kIPerWave], sI[kIPerWave];
  int nI[kIPerWave];
  int64_t pos[kIPerWave];
#pragma unroll
  for (int u = 0; u < kIPerWave; ++u) {
    const int64_t I = I0 + u;
    const bool v = I < ncfg;
    dI[u] = v ? dm[I] : 0;
    sI[u] = v ? sm[I] : 0;
    nI[u] = v ? nopen[I] : 0;
    pos[u] = (FILL && v) ? off[I] : 0;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  for (int64_t T0 = 0; T0 < ncfg; T0 += kTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kTile; t += blockDim.x) {
      const int64_t J = T0 + t;
      if (J < ncfg) {
        tdm[t] = dm[J];
        tsm[t] = sm[J];
        tno[t] = nopen[J];
      } else {
        tdm[t] = 0;
        tsm[t] = 0;
        tno[t] = 100;  // never within 4 open shells of a real configuration
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kIPerWave; ++u) {
      if (I0 + u >= ncfg) break;  // uniform per wavefront
      for (int s = 0; s < kTile; s += kWave) {
        const int t = s + lane;
        const int dn = nI[u] - tno[t];
        const bool ok = dn <= 4 && dn >= -4 && moved(dI[u], sI[u], tdm[t], tsm[t]) <= 2;
        const uint64_t m = __ballot(ok);
        if (FILL && ok) out[pos[u] + __popcll(m & below)] = (int32_t)(T0 + t);
        pos[u] += __popcll(m);
      }
    }
  }
  if (!FILL && lane == 0) {
#pragma unroll
    for (int u = 0; u < kIPerWave; ++u)
      if (I0 + u < ncfg) cnt[I0 + u] = (int32_t)pos[u];
  }
}

__global__ void k_cnt64(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

// Record order within a configuration's segment: doubles, then singles, then
// the rest (pairs the reference skips, and the diagonal block); within each
// group by the number of open shells new in J and by J's open-shell count, so
// that the records the sigma kernel runs side by side enumerate alike.
// In an ODD-class space a closed shell keeps its Ms = 0 determinant while
// open-shell configurations have odd Ms only, so a closed-shell / open-shell
// pair has no connected determinants: such records go to the rest.
constexpr int kNClass = 90;
__device__ __forceinline__ int rec_class(const PairInfo& P, uint64_t sI, uint64_t sJ, int nJ, int odd) {
  const bool none = odd && ((sI == 0) != (sJ == 0));
  const int g = none ? 2 : (P.kind == 3 ? 0 : (P.kind == 2 ? 1 : 2));
  return g * 30 + popc(sJ & ~sI) * 6 + (nJ >> 1);
}

__device__ __forceinline__ void wave_sync_lds() {
  // LDS operations of one wavefront complete in order; the fences keep the
  // compiler from moving LDS reads and writes across this point
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// One record per ordered pair, in the reference orientation, sorted by class
// (stable: adjacency order within a class).  A wavefront per configuration:
// a class histogram in LDS, its scan, then placement by ballot rank.
__global__ __launch_bounds__(256) void k_records(int64_t ncfg, const int64_t* __restrict__ adjoff,
                                                 const int32_t* __restrict__ adj, const uint64_t* __restrict__ dm,
                                                 const uint64_t* __restrict__ sm, const int8_t* __restrict__ nopen,
                                                 const int32_t* __restrict__ nsef, int nbf,
                                                 const double* __restrict__ one, const double* __restrict__ two,
                                                 int odd, PairRec* __restrict__ rec, double* __restrict__ cdiag,
                                                 int32_t* __restrict__ ndbl, int32_t* __restrict__ nsgl,
                                                 unsigned long long* __restrict__ nelems) {
  __shared__ int s_base[4][kNClass];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t I = (int64_t)blockIdx.x * (blockDim.x / kWave) + wave;
  if (I >= ncfg) return;  // uniform per wavefront; wavefront-level synchronisation only
  int* base = s_base[wave];
  for (int c = lane; c < kNClass; c += kWave) base[c] = 0;
  wave_sync_lds();
  const int64_t a0 = adjoff[I], a1 = adjoff[I + 1];
  const uint64_t dI = dm[I], sI = sm[I];
  const int nI = nopen[I];
  auto setup = [&](int J, PairInfo& P, bool& rowI) {
    rowI = J <= I;
    const int R = rowI ? (int)I : J, C = rowI ? J : (int)I;
    pair_setup(P, R, C, rowI ? dI : dm[J], rowI ? sI : sm[J], rowI ? nI : nopen[J], rowI ? dm[J] : dI,
               rowI ? sm[J] : sI, rowI ? nopen[J] : nI, nbf, one, two);
  };
  for (int64_t t = a0 + lane; t < a1; t += kWave) {
    const int J = adj[t];
    PairInfo P;
    bool rowI;
    setup(J, P, rowI);
    atomicAdd(&base[rec_class(P, sI, sm[J], nopen[J], odd)], 1);
  }
  wave_sync_lds();
  if (lane == 0) {
    int run = 0;
    for (int c = 0; c < kNClass; ++c) {
      const int n = base[c];
      base[c] = run;
      run += n;
    }
    ndbl[I] = base[30];
    nsgl[I] = base[60] - base[30];
  }
  wave_sync_lds();
  const uint64_t below = (1ull << lane) - 1ull;
  unsigned long long ne = 0;
  for (int64_t t0 = a0; t0 < a1; t0 += kWave) {
    const int64_t t = t0 + lane;
    bool act = t < a1;
    int J = 0, cls = -1;
    PairInfo P;
    bool rowI = false;
    if (act) {
      J = adj[t];
      setup(J, P, rowI);
      cls = rec_class(P, sI, sm[J], nopen[J], odd);
      ne += (unsigned long long)nsef[J];
    }
    uint64_t pend = __ballot(act);
    while (pend) {
      const int k0 = __shfl(cls, ctz(pend), kWave);
      const uint64_t m = __ballot(act && cls == k0);
      if (act && cls == k0) {
        const int pos = base[k0] + popc(m & below);
        PairRec r = make_rec(P, J, rowI);
        const NewShells w = new_shells(sI, sm[J]);
        r.norb = w.norb;
        r.npos = w.npos | ((uint32_t)w.nn << 16);
        rec[a0 + pos] = r;
        if (J == (int)I) cdiag[I] = P.v0;
        act = false;
      }
      wave_sync_lds();
      if (lane == 0) base[k0] += popc(m);
      wave_sync_lds();
      pend &= ~m;
    }
  }
  for (int off = kWave / 2; off > 0; off >>= 1) ne += __shfl_xor(ne, off, kWave);
  if (lane == 0) atomicAdd(nelems, ne * (unsigned long long)nsef[I]);
}

template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c);
template <>
__device__ __forceinline__ double fmadd<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

struct DirectArgs {
  int64_t ncfg;
  const int64_t* recoff;
  const PairRec* rec;
  const ConfRec* conf;
  const uint64_t* detao;
  const double* xt;
  int nbf;
  double mult;
  const double* two;
  const int16_t* didx;
  int ndidx;
  const double* cdiag;
  const int32_t* ndbl;
  const int32_t* nsgl;
  int frac;              // one determinant per SEF (odd electron count)
  const int32_t* order;  // row processing order (null: natural)
  int chunk;             // XCD run length in workgroups (0: plain round-robin)
  int64_t row0, nrows;   // the slice of the row order to process (all rows but when tuning)
  int split;             // rows are work items I << 2 | q: SEFs [64q, 64q + 64) of I
  unsigned long long* count;  // K = 0 (counting launch): connected determinant pairs visited
};

__device__ __forceinline__ ConfView view_of(const ConfRec& r) {
  ConfView v;
  v.d = r.d;
  v.s = r.s;
  v.nopen = r.nopen;
  v.ph = r.cph;
  v.pat = nullptr;
  return v;
}

// The row (configuration, or work item of a split space) of this
// wavefront (-1: none).  Workgroups are dealt
// to the 8 XCDs round-robin; with chunk C > 0 each XCD takes runs of C
// consecutive workgroups of the row order in turn (the grid is a multiple of
// 8C), so rows that share column configurations meet in the same L2 while
// the XCDs stay evenly loaded.
__device__ __forceinline__ int64_t row_of_wave(const DirectArgs& a) {
  // uniform per wavefront; saying so keeps the row's configuration record,
  // its segment bounds and the loop counters in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  int64_t wg = blockIdx.x;
  if (a.chunk > 0) {
    const int64_t q = blockIdx.x / 8, x = blockIdx.x % 8;
    wg = (q / a.chunk) * 8 * a.chunk + x * a.chunk + q % a.chunk;
  }
  const int64_t slotw = wg * (blockDim.x / kWave) + wave;
  if (slotw >= a.nrows) return -1;
  return a.order ? a.order[a.row0 + slotw] : a.row0 + slotw;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef f32x4 type;
};
template <>
struct Vec16<double> {
  typedef f64x2 type;
};

// acc += hv * x[0..K): 16-byte loads when the rows are aligned
template <typename T, int K, bool VEC>
__device__ __forceinline__ void axpy_row(T hv, const T* __restrict__ x, T (&acc)[K]) {
  constexpr int W = 16 / (int)sizeof(T);
  if constexpr (VEC && K % W == 0) {
    typedef typename Vec16<T>::type V;
    const V* __restrict__ xv = reinterpret_cast<const V*>(x);
#pragma unroll
    for (int q = 0; q < K / W; ++q) {
      const V u = xv[q];
#pragma unroll
      for (int w = 0; w < W; ++w) acc[q * W + w] = fmadd(hv, u[w], acc[q * W + w]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] = fmadd(hv, x[v], acc[v]);
  }
}

// Contribution of one pair record (kind KIND) to the lane's row, determinant
// aa of I: coef * D(aa, b) * x_J[b / 2] for every connected determinant b.
template <typename T, int K, bool VEC, int KIND, bool FRAC>
__device__ __forceinline__ void pair_rows(const DirectArgs& a, const PairRec& r, const ConfView& A, uint64_t aa,
                                          const int16_t* __restrict__ s_didx, const T* __restrict__ X, int ldx,
                                          T (&acc)[K > 0 ? K : 1], unsigned long long& nconn) {
  PairInfo P;
  bool rowI;
  int J;
  unpack_rec(r, P, J, rowI);
  const ConfRec rj = a.conf[J];
  const ConfView B = view_of(rj);
  const double* __restrict__ xrow = KIND == 2 ? a.xt + ((int64_t)P.e0 * a.nbf + P.f0) * a.nbf : nullptr;
  const PairCoef pc = pair_coef(A.nopen, B.nopen, a.mult, FRAC);
  constexpr int sh = FRAC ? 0 : 1;  // determinant b -> SEF b >> sh
  const T* __restrict__ xJ = X + (int64_t)rj.sefoff * ldx;
  const NewShells w{r.norb, r.npos & 0xffffu, (int)(r.npos >> 16)};
  for_each_conn<KIND == 3 ? 2 : 1>(A, B, aa, w, s_didx + (1 << B.nopen) - 1, [&](int b, uint64_t ab) {
    if constexpr (K == 0) {
      ++nconn;
    } else {
      const double v = conn_value<KIND>(P, xrow, rowI, A, B, aa, ab, a.two);
      if (v == 0.0) return;
      const T hv = (T)(((b & 1) ? pc.co : pc.ce) * v);
      axpy_row<T, K, VEC>(hv, xJ + (int64_t)(b >> sh) * ldx, acc);
    }
  });
}

// One wavefront per row configuration I; lanes are (record slot, SEF i of I)
// with nsef_I * nslot = 64, so 64 / nsef_I records are in flight per step.
// Each lane takes determinant 2i of I and, per record, enumerates the
// determinants of J it connects to (for_each_conn), adding
// coef * D(2i, b) * x_J[b / 2].  Lanes of one i are summed across slots at
// the end; the order is fixed, so results are deterministic.  A
// configuration with more than 64 SEFs (9 or 10 open shells: 128 or 256)
// is split over wavefronts of 64 SEFs each, one record slot per lane; every
// SEF count is a power of two, so the slots divide the wavefront.  K = 0 is
// a counting launch: the same enumeration, no element values or products,
// the connected determinant pairs summed into a.count.
template <typename T, int K, bool VEC, bool FRAC>
__global__ __launch_bounds__(256) void k_direct(DirectArgs a, const T* __restrict__ X, int ldx, T* __restrict__ Y,
                                                int ldy) {
  __shared__ int16_t s_didx[2047];
  for (int t = threadIdx.x; t < a.ndidx; t += blockDim.x) s_didx[t] = a.didx[t];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t item = row_of_wave(a);
  if (item < 0) return;  // uniform per wavefront; no block-wide barrier below
  const int64_t I = a.split ? item >> 2 : item;
  const int i0 = a.split ? (int)(item & 3) * kWave : 0;
  const ConfRec ri = a.conf[I];
  const ConfView A = view_of(ri);
  const int nW = ri.nsef - i0 < kWave ? ri.nsef - i0 : kWave;  // SEFs of I in this wavefront
  const int nslot = kWave / nW;
  const int slot = lane / nW, i = i0 + lane - slot * nW;
  constexpr int sh = FRAC ? 0 : 1;  // SEF i <- determinant i << sh
  const uint64_t aa = A.nopen ? a.detao[ri.detoff + (i << sh)] : 0ull;
  T acc[K > 0 ? K : 1];
#pragma unroll
  for (int v = 0; v < K; ++v) acc[v] = (T)0;
  unsigned long long nconn = 0;
  if (slot == 0) {
    // the diagonal block (J = I): determinant 2i itself and its alpha-beta
    // spin exchanges, from the configuration's vnt[0] and the integral table
    PairInfo P;
    P.I = P.J = (int32_t)I;
    P.kind = 1;
    P.e0 = P.e1 = P.f0 = P.f1 = -1;
    P.v0 = a.cdiag[I];
    P.v1 = 0.0;
    const PairCoef pc = pair_coef(A.nopen, A.nopen, a.mult, FRAC);
    const T* __restrict__ xI = X + (int64_t)ri.sefoff * ldx;
    for_each_conn_diag(A, aa, s_didx + (1 << A.nopen) - 1, [&](int b, uint64_t ab) {
      if constexpr (K == 0) {
        ++nconn;
      } else {
        const double v = detval_ao(P, A, A, aa, ab, a.two);
        if (v == 0.0) return;
        const T hv = (T)(((b & 1) ? pc.co : pc.ce) * v);
        axpy_row<T, K, VEC>(hv, xI + (int64_t)(b >> sh) * ldx, acc);
      }
    });
  }
  // double excitations, then single excitations (the segment's order)
  const int64_t r0 = a.recoff[I];
  const int64_t r1 = r0 + a.ndbl[I], r2 = r1 + a.nsgl[I];
  for (int64_t t0 = r0; t0 < r1; t0 += nslot) {
    const int64_t t = t0 + slot;
    if (t < r1) pair_rows<T, K, VEC, 3, FRAC>(a, a.rec[t], A, aa, s_didx, X, ldx, acc, nconn);
  }
  for (int64_t t0 = r1; t0 < r2; t0 += nslot) {
    const int64_t t = t0 + slot;
    if (t < r2) pair_rows<T, K, VEC, 2, FRAC>(a, a.rec[t], A, aa, s_didx, X, ldx, acc, nconn);
  }
  if constexpr (K == 0) {
    for (int off = kWave / 2; off > 0; off >>= 1) nconn += __shfl_xor(nconn, off, kWave);
    if (lane == 0) atomicAdd(a.count, nconn);
    return;
  }
  for (int off = nW; off < kWave; off <<= 1) {
#pragma unroll
    for (int v = 0; v < K; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
  }
  if (slot == 0) {
    T* y = Y + ((int64_t)ri.sefoff + i) * (int64_t)ldy;
#pragma unroll
    for (int v = 0; v < K; ++v) y[v] = acc[v];
  }
}


DirectArgs direct_args(Ctx& c, int pi) {
  DirectArgs a;
  a.ncfg = c.sp.ncfg;
  a.recoff = c.d_adjoff.p;
  a.rec = c.d_rec.p;
  a.conf = c.d_conf.p;
  a.detao = c.d_detao.p;
  a.xt = c.d_xt.p;
  a.nbf = c.sp.nbf;
  a.mult = c.sp.mult;
  a.frac = c.sp.frac ? 1 : 0;
  a.two = c.d_two.p;
  a.didx = c.d_didx.p;
  a.ndidx = (2 << c.sp.max_open) - 1;
  a.cdiag = c.d_cdiag.p;
  a.ndbl = c.d_ndbl.p;
  a.nsgl = c.d_nsgl.p;
  a.order = c.direct_order[pi] ? c.d_order.p : (c.direct_split ? c.d_items.p : nullptr);
  a.chunk = c.direct_chunk[pi];
  a.row0 = 0;
  a.nrows = c.direct_nrows;
  a.split = c.direct_split ? 1 : 0;
  a.count = nullptr;
  return a;
}

template <typename T, int K>
void launch_direct(Ctx& c, const T* X, int ldx, T* Y, int ldy, int64_t row0 = 0, int64_t nrows = -1,
                   unsigned long long* count = nullptr) {
  DirectArgs a = direct_args(c, sizeof(T) == 4 ? 1 : 0);
  a.count = count;
  if (nrows >= 0) {
    a.row0 = row0;
    a.nrows = nrows;
  }
  const bool vec = ((uintptr_t)X % 16 == 0) && ((size_t)ldx * sizeof(T)) % 16 == 0;
  const int m = 8 * (a.chunk > 0 ? a.chunk : 1);
  const unsigned grid = m * nblk(nblk(a.nrows, 4), m);
  if (a.frac) {
    if (vec)
      hipLaunchKernelGGL((k_direct<T, K, true, true>), dim3(grid), dim3(256), 0, c.stream, a, X, ldx, Y, ldy);
    else
      hipLaunchKernelGGL((k_direct<T, K, false, true>), dim3(grid), dim3(256), 0, c.stream, a, X, ldx, Y, ldy);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_direct<T, K, true, false>), dim3(grid), dim3(256), 0, c.stream, a, X, ldx, Y, ldy);
    else
      hipLaunchKernelGGL((k_direct<T, K, false, false>), dim3(grid), dim3(256), 0, c.stream, a, X, ldx, Y, ldy);
  }
  PSOCI_HIP(hipGetLastError());
}

template <typename T>
void direct_any_k(Ctx& c, int k, const T* X, int ldx, T* Y, int ldy, const Staging& st) {
  pad::blocks<T>(c, c.sp.nsef_total, k, 32, X, ldx, Y, ldy, [&](int K, const T* x, int lx, T* y, int ly) {
    switch (K) {
      case 32: launch_direct<T, 32>(c, x, lx, y, ly); break;
      case 24: launch_direct<T, 24>(c, x, lx, y, ly); break;
      case 16: launch_direct<T, 16>(c, x, lx, y, ly); break;
      case 8: launch_direct<T, 8>(c, x, lx, y, ly); break;
      case 4: launch_direct<T, 4>(c, x, lx, y, ly); break;
      case 2: launch_direct<T, 2>(c, x, lx, y, ly); break;
      default: launch_direct<T, 1>(c, x, lx, y, ly); break;
    }
  }, true, st);
}

// Schedule autotuner: the row order (natural, or by occupation string from
// orbital 0) and the XCD run length decide how often column configurations
// are found in L2, which depends on the space's structure and on the x-row
// size, so each precision is tuned on its own: every candidate runs the
// 16-column kernel of that precision on the middle eighth of its row order
// (at least 65,536 rows, or all) twice; the fastest is kept.
// PSOCI_DIRECT_ORDER / PSOCI_DIRECT_CHUNK fix the choice instead.
template <typename T>
void tune_direct_prec(Ctx& c, int pi, hipEvent_t e0, hipEvent_t e1) {
  const int64_t nr = c.direct_nrows, n = c.sp.nsef_total;
  const int64_t rows = std::min<int64_t>(nr, std::max<int64_t>(nr / 8, 65536));
  const int64_t row0 = (nr - rows) / 2;
  const int K = 16;
  DevBuf<T> X, Y;
  X.alloc((size_t)n * K);
  Y.alloc((size_t)n * K);
  PSOCI_HIP(hipMemsetAsync(X.p, 0, (size_t)n * K * sizeof(T), c.stream));
  const int chunks[] = {0, 4, 16, 64};
  double best = 1e300;
  bool best_order = false;
  int best_chunk = 0;
  for (int o = 0; o < 2; ++o)
    for (int ch : chunks) {
      c.direct_order[pi] = o == 1;
      c.direct_chunk[pi] = ch;
      double t = 1e300;
      for (int r = 0; r < 2; ++r) {
        PSOCI_HIP(hipEventRecord(e0, c.stream));
        launch_direct<T, 16>(c, X.p, K, Y.p, K, row0, rows);
        PSOCI_HIP(hipEventRecord(e1, c.stream));
        PSOCI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        PSOCI_HIP(hipEventElapsedTime(&ms, e0, e1));
        t = std::min(t, (double)ms);
      }
      if (t < best) {
        best = t;
        best_order = c.direct_order[pi];
        best_chunk = ch;
      }
    }
  c.direct_order[pi] = best_order;
  c.direct_chunk[pi] = best_chunk;
  c.direct_tune_ms[pi] = best;
}

void tune_direct(Ctx& c) {
  const char* eo = std::getenv("PSOCI_DIRECT_ORDER");
  const char* ec = std::getenv("PSOCI_DIRECT_CHUNK");
  if (eo || ec) {
    for (int pi = 0; pi < 2; ++pi) {
      c.direct_order[pi] = eo && eo[0] == '1';
      c.direct_chunk[pi] = ec ? std::max(0, std::atoi(ec)) : 0;
      c.direct_tune_ms[pi] = -1.0;
    }
    return;
  }
  hipEvent_t e0, e1;
  PSOCI_HIP(hipEventCreate(&e0));
  PSOCI_HIP(hipEventCreate(&e1));
  tune_direct_prec<float>(c, 1, e0, e1);
  tune_direct_prec<double>(c, 0, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

}  // namespace

void build_pairs(Ctx& c) {
  const Space& sp = c.sp;
  c.direct_det_pairs = -1;
  // the pair records and the adjoint take one determinant class; spaces with
  // spin-orbit blocks (configurations of more than one jsym) run stored
  PSOCI_CHECK(!sp.mixed, PSOCI_ERR_UNSUPPORTED,
              "the direct sigma and the adjoint take configurations of one jsym; a space with spin-orbit blocks "
              "needs the stored Hamiltonian (psoci_build_h)");
  PSOCI_CHECK(sp.ndet_total < ((int64_t)1 << 31), PSOCI_ERR_UNSUPPORTED, "more than 2^31 determinants");
  const int64_t ncfg = sp.ncfg;
  const int max_nsef = *std::max_element(sp.nsef.begin(), sp.nsef.end());
  PSOCI_CHECK(max_nsef <= 4 * kWave, PSOCI_ERR_UNSUPPORTED, "direct sigma supports at most 256 SEFs per configuration");
  c.direct_split = max_nsef > kWave;
  if (c.direct_split)
    PSOCI_CHECK(ncfg < ((int64_t)1 << 29), PSOCI_ERR_UNSUPPORTED,
                "direct sigma with more than 64 SEFs per configuration takes fewer than 2^29 configurations");
  // a split space's rows are work items: configuration I << 2 | q for each
  // 64-SEF quarter q of I
  auto items_of = [&](const std::vector<int32_t>& rows) {
    std::vector<int32_t> it;
    it.reserve(rows.size());
    for (int32_t r : rows)
      for (int q = 0; q * kWave < sp.nsef[r]; ++q) it.push_back((int32_t)(((int64_t)r << 2) | q));
    return it;
  };
  std::vector<int32_t> ord, nat_items;  // host sources of the uploads below, alive until the sync
  {
    // alternative row order: lexicographic in the occupation string from
    // orbital 0 up (in an MR-CISD list with the internal orbitals first this
    // groups rows by internal part, across external patterns)
    std::vector<unsigned __int128> key(ncfg);
    for (int64_t i = 0; i < ncfg; ++i) {
      unsigned __int128 k = 0;
      for (int o = 0; o < sp.nbf; ++o) {
        const unsigned v = 2u * (unsigned)((sp.dm[i] >> o) & 1) + (unsigned)((sp.sm[i] >> o) & 1);
        k |= (unsigned __int128)v << (2 * (63 - o));
      }
      key[i] = k;
    }
    ord.resize(ncfg);
    for (int64_t i = 0; i < ncfg; ++i) ord[i] = (int32_t)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) { return key[x] > key[y]; });
    if (c.direct_split) {
      std::vector<int32_t> nat(ncfg);
      for (int64_t i = 0; i < ncfg; ++i) nat[i] = (int32_t)i;
      nat_items = items_of(nat);
      ord = items_of(ord);
      c.d_items.upload(nat_items, c.stream);
      c.direct_nrows = (int64_t)nat_items.size();
    } else {
      c.direct_nrows = ncfg;
    }
    c.d_order.upload(ord, c.stream);
  }
  c.d_detoff.upload(sp.det_off, c.stream);
  c.d_detao.alloc(sp.ndet_total);
  hipLaunchKernelGGL(k_detao, dim3(nblk(ncfg, 4)), dim3(256), 0, c.stream, ncfg, c.d_detoff.p, c.d_sm.p, c.d_nopen.p,
                     c.d_pat.p, c.d_patoff.p, c.d_detao.p);
  PSOCI_HIP(hipGetLastError());
  c.d_conf.alloc(ncfg);
  hipLaunchKernelGGL(k_confrec, dim3(nblk(ncfg, 256)), dim3(256), 0, c.stream, ncfg, c.d_dm.p, c.d_sm.p, c.d_nopen.p,
                     c.d_cph.p, c.d_nsef.p, c.d_sefoff.p, c.d_detoff.p, c.d_conf.p);
  PSOCI_HIP(hipGetLastError());
  const int nbf = sp.nbf;
  const int64_t n3 = (int64_t)nbf * nbf * nbf;
  c.d_xt.alloc(n3);
  hipLaunchKernelGGL(k_xt, dim3(nblk(n3, 256)), dim3(256), 0, c.stream, nbf, c.d_two.p, c.d_xt.p);
  PSOCI_HIP(hipGetLastError());
  // adjacency: count, scan, fill
  DevBuf<int32_t> cnt, adj;
  DevBuf<int64_t> cnt64;
  cnt.alloc(ncfg);
  cnt64.alloc(ncfg + 1);
  const unsigned grid = nblk(ncfg, 4 * kIPerWave);
  hipLaunchKernelGGL(k_adj<false>, dim3(grid), dim3(256), 0, c.stream, ncfg, c.d_dm.p, c.d_sm.p, c.d_nopen.p, cnt.p,
                     nullptr, nullptr);
  PSOCI_HIP(hipGetLastError());
  PSOCI_HIP(hipMemsetAsync(cnt64.p + ncfg, 0, sizeof(int64_t), c.stream));
  hipLaunchKernelGGL(k_cnt64, dim3(nblk(ncfg, 256)), dim3(256), 0, c.stream, ncfg, cnt.p, cnt64.p);
  c.d_adjoff.alloc(ncfg + 1);
  c.scan(cnt64.p, c.d_adjoff.p, ncfg + 1);
  int64_t total = 0;
  PSOCI_HIP(hipMemcpyAsync(&total, c.d_adjoff.p + ncfg, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
  c.sync();
  adj.alloc(total);
  hipLaunchKernelGGL(k_adj<true>, dim3(grid), dim3(256), 0, c.stream, ncfg, c.d_dm.p, c.d_sm.p, c.d_nopen.p, nullptr,
                     c.d_adjoff.p, adj.p);
  PSOCI_HIP(hipGetLastError());
  // pair records (the adjacency is folded into them and released)
  c.d_rec.alloc(total);
  c.d_cdiag.alloc(ncfg);
  c.d_ndbl.alloc(ncfg);
  c.d_nsgl.alloc(ncfg);
  DevBuf<unsigned long long> ne;
  ne.alloc(1);
  PSOCI_HIP(hipMemsetAsync(ne.p, 0, sizeof(unsigned long long), c.stream));
  hipLaunchKernelGGL(k_records, dim3(nblk(ncfg, 4)), dim3(256), 0, c.stream, ncfg, c.d_adjoff.p, adj.p, c.d_dm.p,
                     c.d_sm.p, c.d_nopen.p, c.d_nsef.p, nbf, c.d_one.p, c.d_two.p, sp.odd ? 1 : 0, c.d_rec.p,
                     c.d_cdiag.p, c.d_ndbl.p, c.d_nsgl.p, ne.p);
  PSOCI_HIP(hipGetLastError());
  unsigned long long nelems = 0;
  ne.download(&nelems, 1, c.stream);
  c.sync();
  c.direct_elems = (int64_t)nelems;
  c.nadj = total;
  c.have_pairs = true;
  tune_direct(c);
}

int64_t direct_det_pairs(Ctx& c) {
  if (!c.have_pairs) build_pairs(c);
  if (c.direct_det_pairs < 0) {
    DevBuf<unsigned long long> cnt;
    cnt.alloc(1);
    PSOCI_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), c.stream));
    launch_direct<double, 0>(c, nullptr, 0, nullptr, 0, 0, -1, cnt.p);
    unsigned long long v = 0;
    cnt.download(&v, 1, c.stream);
    c.sync();
    c.direct_det_pairs = (int64_t)v;
  }
  return c.direct_det_pairs;
}

void sigma_direct(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy, const Staging& st) {
  if (st.p) PSOCI_CHECK(direct_engine(c), PSOCI_ERR_STATE, "the direct sigma's operator is not built");
  else if (direct_needs_build(c)) build_direct(c);
  if (direct_engine(c) == 2) {
    sigma_ras(c, prec, k, X, ldx, Y, ldy);
    return;
  }
  if (prec == PSOCI_FP32)
    direct_any_k<float>(c, k, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy, st);
  else
    direct_any_k<double>(c, k, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy, st);
}

}  // namespace psoci
