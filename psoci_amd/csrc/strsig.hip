// String-driven direct sigma on the GPU (the plan and its index arithmetic
// are in ras.hpp / ras.cpp).
//
// Replaces, for complete RAS spaces (full CI and MR-CISD lists), the direct
// path of the reference (matrixVectorProductsScatterAccumulateLowLevel...
// DirectFetch, PsociGAhamiltonian.C:7549-7823, with matrixVectorUpdate
// :13036-13115), which regenerates every configuration-pair block.  One
// sigma call of K columns is, per chunk of KC <= 16 columns (FP64: 8 when
// device memory is short):
//   k_ras_x2c   SEF vectors -> determinant vectors c[2 det + s][KC] = (-1)^s
//               c_det (each determinant belongs to one SEF: a scatter
//               without collisions; both signs, so a product reads an entry
//               of either sign without arithmetic), rows 2 ndet and 2 ndet +
//               1 zero;
//   k_ras_gemm  per block: W[row][p-pair][KC] = sum over q-pairs of
//               G[p-pair][q-pair] D[row][q-pair][KC], D gathered from c into
//               LDS through the signed row of each entry (built once from
//               the string addition tables by k_ras_didx), the product on
//               MFMA (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64);
//   k_ras_y     per SEF and column: y = scale * sum of +-W over the RI
//               terms of its determinants (the term list -- signed W offsets
//               -- is built once at setup by k_ras_terms), written once (a
//               gather: deterministic, no atomics).
// HBM layout: c [2 ndet + 2][KC] (64 B per row at KC = 16 FP32), W [block
// rows][nqp][KC] so that a term of k_ras_y reads KC contiguous values.
#include <algorithm>
#include <cstdlib>

#include "ctx.hpp"
#include "ras.hpp"

namespace psoci {

namespace {

// GEMM tile shape, LDS layout and schedule (compile-time tuning knobs; the
// defaults are the fastest measured on mrci10x24: FP32 16-pair chunks
// without register prefetch, which keeps 6 wavefronts per SIMD, and FP64
// 16-pair chunks with the next chunk's loads issued before the MFMA steps)
#ifndef PSOCI_RAS_QC
#define PSOCI_RAS_QC 16
#endif
#ifndef PSOCI_RAS_QC64
#define PSOCI_RAS_QC64 16
#endif
#ifndef PSOCI_RAS_GPAD
#define PSOCI_RAS_GPAD 1
#endif
#ifndef PSOCI_RAS_DPAD
#define PSOCI_RAS_DPAD 4
#endif
#ifndef PSOCI_RAS_VSTORE
#define PSOCI_RAS_VSTORE 1
#endif
#ifndef PSOCI_RAS_PIPE
#define PSOCI_RAS_PIPE 0
#endif
#ifndef PSOCI_RAS_PIPE64
#define PSOCI_RAS_PIPE64 1
#endif
// 16- and 8-column chunks through the LDS-copy product k_ras_gemm_lds (0:
// the register-staged k_ras_gemm for every width)
#ifndef PSOCI_RAS_DMA
#define PSOCI_RAS_DMA 1
#endif
// FP32 LDS-copy products in chunks of 32 pairs (0: 16)
#ifndef PSOCI_RAS_LQC32
#define PSOCI_RAS_LQC32 0
#endif
constexpr int kMT = 128;            // (row, column) values per GEMM tile
// q-pairs per LDS chunk (FP32, FP64)
template <typename T>
constexpr int ras_qc() {
  return sizeof(T) == 8 ? PSOCI_RAS_QC64 : PSOCI_RAS_QC;
}

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

// The signed rows of every SEF's determinant coefficients (built once):
// rows.x = 2 det + negative of its first determinant; rows.y the same for
// the second, or -1 (one determinant, coefficient 1/sqrt(2)) or -2 (a
// closed shell, coefficient 1); -3 marks a determinant outside the plan
// (never, for a plan that passes its host checks).
__global__ __launch_bounds__(256) void k_ras_sefrows(const RasView* __restrict__ vp, int2* __restrict__ rows) {
  const RasView& v = *vp;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= v.nsef) return;
  const RasSef e = v.sef[s];
  int r[2] = {-3, (e.flag[0] & 2) ? -2 : -1};
  for (int d = 0; d < e.n; ++d) {
    const RasSectorDev& S = v.sec[e.sec[d]];
    const int ca = ras_cls_of(v, S.na, e.a[d]), cb = ras_cls_of(v, S.nb, e.b[d]);
    const int64_t det =
        ras_det(v, e.sec[d], ca, e.a[d] - v.lev[S.na].cls_off[ca], cb, e.b[d] - v.lev[S.nb].cls_off[cb]);
    r[d] = (det >= 0 && det < v.ndet) ? (int)(2 * det + (e.flag[d] & 1)) : -3;
  }
  rows[s] = make_int2(r[0], r[1]);
}

// c[row][KC] = (-1)^(row & 1 ^ negative) * magnitude * x[sef][col0 + q]
// (zero for q >= kw) for both signed rows of each of the SEF's
// determinants; the zero rows 2 ndet, 2 ndet + 1.  X may be wider than T
// (FP64 vectors rounded to the FP32 computation on the fly).
template <typename T, int KC, typename TX = T>
__global__ __launch_bounds__(256) void k_ras_x2c(int64_t nsef, int64_t ndet, const int2* __restrict__ rows,
                                                 const TX* __restrict__ X, int ldx, int col0, int kw,
                                                 T* __restrict__ c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * KC) c[2 * ndet * KC + t] = (T)0;
  const int64_t s = t / KC;
  const int q = (int)(t % KC);
  if (s >= nsef) return;
  const int2 r = rows[s];
  const T x = q < kw ? (T)X[s * (int64_t)ldx + col0 + q] : (T)0;
  const T y = (r.y == -2 ? (T)1 : (T)0x1.6a09e667f3bcdp-1) * x;
  if (r.x >= 0) {
    c[(int64_t)r.x * KC + q] = y;
    c[(int64_t)(r.x ^ 1) * KC + q] = -y;
  }
  if (r.y >= 0) {
    c[(int64_t)r.y * KC + q] = y;
    c[(int64_t)(r.y ^ 1) * KC + q] = -y;
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

// Signed vector row of every D entry (built once): didx[dbase + row * nqp +
// j] = 2 det + negative for pair j of block row `row`, the zero row 2 ndet
// when the pair cannot be added to the row's strings (ras_d_entry) and for
// the padding rows past nrows (up to a multiple of 16, so that whole tiles
// read their entries without bounds checks).
__global__ __launch_bounds__(256) void k_ras_didx(const RasView* __restrict__ vp, int bi, int32_t* __restrict__ didx) {
  const RasView& v = *vp;
  const RasBlockDev& B = v.blk[bi];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (B.nrows + 15) / 16 * 16 * B.nqp) return;
  const int64_t row = e / B.nqp;
  const int j = (int)(e % B.nqp);
  int neg = 0;
  const int64_t det = row < B.nrows ? ras_d_entry(v, B, row, j, neg) : -1;
  didx[B.dbase + e] = (det < 0 || det >= v.ndet) ? (int32_t)(2 * v.ndet) : (int32_t)(2 * det + neg);
}

// One tile of W = D G^T for block bi: PT p-pairs x kMT (row, column)
// values.  Four wavefronts, each PT p-pairs x 32 values (PT / 16 x 2 MFMA
// tiles); PT = 64, or the whole (16-padded) pair count when it is at most
// 112, so that no tile computes rows past it.
// Per chunk of kQC q-pairs: G[p][q] rows and the D entries (c rows of the
// determinants in didx) into LDS, then kQC / 4 MFMA steps; the global loads
// of the next chunk are issued before the MFMA steps of the current one
// (software pipelining through registers).  Tiles are dealt so that XCD x
// runs the contiguous tiles [x per, (x + 1) per): the p-tiles of one row
// tile, which gather the same c rows, share an L2.
template <typename T, int KC, int PT>
__global__ __launch_bounds__(256) void k_ras_gemm(const RasBlockDev* __restrict__ blk, int bi,
                                                  const int32_t* __restrict__ didx, const T* __restrict__ G,
                                                  const T* __restrict__ c, int32_t zrow, T* __restrict__ W,
                                                  int64_t ntiles, int npt) {
  typedef Mfma<T> M;
  constexpr int kQC = ras_qc<T>();
  typedef typename M::v4 v4;
  constexpr int RT = kMT / KC;          // block rows per tile
  constexpr int DPT = kMT * kQC / 256;        // D values per thread and chunk
  constexpr int VPT = KC < DPT ? KC : DPT;     // consecutive columns one thread gathers per entry
  constexpr int CPT = DPT / VPT;               // (row, pair) entries per thread and chunk
  constexpr int TPC = KC / VPT;                // threads per entry
  constexpr int GN = (PT * kQC + 255) / 256;   // G values per thread and chunk (at most)
  constexpr int NI = PT / 16;                  // p-subtiles per wavefront
  constexpr bool pipe = sizeof(T) == 8 ? PSOCI_RAS_PIPE64 : PSOCI_RAS_PIPE;
  // row strides of the LDS tiles (compile-time knobs)
  __shared__ __align__(16) T Gs[PT][kQC + PSOCI_RAS_GPAD];
  __shared__ __align__(16) T Ds[kQC][kMT + PSOCI_RAS_DPAD];
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;  // uniform per workgroup
  const int64_t mt = tile / npt;
  const int pt = (int)(tile % npt);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const RasBlockDev& B = blk[bi];
  const int64_t nrows = B.nrows;
  const int nqp = B.nqp;
  const T* __restrict__ Gb = G + B.goff;
  const int32_t* __restrict__ dx = didx + B.dbase;
  v4 acc[NI][2];
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};

  // per-thread staging of one chunk: G values tid + 256 i (row e / kQC,
  // pair e % kQC of the tile's G slice), D entries below
  T gv[GN];
  T dv[CPT][VPT];
  // chunks of kQC pairs; nqp is a multiple of 16, so the last chunk may
  // hold 16: its other half is loaded as zeros
  auto load = [&](int q0) {
    const int wq = nqp - q0 < kQC ? nqp - q0 : kQC;
#pragma unroll
    for (int i = 0; i < GN; ++i) {
      const int e = tid + 256 * i, gp = e / kQC, gq = e % kQC, gpg = pt * PT + gp;
      gv[i] = (e < PT * kQC && gpg < nqp && gq < wq) ? Gb[(int64_t)gpg * nqp + q0 + gq] : (T)0;
    }
#pragma unroll
    for (int ci = 0; ci < CPT; ++ci) {
      const int ent = (tid / TPC) * CPT + ci;  // (row, pair) entry: rl * kQC + jq
      const int v0 = (tid % TPC) * VPT;
      const int rl = ent / kQC, jq = ent % kQC;
      const int64_t row = mt * RT + rl;
      // the entry's signed row (the zero row past the block or the pairs)
      const int32_t e = (row < nrows && jq < wq) ? dx[row * nqp + q0 + jq] : zrow;
      const T* __restrict__ src = c + (int64_t)e * KC + v0;
      if constexpr (VPT * sizeof(T) % 16 == 0) {
        typedef typename VecN<T, 16 / sizeof(T)>::type V;
        constexpr int W16 = 16 / sizeof(T);
#pragma unroll
        for (int i = 0; i < VPT / W16; ++i) {
          const V x = reinterpret_cast<const V*>(src)[i];
#pragma unroll
          for (int w = 0; w < W16; ++w) dv[ci][i * W16 + w] = x[w];
        }
      } else {
#pragma unroll
        for (int i = 0; i < VPT; ++i) dv[ci][i] = src[i];
      }
    }
  };
  auto store = [&]() {
    typedef typename VecN<T, 16 / sizeof(T)>::type V;
    constexpr int W16 = 16 / sizeof(T);
    constexpr bool vec = PSOCI_RAS_VSTORE && (kMT + PSOCI_RAS_DPAD) % W16 == 0;
#pragma unroll
    for (int i = 0; i < GN; ++i) {
      const int e = tid + 256 * i;
      if (e < PT * kQC) Gs[e / kQC][e % kQC] = gv[i];
    }
#pragma unroll
    for (int ci = 0; ci < CPT; ++ci) {
      const int ent = (tid / TPC) * CPT + ci;
      const int v0 = (tid % TPC) * VPT;
      const int rl = ent / kQC, jq = ent % kQC;
      if constexpr (vec && VPT % W16 == 0) {
#pragma unroll
        for (int i = 0; i < VPT / W16; ++i) {
          V x;
#pragma unroll
          for (int w = 0; w < W16; ++w) x[w] = dv[ci][i * W16 + w];
          *reinterpret_cast<V*>(&Ds[jq][rl * KC + v0 + i * W16]) = x;
        }
      } else {
#pragma unroll
        for (int i = 0; i < VPT; ++i) Ds[jq][rl * KC + v0 + i] = dv[ci][i];
      }
    }
  };
  if (pipe) {
    load(0);
    store();
    __syncthreads();
  }
  for (int q0 = 0; q0 < nqp; q0 += kQC) {
    const bool more = q0 + kQC < nqp;
    if (pipe) {
      if (more) load(q0 + kQC);  // in flight during the MFMA steps below
    } else {
      load(q0);
      store();
      __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < kQC / 4; ++kk) {
      T a[NI], b[2];
#pragma unroll
      for (int i = 0; i < NI; ++i) a[i] = Gs[16 * i + (lane & 15)][4 * kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Ds[4 * kk + (lane >> 4)][32 * wave + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::op(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
    if (pipe) {
      if (more) store();
      __syncthreads();
    }
  }
  const int64_t wb = B.wbase;
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int m = 32 * wave + 16 * j + (lane & 15);
      const int64_t row = mt * RT + m / KC;
      const int vq = m % KC;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int p = pt * PT + 16 * i + M::row(t, lane);
        if (row < nrows && p < nqp) W[(wb + row * nqp + p) * KC + vq] = acc[i][j][t];
      }
    }
}

// The 16- and 8-column product with its operands copied straight from global
// memory into LDS (global_load_lds_dwordx4: each lane names its own 16-byte
// source, the wavefront's 64 pieces land contiguously at one LDS address),
// into two LDS stages: the copies of chunk n + 1 are in flight while the
// MFMA steps consume chunk n, with one barrier per chunk and no staging
// registers.  Layouts in LDS:
//   G stage [q][PT] -- q-pair rows of the tile's p-pairs, read from G's q
//     rows (G is symmetric, so row q holds column q); at FP32 the 16-byte
//     pieces of odd rows swap halves of every 128 bytes when PT / 16 is even,
//     which keeps the two rows one 32-lane operand read touches in separate
//     banks;
//   D stage [row][q][KC] -- one KC-value piece per (row, pair) entry, copied
//     from the entry's signed row of c (the zero row for a missing
//     determinant or a padding row), so no arithmetic touches the operands
//     between the copy and the MFMA; wave w copies and reads only its own
//     rows of the tile (2w, 2w + 1 at 16 columns; 4w .. 4w + 3 at 8).
// wavefronts per SIMD the LDS-copy product is compiled for (a register
// budget: the compiler would otherwise hold more k-steps' operands in flight
// than occupancy can afford; at FP32 and 64 p-pairs the budget of 6 spills a
// few set-up values outside the chunk loop, measured 2 % faster than 5)
template <typename T, int PT>
constexpr int ras_lds_waves() {
  return sizeof(T) == 8 ? (PT <= 64 ? 3 : 2) : (PT <= 64 ? 6 : 4);
}

template <int N>
struct IntC {
  static constexpr int value = N;
};

// 16 bytes from this lane's global address to LDS at byte address lds + 16
// * lane (lds uniform over the wavefront); completion is counted by vmcnt.  Written as
// an asm statement so that the compiler, which cannot tell the two LDS
// stages apart, does not wait for every copy in flight before each LDS read:
// the kernel waits for its copies itself (vmcnt(0) before the barrier that
// publishes a stage).  M0 carries the LDS address; nothing else in the
// kernels that use this keeps a value in M0.
__device__ __forceinline__ void lds_copy16(const void* src, uint32_t lds) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" ::"v"(src), "s"(lds) : "memory", "m0");
}
// LDS byte address of a __shared__ location, uniform over the wavefront
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p);
}

template <typename T, int KC, int PT, int QC = 16, int WPE = ras_lds_waves<T, PT>()>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_ras_gemm_lds(const RasBlockDev* __restrict__ blk, int bi, const int32_t* __restrict__ didx,
                    const T* __restrict__ G, const T* __restrict__ c, int32_t zrow, T* __restrict__ W, int64_t ntiles,
                    int npt, const uint64_t* __restrict__ qm, const uint64_t* __restrict__ smask) {
  typedef Mfma<T> M;
  typedef typename M::v4 v4;
  static_assert(KC == 8 || KC == 16, "8 or 16 columns");
  static_assert(QC == 16 || QC == 32, "chunks of 16 or 32 pairs");
  constexpr int RT = kMT / KC;           // 8 or 16 block rows per tile
  constexpr int RW = RT / 4;             // rows per wavefront
  constexpr int NI = PT / 16;            // p-subtiles per wavefront
  constexpr int SL = 16 / sizeof(T);     // values per 16-byte piece
  constexpr int SPE = KC / SL;           // pieces per D entry
  constexpr int EPI = 64 / SPE;          // D entries per copy instruction
  constexpr int DI = RW * QC / EPI;      // D copies per wavefront and chunk
  constexpr int GSR = PT / SL;           // pieces per G row
  constexpr int GI = QC * GSR / 64;      // G copies per chunk
  constexpr int GW = (GI + 3) / 4;       // G copies per wavefront (at most)
  constexpr bool SWZ = sizeof(T) == 4 && (PT / 16) % 2 == 0;
  static_assert(QC * GSR % 64 == 0 && RW * QC % EPI == 0, "copy shapes");
  __shared__ __align__(16) T Gs[2][QC * PT];
  __shared__ __align__(16) T Ds[2][RT * QC * KC];
  const int64_t per = (ntiles + 7) / 8;
  const int64_t tile = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;  // uniform per workgroup
  const int64_t mt = tile / npt;
  const int pt = (int)(tile % npt);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const RasBlockDev& B = blk[bi];
  const int64_t nrows = B.nrows;
  // chunks of QC pairs; with QC = 32 the last one may hold 16 (nqp is a
  // multiple of 16): its other half reads the zero row and G row nqp - 1
  const int nqp = B.nqp, nch = (nqp + QC - 1) / QC;

  // this lane's sources as offsets advanced by one chunk per step: its D
  // entries in the index table (rows padded to whole tiles: no bounds
  // checks) and its G pieces (rows q of G's symmetric tile; a piece past nqp
  // repeats its row's first, feeding only output rows that are not stored)
  const int32_t* __restrict__ dx = didx + B.dbase;
  const T* __restrict__ Gb = G + B.goff;
  int dxo[DI];
#pragma unroll
  for (int d = 0; d < DI; ++d) {
    const int en = d * EPI + lane / SPE;  // entry among the wavefront's RW x QC
    dxo[d] = (int)((mt * RT + RW * wave + en / QC) * nqp + en % QC);
  }
  int go[GW];
#pragma unroll
  for (int gi = 0; gi < GW; ++gi) {
    const int sl = 64 * (wave + 4 * gi) + lane, k = sl / GSR;
    const int r = SWZ ? (sl % GSR) ^ (4 * (k & 1)) : sl % GSR;
    const int p = pt * PT + r * SL;
    go[gi] = k * nqp + (p < nqp ? p : 0);
  }
  const int v0 = SL * (lane % SPE);
  // LDS addresses of this wavefront's first D copy and first G copy (stage 0)
  const uint32_t dl0 = lds_addr(&Ds[0][RW * wave * QC * KC]);
  const uint32_t gl0 = lds_addr(&Gs[0][64 * wave * SL]);
  auto fetch_idx = [&](int ch, int32_t (&ix)[DI]) {
#pragma unroll
    for (int d = 0; d < DI; ++d) {
      if (QC == 16 || ch * QC + (d * EPI + lane / SPE) % QC < nqp) ix[d] = dx[dxo[d] + ch * QC];
      else ix[d] = zrow;
    }
  };
  // the copies of chunk ch into stage st: D pieces first (they use the
  // indices loaded a chunk earlier, which the compiler waits for), then G
  auto issue = [&](int ch, int st, const int32_t (&ix)[DI]) {
    // every source address is formed (its index load waited for) before
    // the first copy issues: a wait for an index load placed between two
    // copies would also wait for the copy before it
    const T* src[DI];
#pragma unroll
    for (int d = 0; d < DI; ++d) src[d] = c + (int64_t)ix[d] * KC + v0;
#pragma unroll
    for (int d = 0; d < DI; ++d) asm volatile("" : "+v"(src[d]));
#pragma unroll
    for (int d = 0; d < DI; ++d)
      lds_copy16(src[d], dl0 + (uint32_t)((st * RT * QC * KC + d * EPI * KC) * sizeof(T)));
    const int q0 = ch * QC;
#pragma unroll
    for (int gi = 0; gi < GW; ++gi) {
      if (wave + 4 * gi >= GI) continue;
      const int k = (64 * (wave + 4 * gi) + lane) / GSR;  // row of the chunk
      const int qo = (QC == 16 || q0 + k < nqp) ? q0 * nqp : (nqp - 1 - k) * nqp;
      lds_copy16(Gb + (go[gi] + qo), gl0 + (uint32_t)((st * QC * PT + 64 * 4 * gi * SL) * sizeof(T)));
    }
  };

  // lane offsets of the MFMA operand reads (k row kr = lane / 16 within a
  // step): p-subtile i of G sits at a0[i & 1] + 16 i (with the swap of
  // 64-byte halves on odd rows: 16 (i ^ 1) = 16 i -+ 16); the B operand's
  // column lane & 15 is value vl of row RW wave + rl (+ 16 j / KC for
  // subtile j)
  const int kr = lane >> 4;
  const int sw = SWZ ? 16 * (kr & 1) : 0;
  const int a0[2] = {kr * PT + (lane & 15) + sw, kr * PT + (lane & 15) - sw};
  const int rl = (lane & 15) / KC, vl = (lane & 15) % KC;
  const int boff = ((RW * wave + rl) * QC + kr) * KC + vl;
  // pairs past nq are padding (zero G rows and columns, zero D entries): a
  // tile's last p-subtiles that hold only padding are skipped
  const int nq = B.nq;
  const int niv = ((nq - pt * PT < PT ? nq - pt * PT : PT) + 15) / 16;
  // chunks with only zero D entries are skipped: when every row of the tile
  // has the same first string u (alpha-beta and alpha-alpha products, rows
  // u-major), a chunk whose pairs all add an orbital u holds (bit of qm not
  // in u's mask) contributes nothing.  live[0..1]: bits of the first 128
  // chunks (later ones are always live); uniform over the workgroup.
  uint64_t live[2] = {~0ull, ~0ull};
  if (B.kind != 2) {
    const int64_t r0 = mt * RT, r1 = (r0 + RT < nrows ? r0 + RT : nrows) - 1;
    int64_t ua, ub, wdum;
    if (B.tri) {
      ras_tri_uw(r0, B.nu, ua, wdum);
      ras_tri_uw(r1, B.nu, ub, wdum);
    } else {
      ua = r0 / B.nw;
      ub = r1 / B.nw;
    }
    if (ua == ub) {
      const uint64_t um = smask[B.umoff + B.u0 + ua];
      constexpr int H = QC / 16;  // 16-pair pieces of a chunk (qm is per piece)
      for (int ch = 0; ch < nch && ch < 128; ++ch) {
        bool dead = true;
        for (int h = 0; h < H; ++h)
          if (ch * H + h < nqp / 16 && (qm[B.qmoff + ch * H + h] & ~um) != 0) dead = false;
        if (dead) live[ch >> 6] &= ~(1ull << (ch & 63));
      }
    }
  }
  // first live chunk after chunk x (nch when none): a find-first-set on
  // the masks (bits past nch are set)
  auto next_live = [&](int x) {
    int y = x + 1;
    if (y < 64) {
      const uint64_t m = live[0] >> y;
      if (m) return min(y + (int)__builtin_ctzll(m), nch);
      y = 64;
    }
    if (y < 128) {
      const uint64_t m = live[1] >> (y - 64);
      if (m) return min(y + (int)__builtin_ctzll(m), nch);
      y = 128;
    }
    return min(y, nch);
  };

  // the whole tile with NV live p-subtiles
  auto run = [&](auto nvc) {
    constexpr int NV = decltype(nvc)::value;
    v4 acc[NV][2];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = v4{0, 0, 0, 0};
    // one k-step of 4 pairs on stage st
    auto step = [&](int st, int kk) {
      const T* __restrict__ gs = Gs[st];
      const T* __restrict__ ds = Ds[st];
      T a[NV], b[2];
#pragma unroll
      for (int i = 0; i < NV; ++i) a[i] = gs[a0[i & 1] + 16 * i + 4 * kk * PT];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = ds[boff + (16 / KC * j * QC + 4 * kk) * KC];
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = M::op(a[i], b[j], acc[i][j]);
    };
    // the live chunks in order: c0 (in the stage being read), c1 (copies in
    // flight), c2 (indices in flight)
    int32_t ixn[DI];
    int c0 = next_live(-1);
    int c1 = c0 < nch ? next_live(c0) : nch;
    int c2 = c1 < nch ? next_live(c1) : nch;
    if (c0 < nch) {
      fetch_idx(c0, ixn);
      issue(c0, 0, ixn);
      if (c1 < nch) fetch_idx(c1, ixn);
    }
    for (int it = 0; c0 < nch; ++it) {
      const int st = it & 1;
      // this wavefront's copies of chunk c0 have landed, every wavefront's
      // have after the barrier, and stage st ^ 1 is free (its reads of the
      // chunk before ended before the barrier)
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (c1 < nch) {
        issue(c1, st ^ 1, ixn);
        if (c2 < nch) fetch_idx(c2, ixn);
      }
#pragma unroll
      for (int kk = 0; kk < QC / 4; ++kk) step(st, kk);
      c0 = c1;
      c1 = c2;
      c2 = c2 < nch ? next_live(c2) : nch;
    }
    const int64_t wb = B.wbase;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t row = mt * RT + RW * wave + 16 / KC * j + rl;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int p = pt * PT + 16 * i + M::row(t, lane);
          if (row < nrows && p < nqp) W[(wb + row * nqp + p) * KC + vl] = acc[i][j][t];
        }
      }
  };
  // tiles cover the pairs exactly except, for some widths, the last one
  if (niv >= NI) {
    run(IntC<NI>{});
  } else if constexpr (NI > 1) {
    if (niv == 1) run(IntC<1>{});
    else if (NI > 2 && niv == 2) run(IntC<(NI > 2 ? 2 : 1)>{});
    else if (NI > 3 && niv == 3) run(IntC<(NI > 3 ? 3 : 1)>{});
    else run(IntC<NI>{});
  }
}

// The term list of every SEF (built once): FILL = false counts each SEF's
// terms into cnt[s]; FILL = true writes them, as the W offset per column
// with bit 31 set for a negative sign (the determinants' coefficient signs
// folded in), at toff[s] (the scanned counts), and the SEF's scale code.
template <bool FILL>
__global__ __launch_bounds__(256) void k_ras_terms(const RasView* __restrict__ vp, int64_t* __restrict__ cnt,
                                                   const int64_t* __restrict__ toff, int32_t* __restrict__ term,
                                                   int8_t* __restrict__ scl, int* __restrict__ bad) {
  const RasView& v = *vp;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= v.nsef) return;
  const RasSef e = v.sef[s];
  if (!FILL) {
    int64_t n = 0;
    ras_sef_terms(v, e, bad, [&](int64_t, int) { ++n; });
    cnt[s] = n;
    return;
  }
  int64_t o = toff[s];
  const int64_t end = toff[s + 1];
  const int code = ras_sef_terms(v, e, bad, [&](int64_t off, int neg) {
    if (o >= end) {
      *bad = 1;
      return;
    }
    term[o++] = (int32_t)((uint32_t)off | ((uint32_t)neg << 31));
  });
  scl[s] = (int8_t)code;
  if (o != end) *bad = 1;
}

// y[sef][col0 + q] = scale * sum of +-W[off][q] over the SEF's terms (Y may
// be wider than T: the FP32 result widened exactly)
template <typename T, int KC, typename TY = T>
__global__ __launch_bounds__(256) void k_ras_y(const int64_t* __restrict__ toff, const int32_t* __restrict__ term,
                                               const int8_t* __restrict__ scl, const T* __restrict__ W, int64_t nsef,
                                               TY* __restrict__ Y, int ldy, int col0, int kw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t s = t / KC;
  const int q = (int)(t % KC);
  if (s >= nsef) return;
  const int64_t b = toff[s], e = toff[s + 1];
  T a0 = (T)0, a1 = (T)0;
  int64_t i = b;
  for (; i + 4 <= e; i += 4) {
    const int32_t t0 = term[i], t1 = term[i + 1], t2 = term[i + 2], t3 = term[i + 3];
    const T w0 = W[(int64_t)(t0 & 0x7fffffff) * KC + q];
    const T w1 = W[(int64_t)(t1 & 0x7fffffff) * KC + q];
    const T w2 = W[(int64_t)(t2 & 0x7fffffff) * KC + q];
    const T w3 = W[(int64_t)(t3 & 0x7fffffff) * KC + q];
    a0 += t0 < 0 ? -w0 : w0;
    a1 += t1 < 0 ? -w1 : w1;
    a0 += t2 < 0 ? -w2 : w2;
    a1 += t3 < 0 ? -w3 : w3;
  }
  for (; i < e; ++i) {
    const int32_t t0 = term[i];
    const T w0 = W[(int64_t)(t0 & 0x7fffffff) * KC + q];
    a0 += t0 < 0 ? -w0 : w0;
  }
  const int8_t sc = scl[s];
  const T f = sc == 2 ? (T)1 : (sc == 1 ? (T)0x1.6a09e667f3bcdp+0 : (T)0x1.6a09e667f3bcdp-1);
  if (q < kw) Y[s * (int64_t)ldy + col0 + q] = (TY)(f * (a0 + a1));
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

// Optional per-launch events of a timed sigma: group 0 the SEF ->
// determinant scatter, 1..3 the products of each kind, 4 the SEF gather.
struct RasTiming {
  std::vector<hipEvent_t> ev;
  std::vector<int> group;  // group of the launch that ends at ev[i + 1]
  void mark(hipStream_t s, int g) {
    hipEvent_t e;
    PSOCI_HIP(hipEventCreate(&e));
    PSOCI_HIP(hipEventRecord(e, s));
    ev.push_back(e);
    if (g >= 0) group.push_back(g);
  }
  ~RasTiming() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// p-pairs per product tile: one tile when there are at most 112, else the
// first of 64, 96, 80, 112 and 48 that divides the (16-padded) pair count,
// else 64 with a narrower last tile
inline int ras_tile_width(int nqp) {
  if (nqp <= 112) return nqp;
  for (int w : {64, 96, 80, 112, 48})
    if (nqp % w == 0) return w;
  return 64;
}

template <typename T, int KC, typename TIO = T>
void ras_chunk(Ctx& c, const TIO* X, int ldx, TIO* Y, int ldy, int col0, int kw, RasTiming* tm = nullptr) {
  const RasPlan& P = c.ras;
  T* cv = reinterpret_cast<T*>(c.d_rasc.p);
  T* W = reinterpret_cast<T*>(c.d_rasw.p);
  const T* G = sizeof(T) == 4 ? reinterpret_cast<const T*>(c.d_rasg32.p) : reinterpret_cast<const T*>(c.d_rasg64.p);
  const int64_t nsef = (int64_t)P.sef.size();
  if (tm && tm->ev.empty()) tm->mark(c.stream, -1);
  hipLaunchKernelGGL((k_ras_x2c<T, KC, TIO>), dim3(nblk(nsef * KC, 256)), dim3(256), 0, c.stream, nsef, P.ndet,
                     c.d_rasrows.p, X, ldx, col0, kw, cv);
  PSOCI_HIP(hipGetLastError());
  if (tm) tm->mark(c.stream, 0);
  for (int bi = 0; bi < (int)P.blk.size(); ++bi) {
    const RasBlockDev& B = P.blk[bi];
    const int pt_w = ras_tile_width(B.nqp);
    const int npt = (B.nqp + pt_w - 1) / pt_w;
    const int64_t nmt = (B.nrows * KC + kMT - 1) / kMT;
    const int64_t ntiles = nmt * npt;
    const unsigned grid = (unsigned)(8 * ((ntiles + 7) / 8));
    const RasBlockDev* blk = c.d_rasblk.p;
    const int32_t* didx = c.d_rasdidx.p;
    const int32_t zrow = (int32_t)(2 * P.ndet);  // the zero row of c
    const uint64_t* qm = c.d_rasqm.p;
    const uint64_t* smask = c.d_rasmask.p;
    if constexpr ((KC == 16 || KC == 8) && PSOCI_RAS_DMA) {
      constexpr int LQC = sizeof(T) == 4 && PSOCI_RAS_LQC32 ? 32 : 16;
      switch (pt_w) {
#define PSOCI_RAS_GEMML(PTV)                                                                                    \
  case PTV:                                                                                                    \
    hipLaunchKernelGGL((k_ras_gemm_lds<T, KC, PTV, LQC>), dim3(grid), dim3(256), 0, c.stream, blk, bi, didx, G, cv, \
                       zrow, W, ntiles, npt, qm, smask);                                                       \
    break;
        PSOCI_RAS_GEMML(16)
        PSOCI_RAS_GEMML(32)
        PSOCI_RAS_GEMML(48)
        PSOCI_RAS_GEMML(80)
        PSOCI_RAS_GEMML(96)
        PSOCI_RAS_GEMML(112)
#undef PSOCI_RAS_GEMML
        default:
          hipLaunchKernelGGL((k_ras_gemm_lds<T, KC, 64, LQC>), dim3(grid), dim3(256), 0, c.stream, blk, bi, didx, G,
                             cv, zrow, W, ntiles, npt, qm, smask);
      }
    } else
    switch (pt_w) {
#define PSOCI_RAS_GEMM(PTV)                                                                                  \
  case PTV:                                                                                                 \
    hipLaunchKernelGGL((k_ras_gemm<T, KC, PTV>), dim3(grid), dim3(256), 0, c.stream, blk, bi, didx, G, cv, zrow, \
                       W, ntiles, npt);                                                                     \
    break;
      PSOCI_RAS_GEMM(16)
      PSOCI_RAS_GEMM(32)
      PSOCI_RAS_GEMM(48)
      PSOCI_RAS_GEMM(80)
      PSOCI_RAS_GEMM(96)
      PSOCI_RAS_GEMM(112)
#undef PSOCI_RAS_GEMM
      default:
        hipLaunchKernelGGL((k_ras_gemm<T, KC, 64>), dim3(grid), dim3(256), 0, c.stream, blk, bi, didx, G, cv, zrow,
                           W, ntiles, npt);
    }
    PSOCI_HIP(hipGetLastError());
    if (tm) tm->mark(c.stream, 1 + B.kind);
  }
  hipLaunchKernelGGL((k_ras_y<T, KC, TIO>), dim3(nblk(nsef * KC, 256)), dim3(256), 0, c.stream, c.d_rastoff.p,
                     c.d_rasterm.p, c.d_rasscl.p, W, nsef, Y, ldy, col0, kw);
  PSOCI_HIP(hipGetLastError());
  if (tm) tm->mark(c.stream, 4);
}

// T: the computation's type; TIO: the vectors' (FP64 vectors with an FP32
// computation: rounded on the way in, widened on the way out)
template <typename T, typename TIO = T>
void ras_any_k(Ctx& c, int k, const TIO* X, int ldx, TIO* Y, int ldy, RasTiming* tm = nullptr) {
  const int kmax = sizeof(T) == 4 ? 16 : c.ras_kc64;
  for (int c0 = 0; c0 < k; c0 += kmax) {
    const int kw = std::min(kmax, k - c0);
    if (kw > 8)
      ras_chunk<T, 16, TIO>(c, X, ldx, Y, ldy, c0, kw, tm);
    else if (kw > 4) ras_chunk<T, 8, TIO>(c, X, ldx, Y, ldy, c0, kw, tm);
    else if (kw > 2) ras_chunk<T, 4, TIO>(c, X, ldx, Y, ldy, c0, kw, tm);
    else if (kw > 1) ras_chunk<T, 2, TIO>(c, X, ldx, Y, ldy, c0, kw, tm);
    else ras_chunk<T, 1, TIO>(c, X, ldx, Y, ldy, c0, kw, tm);
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
    // device memory: W and the determinant vectors (16 FP32 / 8 or 16 FP64
    // columns), the term lists; leave room for the caller's vectors
    int64_t terms = 0;
    for (const RasSef& e : P.sef)
      for (int d = 0; d < e.n; ++d) {
        const RasSectorDev& S = P.sec[e.sec[d]];
        terms += (int64_t)S.na * S.nb + (int64_t)S.na * (S.na - 1) / 2 + (int64_t)S.nb * (S.nb - 1) / 2;
      }
    const double need = (double)P.wunit * 64 + (double)P.dunit * 4 + (double)P.ndet * 128 + (double)terms * 4 +
                        (double)P.sef.size() * 9;
    size_t fr = 0, tot = 0;
    PSOCI_HIP(hipMemGetInfo(&fr, &tot));
    // FP64 in chunks of 16 columns (whole 128-byte W and c pieces) when W and
    // c fit twice over in 60 % of the free memory, else chunks of 8
    c.ras_kc64 = need + (double)P.wunit * 64 + (double)P.ndet * 128 <= 0.6 * (double)fr ? 16 : 8;
    if (need > 0.6 * (double)fr) {
      c.ras_why = "needs " + std::to_string((int64_t)(need / 1e9)) + " GB of device memory, " +
                  std::to_string((int64_t)(fr / 1000000000)) + " GB free";
      c.ras = RasPlan();
      return false;
    }
  }
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
  c.d_rasqm.upload(P.qm, c.stream);
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
  // determinant vectors and W: 16 FP32 or ras_kc64 FP64 columns per chunk
  c.d_rasc.alloc((size_t)(2 * P.ndet + 2) * (c.ras_kc64 == 16 ? 16 * sizeof(double) : 16 * sizeof(float)));
  // the D entries' signed rows
  c.d_rasdidx.alloc(P.dunit);
  for (int bi = 0; bi < (int)P.blk.size(); ++bi) {
    const int64_t n = (P.blk[bi].nrows + 15) / 16 * 16 * P.blk[bi].nqp;
    hipLaunchKernelGGL(k_ras_didx, dim3(nblk(n, 256)), dim3(256), 0, c.stream, c.d_rasview.p, bi, c.d_rasdidx.p);
    PSOCI_HIP(hipGetLastError());
  }
  // the SEF term lists (merged where an SEF's two determinants read the
  // same W elements): count, scan, fill
  {
    const int64_t nsef = (int64_t)P.sef.size();
    DevBuf<int64_t> cnt;
    cnt.alloc(nsef + 1);
    DevBuf<int> bad;
    bad.alloc(1);
    PSOCI_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), c.stream));
    PSOCI_HIP(hipMemsetAsync(cnt.p + nsef, 0, sizeof(int64_t), c.stream));
    hipLaunchKernelGGL(k_ras_terms<false>, dim3(nblk(nsef, 256)), dim3(256), 0, c.stream, c.d_rasview.p, cnt.p,
                       nullptr, nullptr, nullptr, bad.p);
    PSOCI_HIP(hipGetLastError());
    c.d_rastoff.alloc(nsef + 1);
    c.scan(cnt.p, c.d_rastoff.p, nsef + 1);
    int64_t nterm = 0;
    PSOCI_HIP(hipMemcpyAsync(&nterm, c.d_rastoff.p + nsef, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    c.d_rasterm.alloc(nterm);
    c.d_rasscl.alloc(nsef);
    c.d_rasrows.alloc(nsef);
    hipLaunchKernelGGL(k_ras_sefrows, dim3(nblk(nsef, 256)), dim3(256), 0, c.stream, c.d_rasview.p, c.d_rasrows.p);
    PSOCI_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ras_terms<true>, dim3(nblk(nsef, 256)), dim3(256), 0, c.stream, c.d_rasview.p, nullptr,
                       c.d_rastoff.p, c.d_rasterm.p, c.d_rasscl.p, bad.p);
    PSOCI_HIP(hipGetLastError());
    int b = 0;
    bad.download(&b, 1, c.stream);
    c.sync();
    PSOCI_CHECK(!b, PSOCI_ERR_STATE, "string-driven sigma: inconsistent term lists");
    c.ras_nterm = nterm;
  }
  c.d_rasw.alloc((size_t)P.wunit * (c.ras_kc64 == 16 ? 16 * sizeof(double) : 16 * sizeof(float)));
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

void ras_timing(Ctx& c, int prec, int k, int reps, double* ms, double* flops) {
  PSOCI_CHECK(c.have_ras, PSOCI_ERR_STATE, "the string-driven direct sigma is not set up");
  PSOCI_CHECK(k >= 1 && reps >= 1, PSOCI_ERR_ARG, "k and reps must be positive");
  const int64_t n = c.sp.nsef_total;
  const size_t es = prec == PSOCI_FP32 ? 4 : 8;
  DevBuf<char> X, Y;
  X.alloc((size_t)n * k * es);
  Y.alloc((size_t)n * k * es);
  PSOCI_HIP(hipMemsetAsync(X.p, 0, (size_t)n * k * es, c.stream));
  double acc[5] = {0, 0, 0, 0, 0};
  for (int r = 0; r < reps; ++r) {
    RasTiming tm;
    if (prec == PSOCI_FP32)
      ras_any_k<float>(c, k, reinterpret_cast<const float*>(X.p), k, reinterpret_cast<float*>(Y.p), k, &tm);
    else
      ras_any_k<double>(c, k, reinterpret_cast<const double*>(X.p), k, reinterpret_cast<double*>(Y.p), k, &tm);
    PSOCI_HIP(hipEventSynchronize(tm.ev.back()));
    for (size_t i = 0; i < tm.group.size(); ++i) {
      float t = 0.f;
      PSOCI_HIP(hipEventElapsedTime(&t, tm.ev[i], tm.ev[i + 1]));
      acc[tm.group[i]] += t;
    }
  }
  for (int g = 0; g < 5; ++g) ms[g] = acc[g] / reps;
  if (flops) {
    for (int g = 0; g < 3; ++g) flops[g] = 0.0;
    for (const RasBlockDev& B : c.ras.blk) flops[B.kind] += 2.0 * (double)B.nrows * B.nq * B.nq * k;
  }
}

void sigma_ras(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy) {
  PSOCI_CHECK(c.have_ras, PSOCI_ERR_STATE, "the string-driven direct sigma is not set up");
  if (prec == PSOCI_FP32)
    ras_any_k<float>(c, k, static_cast<const float*>(X), ldx, static_cast<float*>(Y), ldy);
  else
    ras_any_k<double>(c, k, static_cast<const double*>(X), ldx, static_cast<double*>(Y), ldy);
}

}  // namespace psoci
