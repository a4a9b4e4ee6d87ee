// String-driven direct sigma for complete RAS spaces (host plan and the
// index arithmetic shared by the GPU kernels and the host check).
//
// A configuration list that holds every occupation of N electrons with at
// most m electrons in a set E of "external" orbitals (and any occupation of
// the others) -- a full CI (E empty) or an MR-CISD list such as the
// reference's benchmark spaces -- has a determinant space that factors into
// alpha and beta strings: within each spin sector (n_alpha, n_beta), every
// (alpha string, beta string) pair with ext(alpha) + ext(beta) <= m is a
// determinant of exactly one configuration.  The reference's determinants
// are (-1)^phase times the canonical product (alpha orbitals ascending, then
// beta ascending), its SEFs the Kramers combinations (D_2k + mult
// D_2k+1)/sqrt(2) (PsociHamiltonian.C:1304-1322, 1350-1454), so the SEF
// sigma is C^T H_det C x with H_det the spin-free Hamiltonian over canonical
// determinants.
//
// H_det x is evaluated by resolutions of the identity over strings with
// fewer electrons, which turns element generation into dense products:
//   alpha-beta: sum_pqrs (pq|rs) E^a_pq E^b_rs through (N-1, N-1) string
//     pairs (ka, kb): D[(ka,kb)][(q,s)] = eps eps c(ka+q, kb+s),
//     W = D G^T with G[(p,r),(q,s)] = (pq|rs) (+ the one-electron integrals
//     folded in as h_pq delta_rs / n_b + h_rs delta_pq / n_a),
//     sigma(ka+p, kb+r) += eps eps W[(ka,kb)][(p,r)];
//   same spin: 1/2 sum (pq|rs) a+_p a+_r a_s a_q through (N-2) strings k2 of
//     that spin and the other spin's N strings, pairs p > r, q > s,
//     V[(p,r),(q,s)] = (pq|rs) - (ps|rq) (one-electron integrals folded in
//     here when the other spin has no electrons).
// The pair sets are restricted per class of the row strings (their external
// electron counts) to the pairs that keep ext <= m, so every product is a
// dense GEMM over the class pair.  Only sectors with n_alpha >= n_beta are
// evaluated: a sector with n_alpha > n_beta is the spin flip of its mirror,
// whose sigma is mult times its own.
//
// The reference keeps no such structure (it regenerates every block with
// generateSOCIblocks, PsociHamiltonian.C:401-558); the plan is used only
// where it is exact: EVEN-class spaces of one jsym with an even electron
// count, and no configuration pair whose single-excitation sum vnt[0] is
// exactly zero while its exchange terms are not (the reference skips such a
// block, processSingleExcitation PsociHamiltonian.C:190-261; checked on the
// GPU at setup).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "elem.hpp"

namespace psoci {

constexpr int kRasMaxCls = 7;    // external-electron classes 0..6
constexpr int kRasMaxN = 32;     // electrons
constexpr uint32_t kRasNone = 0xffffffffu;

// add-table entry: sign << 31 | class << 28 | rank within the class
PSOCI_HD uint32_t ras_pack(int sign_neg, int cls, int64_t rank) {
  return ((uint32_t)sign_neg << 31) | ((uint32_t)cls << 28) | (uint32_t)rank;
}
PSOCI_HD int ras_cls(uint32_t e) { return (int)((e >> 28) & 7u); }
PSOCI_HD int64_t ras_rank(uint32_t e) { return (int64_t)(e & 0x0fffffffu); }
PSOCI_HD int ras_neg(uint32_t e) { return (int)(e >> 31); }

struct RasLevelDev {
  int64_t mask_off;                 // first string of the level in mask[] / cls[]
  int64_t add_off;                  // add[]: [count][nbf]
  int64_t rem_off;                  // rem[]: [count][n]
  int64_t cls_off[kRasMaxCls + 1];  // global index (within the level) of each class's first string
};

struct RasSectorDev {
  int na, nb;
  // 1: an Ms = 0 sector whose vectors are symmetric under the spin flip
  // (mult = +1): only the alpha-beta products with class(u) <= class(w) and
  // no beta-beta products are evaluated; the others are their mirrors
  int sym, pad;
  int64_t det0;                          // first determinant of the sector
  int64_t boff[kRasMaxCls][kRasMaxCls];  // block (alpha class, beta class) offset in the sector, -1 none
};

// One dense product: rows (u, w), u of level lu and class cu, w of level lw
// and class cw (global indices u0 + row / nw, w0 + row % nw); pair list of
// shape (0: ordered (q, s); 1: q > s) and budget (external electrons the
// pairs may add), nq pairs padded to nqp.
struct RasBlockDev {
  int kind;  // 0 alpha-beta, 1 alpha-alpha, 2 beta-beta
  int sector;
  int lu, lw, cu, cw;
  int budget, nq, nqp;
  // 1: an alpha-beta product of a symmetric sector with cu == cw, whose rows
  // (u, w) with u > w mirror (w, u): only u <= w are evaluated, in the
  // triangular order ras_tri_row gives (nrows = nu (nu + 1) / 2)
  int tri;
  int64_t u0, nu, w0, nw, nrows;
  int64_t qoff;   // first pair in qlist[]
  int64_t goff;   // first element of G (nqp x nqp, row-major [p-pair][q-pair])
  int64_t wbase;  // W offset per column: W element (wbase + row * nqp + pidx) * KC + v
  int64_t dbase;  // first entry in the D index table (rows padded to a multiple of 16)
  // per chunk of 16 pairs: the bits of the orbitals q the chunk adds first
  // (qm[qmoff + chunk]); a chunk all of whose q the row string u already
  // holds has only zero D entries (kinds 0 and 1, where u is the row's
  // first string: mask[umoff + u])
  int64_t qmoff, umoff;
};

// The determinants of one SEF (one or two): alpha / beta global string
// indices, sector, and flags (bit 0 negative coefficient, bit 1 closed
// shell: coefficient +-1 instead of +-1/sqrt(2)); n = 1 or 2.
struct RasSef {
  int32_t a[2], b[2];
  int8_t sec[2];
  uint8_t flag[2];
  int32_t n;
};

struct RasView {
  int nbf, m, nsec, nblk;
  const uint64_t* mask;
  const int8_t* cls;
  const uint32_t* add;
  const int32_t* rem;
  RasLevelDev lev[kRasMaxN + 1];
  RasSectorDev sec[kRasMaxN / 4 + 1];
  const RasBlockDev* blk;
  const int16_t* qlist;   // q | s << 8, -1 padding
  const int16_t* pidx;    // [shape][budget][nbf][nbf] position in the pair list, -1 none
  int32_t blockid[kRasMaxN / 4 + 1][3][kRasMaxCls][kRasMaxCls];  // -1 none
  const RasSef* sef;
  int64_t nsef;
  int64_t ndet, wunit;  // determinants and W elements per column (bounds of every index)
};

// triangular row of (u, w), u <= w < n (local indices), and its inverse
PSOCI_HD int64_t ras_tri_row(int64_t u, int64_t w, int64_t n) { return u * n - u * (u - 1) / 2 + (w - u); }
PSOCI_HD void ras_tri_uw(int64_t row, int64_t n, int64_t& u, int64_t& w) {
  const double b = 2.0 * (double)n + 1.0;
  int64_t x = (int64_t)((b - sqrt(b * b - 8.0 * (double)row)) * 0.5);
  if (x < 0) x = 0;
  if (x > n - 1) x = n - 1;
  while (x > 0 && ras_tri_row(x, x, n) > row) --x;
  while (x < n - 1 && ras_tri_row(x + 1, x + 1, n) <= row) ++x;
  u = x;
  w = x + (row - ras_tri_row(x, x, n));
}

PSOCI_HD int64_t ras_det(const RasView& v, int s, int ca, int64_t ra, int cb, int64_t rb) {
  const RasSectorDev& S = v.sec[s];
  const int64_t nbw = v.lev[S.nb].cls_off[cb + 1] - v.lev[S.nb].cls_off[cb];
  return S.det0 + S.boff[ca][cb] + ra * nbw + rb;
}

// class of global string index g of level n
PSOCI_HD int ras_cls_of(const RasView& v, int n, int64_t g) { return v.cls[v.lev[n].mask_off + g]; }

// the determinant (and sign) of D-entry (row, pair j) of block B: -1 when
// the pair cannot be added to the row strings
PSOCI_HD int64_t ras_d_entry(const RasView& v, const RasBlockDev& B, int64_t row, int j, int& neg) {
  const int16_t qs = v.qlist[B.qoff + j];
  if (qs < 0) return -1;
  const int q = qs & 0xff, s = (qs >> 8) & 0xff;
  int64_t u, w;
  if (B.tri) {
    ras_tri_uw(row, B.nu, u, w);
    u += B.u0;
    w += B.w0;
  } else {
    u = B.u0 + row / B.nw;
    w = B.w0 + row % B.nw;
  }
  const RasSectorDev& S = v.sec[B.sector];
  if (B.kind == 0) {
    const uint32_t ea = v.add[v.lev[B.lu].add_off + u * v.nbf + q];
    const uint32_t eb = v.add[v.lev[B.lw].add_off + w * v.nbf + s];
    if (ea == kRasNone || eb == kRasNone) return -1;
    neg = ras_neg(ea) ^ ras_neg(eb);
    return ras_det(v, B.sector, ras_cls(ea), ras_rank(ea), ras_cls(eb), ras_rank(eb));
  }
  // same spin: add s, then q (q > s), to the N-2 string; the other spin's
  // string is the row's own
  const int l2 = B.kind == 1 ? B.lu : B.lw;
  const int64_t k2 = B.kind == 1 ? u : w;
  const uint32_t e1 = v.add[v.lev[l2].add_off + k2 * v.nbf + s];
  if (e1 == kRasNone) return -1;
  const int64_t g1 = v.lev[l2 + 1].cls_off[ras_cls(e1)] + ras_rank(e1);
  const uint32_t e2 = v.add[v.lev[l2 + 1].add_off + g1 * v.nbf + q];
  if (e2 == kRasNone) return -1;
  neg = ras_neg(e1) ^ ras_neg(e2);
  if (B.kind == 1) {
    const int cb = ras_cls_of(v, S.nb, w);
    return ras_det(v, B.sector, ras_cls(e2), ras_rank(e2), cb, w - v.lev[S.nb].cls_off[cb]);
  }
  const int ca = ras_cls_of(v, S.na, u);
  return ras_det(v, B.sector, ca, u - v.lev[S.na].cls_off[ca], ras_cls(e2), ras_rank(e2));
}

// Calls f(block index, W element offset per column (wbase + row * nqp +
// pidx), negative sign) for every term of sigma at determinant (alpha
// string a, beta string b) of sector s: alpha-beta removals (p in a, r in
// b), then alpha-alpha and beta-beta pairs (p > r).  A term without a block
// or pair position (never, for a determinant of the plan) is passed as
// f(-1, -1, 0).
//
// Spin-flip symmetry (S.sym): with the determinant vector symmetric under
// exchanging the alpha and beta strings, W[(kb, ka)][(r, p)] of the
// alpha-beta product equals W[(ka, kb)][(p, r)], and the beta-beta W at
// row (a, k2) equals the alpha-alpha W at row (k2, a) (same pair); a term
// of a product that is not evaluated reads its mirror.
template <class F>
PSOCI_HD void ras_terms(const RasView& v, int s, int64_t a, int64_t b, F&& f) {
  const RasSectorDev& S = v.sec[s];
  const uint64_t am = v.mask[v.lev[S.na].mask_off + a];
  const uint64_t bm = v.mask[v.lev[S.nb].mask_off + b];
  const int ca = ras_cls_of(v, S.na, a), cb = ras_cls_of(v, S.nb, b);
  const int nbf = v.nbf;
  const int nb2 = nbf * nbf;
  if (S.na >= 1 && S.nb >= 1) {
    int tp = 0;
    for (uint64_t x = am; x; x &= x - 1, ++tp) {
      const int p = ctz(x);
      const int64_t u = v.rem[v.lev[S.na].rem_off + a * S.na + tp];
      const int cu = ras_cls_of(v, S.na - 1, u);
      int tr = 0;
      for (uint64_t y = bm; y; y &= y - 1, ++tr) {
        const int r = ctz(y);
        const int64_t w = v.rem[v.lev[S.nb].rem_off + b * S.nb + tr];
        const int cw = ras_cls_of(v, S.nb - 1, w);
        const bool mir = S.sym && cu > cw;
        const int bi = mir ? v.blockid[s][0][cw][cu] : v.blockid[s][0][cu][cw];
        if (bi < 0) {
          f(-1, (int64_t)-1, 0);
          continue;
        }
        const RasBlockDev& B = v.blk[bi];
        // local row strings and pair of the evaluated product
        int64_t ul = mir ? w - B.u0 : u - B.u0, wl = mir ? u - B.w0 : w - B.w0;
        int pp = mir ? r : p, rr = mir ? p : r;
        // a triangular product holds rows u <= w; on a diagonal row (u = w)
        // W[(u,u)][(p,r)] = W[(u,u)][(r,p)] (the flip), read as p <= r so
        // that a determinant and its flip read the same element
        if (B.tri && (ul > wl || (ul == wl && pp > rr))) {
          const int64_t t = ul;
          ul = wl;
          wl = t;
          const int tq = pp;
          pp = rr;
          rr = tq;
        }
        const int pi = v.pidx[(0 * (v.m + 1) + B.budget) * nb2 + pp * nbf + rr];
        if (pi < 0) {
          f(-1, (int64_t)-1, 0);
          continue;
        }
        const int64_t row = B.tri ? ras_tri_row(ul, wl, B.nu) : ul * B.nw + wl;
        f(bi, B.wbase + row * B.nqp + pi, (tp + tr) & 1);
      }
    }
  }
  for (int spin = 0; spin < 2; ++spin) {
    const int n = spin == 0 ? S.na : S.nb;
    if (n < 2) continue;
    const uint64_t mk = spin == 0 ? am : bm;
    const int64_t g = spin == 0 ? a : b;
    const int64_t o = spin == 0 ? b : a;
    const int co = spin == 0 ? cb : ca;
    int tp = 0;
    for (uint64_t x = mk; x; x &= x - 1, ++tp) {
      const int p = ctz(x);
      const int64_t u1 = v.rem[v.lev[n].rem_off + g * n + tp];
      int tr = 0;
      for (uint64_t y = mk & ((1ull << p) - 1ull); y; y &= y - 1, ++tr) {
        const int r = ctz(y);
        const int64_t k2 = v.rem[v.lev[n - 1].rem_off + u1 * (n - 1) + tr];
        const int ck = ras_cls_of(v, n - 2, k2);
        const bool mir = S.sym && spin == 1;  // beta-beta read from the alpha-alpha product
        const int bi = (spin == 0 || mir) ? v.blockid[s][1][ck][co] : v.blockid[s][2][co][ck];
        if (bi < 0) {
          f(-1, (int64_t)-1, 0);
          continue;
        }
        const RasBlockDev& B = v.blk[bi];
        const int pi = v.pidx[(1 * (v.m + 1) + B.budget) * nb2 + p * nbf + r];
        if (pi < 0) {
          f(-1, (int64_t)-1, 0);
          continue;
        }
        const int64_t row =
            (spin == 0 || mir) ? (k2 - B.u0) * B.nw + (o - B.w0) : (o - B.u0) * B.nw + (k2 - B.w0);
        f(bi, B.wbase + row * B.nqp + pi, (tp + tr) & 1);
      }
    }
  }
}

// Host plan: the tables above in host memory, the per-block G matrices
// (FP64), and the sizes the device buffers need.
// The W terms of SEF e: f(offset, negative) per term; returns the SEF's
// scale code (0 -> 1/sqrt(2), 1 -> sqrt(2), 2 -> 1).  The two determinants
// of an open-shell SEF of a spin-flip symmetric sector (D and its flip)
// read the same W elements -- every term of the flip is the mirror of one of
// D's, stored once -- so their terms are merged: equal signs give one term
// of twice the weight (scale 2/sqrt(2) = sqrt(2)), opposite signs cancel.
// Otherwise (a term without its mirror, more than kRasMergeMax terms, other
// sectors) the determinants' terms are listed one by one.  *bad = 1 for a
// term with no W element (never, for a plan that passes its host checks).
constexpr int kRasMergeMax = 64;
template <typename F>
PSOCI_HD int ras_sef_terms(const RasView& v, const RasSef& e, int* bad, F&& f) {
  if (e.n == 2 && e.sec[0] == e.sec[1] && v.sec[e.sec[0]].sym) {
    int32_t off[kRasMergeMax];
    int8_t sg[kRasMergeMax];
    int n0 = 0;
    bool ok = true;
    ras_terms(v, e.sec[0], e.a[0], e.b[0], [&](int bi, int64_t o, int neg) {
      if (bi < 0 || o < 0 || o >= v.wunit || n0 >= kRasMergeMax) {
        ok = false;
        return;
      }
      off[n0] = (int32_t)o;
      sg[n0++] = (int8_t)(neg ^ (e.flag[0] & 1));
    });
    if (ok) {
      uint64_t used = 0;
      int npair = 0;
      ras_terms(v, e.sec[1], e.a[1], e.b[1], [&](int bi, int64_t o, int neg) {
        if (!ok) return;
        int k = -1;
        for (int i = 0; i < n0; ++i)
          if (!((used >> i) & 1) && off[i] == o) {
            k = i;
            break;
          }
        if (bi < 0 || k < 0) {
          ok = false;
          return;
        }
        used |= 1ull << k;
        if (sg[k] != (int8_t)(neg ^ (e.flag[1] & 1))) sg[k] = 2;  // the pair cancels
        ++npair;
      });
      if (ok && npair == n0) {
        for (int i = 0; i < n0; ++i)
          if (sg[i] != 2) f((int64_t)off[i], (int)sg[i]);
        return 1;
      }
    }
  }
  for (int d = 0; d < e.n; ++d) {
    const int nd = e.flag[d] & 1;
    ras_terms(v, e.sec[d], e.a[d], e.b[d], [&](int bi, int64_t o, int neg) {
      if (bi < 0 || o < 0 || o >= v.wunit) {
        *bad = 1;
        return;
      }
      f(o, neg ^ nd);
    });
  }
  return (e.flag[0] & 2) ? 2 : (e.sec[0] > 0 ? 1 : 0);
}

struct RasPlan {
  bool ok = false;
  std::string why;  // when !ok: why the space does not qualify
  int nbf = 0, N = 0, m = 0;
  uint64_t emask = 0;
  double mult = 1.0;
  std::vector<uint64_t> mask;
  std::vector<int8_t> cls;
  std::vector<uint32_t> add;
  std::vector<int32_t> rem;
  std::vector<RasLevelDev> lev;  // kRasMaxN + 1
  std::vector<int> lev_cnt_n;    // electrons per built level (diagnostics)
  std::vector<RasSectorDev> sec;
  std::vector<RasBlockDev> blk;
  std::vector<int16_t> qlist, pidx;
  std::vector<uint64_t> qm;  // per block and chunk of 16 pairs (RasBlockDev::qmoff)
  std::vector<double> G;
  std::vector<int32_t> blockid;  // flattened [sector][kind][cu][cw]
  std::vector<RasSef> sef;
  int64_t ndet = 0;   // evaluated determinants (all sectors with n_alpha >= n_beta)
  int64_t wunit = 0;  // W elements per column
  int64_t dunit = 0;  // D index table entries
  double gemm_flops_per_col = 0.0;  // 2 * sum over blocks of rows * nq^2
};

// Builds the plan for the space and integrals; plan.ok = false (with a
// reason) when the space is not a complete RAS space of the supported kind.
struct Space;
void ras_plan(RasPlan& P, const Space& sp, const double* one, const double* two);
// A view of the plan's host tables (host pointers), e.g. for a host check.
RasView ras_host_view(const RasPlan& P);

}  // namespace psoci
