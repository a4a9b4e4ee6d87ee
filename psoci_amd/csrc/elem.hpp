// Configuration-pair matrix elements on CDNA4, in bit-mask form.
//
// A configuration is two 64-bit words: d (doubly occupied orbitals) and s
// (singly occupied).  A determinant of a configuration with n open shells is
// an n-bit pattern p over the open shells (bit t set = the t-th open shell,
// in ascending orbital order, carries alpha spin); its alpha and beta
// orbital sets are
//     A = d | deposit(p, s),   B = d | (s & ~deposit(p, s)).
// The reference keeps the same determinants as sorted orbital lists
// (JOUTFG, PsociDeterminants.hpp:60-87); everything it computes from those
// lists -- excitation counts, the orbitals that differ, and the sign windows
// "count orbitals strictly between a and b" -- is a popcount or a
// count-trailing-zeros of these words.
//
// The arithmetic of every element follows the reference's operation order
// exactly (PsociHamiltonian.C:139-388 for the configuration-level sums,
// :860-1118 for determinant pairs, :1350-1454 for the SEF transform); the
// library is compiled with -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define PSOCI_HD __host__ __device__ inline

namespace psoci {

// 1/sqrt(2) rounded to double: sqrt(0.5) as the reference computes it.
constexpr double kRthlf = 0x1.6a09e667f3bcdp-1;

struct PairInfo {
  int32_t I, J;
  // 0 skipped by the reference, 1 diagonal, 2 single, 3 double, 4 spin-orbit
  // single (configurations of different jsym one electron apart)
  int8_t kind;
  int8_t e0, e1, f0, f1;  // iex[0..1], jex[0..1] (orbitals leaving I / entering J)
  int8_t sop;             // kind 4: bit 0 jsym_I odd, bit 1 jsym_I + jsym_J odd
  int8_t pad[2];
  double v0, v1;  // vnt[0], vnt[1]
};

// Packed configuration (32 bytes: one load per configuration).
struct ConfRec {
  uint64_t d, s;
  int32_t sefoff;
  int32_t detoff;
  int8_t nopen;
  uint8_t cph;
  int16_t nsef;
  int32_t pad;
};

// One ordered configuration pair (I, J) of the direct sigma: J, and the
// reference-orientation block data (row = the later configuration): kind,
// excitation orbitals, vnt[0], vnt[1].  meta bits: kind 0-1, row-is-I 2,
// e0 3-9, e1 10-16, f0 17-23, f1 24-30 (127 = none; orbitals 0-63).
struct PairRec {
  int32_t J;
  uint32_t meta;
  double v0, v1;
  // open shells of J that are not open in the segment's configuration (at
  // most four): orbitals, 6 bits each; their positions in J's determinant
  // pattern, 4 bits each (bits 0-15), and their count (bits 16-18)
  uint32_t norb, npos;
};
// 32 bytes: mrci10x24's 3,006,213,765 records take 96.2 GB of HBM
static_assert(sizeof(PairRec) == 32, "PairRec layout");
static_assert(sizeof(ConfRec) == 32, "ConfRec layout");

struct ConfView {
  uint64_t d, s;
  int nopen;
  int ph;              // configuration phase bit (see space.cpp)
  const uint16_t* pat; // determinant patterns of this open-shell count
};

PSOCI_HD int pair_idx(int p, int q) {
  int a = p > q ? p : q, b = p > q ? q : p;
  return a * (a + 1) / 2 + b;
}

// (pq|rs) at the canonical 8-fold slot (PsociIntegrals::canonicalIndex - 1).
PSOCI_HD double g2(const double* __restrict__ two, int p, int q, int r, int s) {
  int ij = pair_idx(p, q), kl = pair_idx(r, s);
  int a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return two[(int64_t)a * (a + 1) / 2 + b];
}
PSOCI_HD int g2_slot(int p, int q, int r, int s) {
  int ij = pair_idx(p, q), kl = pair_idx(r, s);
  int a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return a * (a + 1) / 2 + b;
}

PSOCI_HD double h1(const double* __restrict__ one, int p, int q) { return one[pair_idx(p, q)]; }

// deposit the low bits of p into the set bits of mask, lowest first
PSOCI_HD uint64_t deposit(uint32_t p, uint64_t mask) {
  uint64_t r = 0;
  while (mask) {
    uint64_t low = mask & (~mask + 1);
    if (p & 1u) r |= low;
    p >>= 1;
    mask ^= low;
  }
  return r;
}

// orbitals strictly between a and b
PSOCI_HD uint64_t between(int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  // empty for hi - lo <= 1 without a branch (2 << 63 wraps to 0)
  return ((1ull << hi) - 1ull) & ~((2ull << lo) - 1ull);
}

PSOCI_HD int popc(uint64_t x) { return __builtin_popcountll(x); }
PSOCI_HD int ctz(uint64_t x) { return __builtin_ctzll(x); }

// Electrons that must move to turn configuration I into J (the reference's
// running iex/jex count, compareOrbMap, PsociGADeterminants.C:2249-2275).
PSOCI_HD int moved(uint64_t dI, uint64_t sI, uint64_t dJ, uint64_t sJ) {
  uint64_t eJ = ~(sJ | dJ);
  return 2 * popc(dI & eJ) + popc(dI & sJ) + popc(sI & eJ);
}

// Configuration-pair setup: generateSOCIblocks' early exits and iex/jex lists
// (PsociHamiltonian.C:436-512) and the configuration-level integral sums
// processDiagExcitation / processSingleExcitation / processDoubleExcitation
// (:306-388, :190-261, :139-186), each in the reference's loop order.
// jsI, jsJ: the configurations' jsym; a single excitation between different
// jsym is the spin-orbit branch of processSingleExcitation (:223-238).
PSOCI_HD void pair_setup(PairInfo& P, int I, int J, uint64_t dI, uint64_t sI, int nI, uint64_t dJ, uint64_t sJ,
                                  int nJ, int nbf, const double* __restrict__ one, const double* __restrict__ two,
                                  int jsI = 1, int jsJ = 1) {
  P.I = I;
  P.J = J;
  P.kind = 0;
  P.e0 = P.e1 = P.f0 = P.f1 = -1;
  P.sop = 0;
  P.v0 = P.v1 = 0.0;
  int dn = nI - nJ;
  if (dn > 4 || dn < -4) return;
  // iex: orbitals where occ_I > occ_J (repeated for a difference of two), ascending
  int ie[2] = {-1, -1}, je[2] = {-1, -1}, ni = 0, nj = 0;
  uint64_t occI = dI | sI, occJ = dJ | sJ;
  uint64_t diff = (dI ^ dJ) | (sI ^ sJ);
  while (diff) {
    int o = ctz(diff);
    diff &= diff - 1;
    int oi = (int)((dI >> o) & 1) * 2 + (int)((sI >> o) & 1);
    int oj = (int)((dJ >> o) & 1) * 2 + (int)((sJ >> o) & 1);
    int m = oi - oj;
    for (; m > 0; --m) {
      if (ni >= 2) return;
      ie[ni++] = o;
    }
    for (; m < 0; ++m) {
      if (nj >= 2) return;
      je[nj++] = o;
    }
  }
  (void)occI;
  (void)occJ;
  P.e0 = (int8_t)ie[0];
  P.e1 = (int8_t)ie[1];
  P.f0 = (int8_t)je[0];
  P.f1 = (int8_t)je[1];
  if (I == J) {
    // processDiagExcitation: vnt[0]; the open-shell exchanges are taken per
    // determinant from the integral table in detval().
    double v = 0.0;
    for (int ib = 0; ib < nbf; ++ib) {
      int oi = (int)((dI >> ib) & 1) * 2 + (int)((sI >> ib) & 1);
      if (!oi) continue;
      if (oi == 1) {
        v += h1(one, ib, ib);
      } else {
        v += 2.0 * h1(one, ib, ib) + g2(two, ib, ib, ib, ib);
      }
      for (int jb = 0; jb < ib; ++jb) {
        int oj = (int)((dI >> jb) & 1) * 2 + (int)((sI >> jb) & 1);
        if (!oj) continue;
        int ninj = oi * oj;
        if (ninj == 1) {
          v += g2(two, ib, ib, jb, jb);
        } else {
          double buf = 2.0 * g2(two, ib, ib, jb, jb);
          buf -= g2(two, ib, jb, ib, jb);
          if (ninj == 2) v += buf;
          else v += 2.0 * buf;
        }
      }
    }
    P.v0 = v;
    P.kind = 1;
    return;
  }
  if (ni == 2) {
    double a = g2(two, ie[0], je[0], ie[1], je[1]);
    double b = g2(two, ie[0], je[1], ie[1], je[0]);
    P.v0 = a;
    P.v1 = b;
    P.kind = (a == 0.0 && b == 0.0) ? 0 : 3;
    return;
  }
  if (ni == 1 && jsI != jsJ) {
    // vnt[0] = +-valone(iex, jex) / 2, negated when jex > iex; the block is
    // computed (ciout runs) even when the integral is zero
    int ex = ie[0], jx = je[0];
    double v = h1(one, ex, jx);
    if (v != 0.0) {
      if (jx > ex) v = -v;
      v = v / 2;
    }
    P.v0 = v;
    P.sop = (int8_t)((jsI % 2 == 1 ? 1 : 0) | ((jsI + jsJ) % 2 == 1 ? 2 : 0));
    P.kind = 4;
    return;
  }
  if (ni == 1) {
    int ex = ie[0], jx = je[0];
    double v = h1(one, ex, jx) - g2(two, ex, jx, ex, ex);
    // occupied orbitals of I in ascending order (the reference loops over all
    // orbitals and skips the empty ones)
    uint64_t occ = dI | sI;
    while (occ) {
      const int i = ctz(occ);
      occ &= occ - 1;
      int oi = (int)((dI >> i) & 1) * 2 + (int)((sI >> i) & 1);
      double buf = g2(two, ex, jx, i, i);
      v += buf;
      if (oi == 2) v += buf;
      if (oi == 2 && ((dJ >> i) & 1)) v -= g2(two, ex, i, i, jx);
    }
    P.v0 = v;
    P.kind = (v == 0.0) ? 0 : 2;
    return;
  }
  // identical occupations under two indices: the reference computes no block
}

PSOCI_HD PairRec make_rec(const PairInfo& P, int J, bool rowI) {
  auto o7 = [](int v) { return (uint32_t)(v < 0 ? 127 : v); };
  PairRec r;
  r.J = J;
  r.meta = (uint32_t)P.kind | ((rowI ? 1u : 0u) << 2) | (o7(P.e0) << 3) | (o7(P.e1) << 10) | (o7(P.f0) << 17) |
           (o7(P.f1) << 24);
  r.v0 = P.v0;
  r.v1 = P.v1;
  return r;
}

PSOCI_HD void unpack_rec(const PairRec& r, PairInfo& P, int& J, bool& rowI) {
  auto u7 = [](uint32_t v) { return (int8_t)(v == 127 ? -1 : (int)v); };
  J = r.J;
  P.kind = (int8_t)(r.meta & 3u);
  rowI = (r.meta >> 2) & 1u;
  P.e0 = u7((r.meta >> 3) & 127u);
  P.e1 = u7((r.meta >> 10) & 127u);
  P.f0 = u7((r.meta >> 17) & 127u);
  P.f1 = u7((r.meta >> 24) & 127u);
  P.v0 = r.v0;
  P.v1 = r.v1;
}

// Spin-orbit determinant-pair element (socig, PsociHamiltonian.C:709-858)
// of a kind-4 pair: +-vnt[0] when the column determinant holds at most one
// spin-orbital the row determinant does not (nmarka, nmarkb <= 1); the sign
// counts the configuration phases, the orbitals of the spin list(s) the
// moved electron passes, and the Lx / Ly / Lz parity rules from jsym.
PSOCI_HD double detval_so(const PairInfo& P, const ConfView& A, const ConfView& B, uint64_t ao, uint64_t bo) {
  const uint64_t aI = A.d | ao, bI = A.d | (A.s & ~ao);
  const uint64_t aJ = B.d | bo, bJ = B.d | (B.s & ~bo);
  const int nmarka = popc(aJ & ~aI);
  if (nmarka + popc(bJ & ~bI) > 1) return 0.0;
  const int iex = P.e0, jex = P.f0;
  const bool jsi_odd = P.sop & 1, sum_odd = (P.sop >> 1) & 1;
  const uint64_t below_i = (1ull << iex) - 1ull, below_j = (1ull << jex) - 1ull;
  const uint64_t above_i = ~((2ull << iex) - 1ull), above_j = ~((2ull << jex) - 1ull);
  const int nai = popc(aI), naj = popc(aJ);
  int ncount = A.ph + B.ph;
  if (nai < naj) {
    // i-beta j-alpha: beta orbitals of ii below iex, alpha of jj above jex
    ncount += popc(bI & below_i) + popc(aJ & above_j);
    if (!sum_odd || jsi_odd) ++ncount;  // Ly, or Lx with jsym_I odd
  } else if (nai > naj) {
    // i-alpha j-beta
    ncount += popc(aI & above_i) + popc(bJ & below_j);
    if (sum_odd && jsi_odd) ++ncount;
  } else {
    // i-alpha j-alpha or i-beta j-beta: orbitals strictly between iex and jex
    if (nmarka == 0) ++ncount;
    if (jsi_odd) ++ncount;
    ncount += popc((nmarka != 0 ? aI : bI) & between(iex, jex));
  }
  return (ncount & 1) ? -P.v0 : P.v0;
}

// One determinant-pair element (generateSEFCIG, PsociHamiltonian.C:860-1118;
// socig for spin-orbit pairs).  ao / bo: the alpha open shells (orbital
// masks) of determinant ii of the row configuration and jj of the column
// configuration.
PSOCI_HD double detval_ao(const PairInfo& P, const ConfView& A, const ConfView& B, uint64_t ao, uint64_t bo,
                          const double* __restrict__ two) {
  if (P.kind == 4) return detval_so(P, A, B, ao, bo);
  uint64_t aI = A.d | ao, bI = A.d | (A.s & ~ao);
  uint64_t aJ = B.d | bo, bJ = B.d | (B.s & ~bo);
  if (popc(aI) != popc(aJ)) return 0.0;  // spin projections differ
  uint64_t xa = aJ & ~aI, xb = bJ & ~bI;
  int ma = popc(xa), mb = popc(xb);
  int tot = ma + mb;
  if (tot > 2) return 0.0;
  if (tot == 0) {
    // same determinant: vnt[0] minus same-spin open-shell exchanges, pairs
    // (singly[i], singly[j]) with j < i in ascending order
    double v = P.v0;
    uint64_t si = A.s;
    uint64_t seen = 0;
    while (si) {
      int io = ctz(si);
      si &= si - 1;
      uint64_t sj = seen;
      int ia = (int)((aI >> io) & 1);
      while (sj) {
        int jo = ctz(sj);
        sj &= sj - 1;
        if (ia == (int)((aI >> jo) & 1)) v -= g2(two, io, jo, jo, io);
      }
      seen |= 1ull << io;
    }
    return v;
  }
  int cnt = A.ph + B.ph;
  double v;
  if (tot == 1) {
    uint64_t S = ma ? aI : bI;
    cnt += popc(S & between(P.e0, P.f0));
    v = P.v0;
    uint64_t T = S & ~A.d & ~(1ull << P.e0) & ~(1ull << P.f0);
    while (T) {
      int ia = ctz(T);
      T &= T - 1;
      v -= g2(two, P.e0, ia, ia, P.f0);
    }
  } else if (ma == 1) {
    int iex2 = ctz(xa), jex2 = ctz(xb);
    int iex3 = ctz(aI & ~aJ), jex3 = ctz(bI & ~bJ);
    cnt += popc(aI & between(iex2, iex3)) + popc(bI & between(jex2, jex3));
    v = g2(two, jex2, jex3, iex2, iex3);
  } else {
    uint64_t SI = ma ? aI : bI, SJ = ma ? aJ : bJ;
    cnt += popc(SJ & between(P.f0, P.f1)) + popc(SI & between(P.e0, P.e1));
    v = P.v0 - P.v1;
  }
  return (cnt & 1) ? -v : v;
}

PSOCI_HD double detval(const PairInfo& P, const ConfView& A, const ConfView& B, uint32_t pa, uint32_t pb,
                       const double* __restrict__ two) {
  return detval_ao(P, A, B, deposit(pa, A.s), deposit(pb, B.s), two);
}

// SEF element (i, j) of the block: transformCDCt (PsociHamiltonian.C:1350-1454)
// applied to the two 2x2 determinant sub-blocks it reads; SEF k pairs
// determinants 2k and 2k+1 (determinantToDoublegroupMaps, :1304-1322);
// mult = -1 when the configuration's symmetry parity differs from the
// state's (ciout, :1487-1495).  aI0/aI1: alpha open shells of determinants
// 2i and 2i+1 of the row configuration (aI1 unused without open shells),
// aJ0/aJ1 likewise for 2j, 2j+1 of the column configuration.
//
// frac: odd electron count, one determinant per SEF (transformCDCt returns
// detdet unchanged, :1386-1391): the element is D(aI0, aJ0).
PSOCI_HD double sef_elem_ao(const PairInfo& P, const ConfView& A, const ConfView& B, uint64_t aI0, uint64_t aI1,
                            uint64_t aJ0, uint64_t aJ1, double multi, double multj, const double* __restrict__ two,
                            bool frac = false) {
  if (frac) return detval_ao(P, A, B, aI0, aJ0, two);
  double ds0, ds1 = 0.0;
  if (B.nopen == 0) {
    ds0 = detval_ao(P, A, B, aI0, aJ0, two);
    if (A.nopen) ds1 = detval_ao(P, A, B, aI1, aJ0, two);
  } else {
    double t = detval_ao(P, A, B, aI0, aJ0, two);
    t += multj * detval_ao(P, A, B, aI0, aJ1, two);
    ds0 = t * kRthlf;
    if (A.nopen) {
      double u = detval_ao(P, A, B, aI1, aJ0, two);
      u += multj * detval_ao(P, A, B, aI1, aJ1, two);
      ds1 = u * kRthlf;
    }
  }
  if (A.nopen == 0) return ds0;
  double t = kRthlf * ds0;
  t += (multi * kRthlf) * ds1;
  return t;
}

// ---------------------------------------------------------------------------
// Off-diagonal elements from a three-index table.
//
// For a single excitation e0 -> f0 every two-electron integral a determinant
// pair needs is (e0 c|c f0) for an open shell c of the row configuration:
// the exchange sum of generateSEFCIG's single branch runs over such c, and
// its alpha-beta branch (the moved electron flips spin while c flips the
// other way) reads (f0 c|c e0) or (c e0|f0 c), the same canonical slot.  For
// a double excitation the alpha-beta branch reads vnt[0] or vnt[1].  So with
// xrow[o] = (e0 o|o f0) (row (e0, f0) of XT[e][f][o] = (e o|o f), nbf^3
// doubles) the element needs no canonical-index arithmetic; the values and
// their summation order are the reference's, so elements stay bit-identical
// to detval_ao().

// XT[(e * nbf + f) * nbf + o] = (e o|o f), on the host
inline void make_xt(int nbf, const double* two, double* xt) {
  for (int e = 0; e < nbf; ++e)
    for (int f = 0; f < nbf; ++f)
      for (int o = 0; o < nbf; ++o) xt[((int64_t)e * nbf + f) * nbf + o] = g2(two, e, o, o, f);
}

// detval_ao() for off-diagonal pairs (kind 2 or 3); xrow as above (unused
// for kind 3).  SAME_OK = false drops the identical-determinant branch, which
// a pair of distinct configurations never reaches.  KIND = 2 or 3 promises
// that kind and a determinant pair within two replacements (as enumerated by
// for_each_conn), dropping the branches that cannot occur: a double
// excitation is always two replacements, and the two-replacement pairs of a
// single excitation are alpha-beta (the moved electron and a flipped open
// shell trade spins).
template <bool SAME_OK = true, int KIND = 0>
PSOCI_HD double detval_x(const PairInfo& P, const double* __restrict__ xrow, const ConfView& A, const ConfView& B,
                         uint64_t ao, uint64_t bo, const double* __restrict__ two) {
  uint64_t aI = A.d | ao, bI = A.d | (A.s & ~ao);
  uint64_t aJ = B.d | bo, bJ = B.d | (B.s & ~bo);
  uint64_t xa = aJ & ~aI, xb = bJ & ~bI;
  int ma = popc(xa), mb = popc(xb);
  int tot = ma + mb;
  if constexpr (KIND == 0) {
    if (popc(aI) != popc(aJ)) return 0.0;
    if (tot > 2 || tot == 0) {
      if constexpr (SAME_OK) return tot == 0 ? detval_ao(P, A, B, ao, bo, two) : 0.0;
      return 0.0;
    }
  }
  int cnt = A.ph + B.ph;
  double v;
  if (KIND != 3 && tot == 1) {
    uint64_t S = ma ? aI : bI;
    cnt += popc(S & between(P.e0, P.f0));
    v = P.v0;
    // the same-spin exchange terms in ascending orbital order, four loads
    // issued before their subtractions (a skipped slot subtracts 0.0, which
    // leaves v unchanged): the result is the plain loop's, bit for bit
    uint64_t T = S & ~A.d & ~(1ull << P.e0) & ~(1ull << P.f0);
    while (T) {
      double x0 = 0.0, x1 = 0.0, x2 = 0.0, x3 = 0.0;
      x0 = xrow[ctz(T)];
      T &= T - 1;
      if (T) {
        x1 = xrow[ctz(T)];
        T &= T - 1;
      }
      if (T) {
        x2 = xrow[ctz(T)];
        T &= T - 1;
      }
      if (T) {
        x3 = xrow[ctz(T)];
        T &= T - 1;
      }
      v -= x0;
      v -= x1;
      v -= x2;
      v -= x3;
    }
  } else if (KIND == 2 || ma == 1) {
    int iex2 = ctz(xa), jex2 = ctz(xb);
    int iex3 = ctz(aI & ~aJ), jex3 = ctz(bI & ~bJ);
    cnt += popc(aI & between(iex2, iex3)) + popc(bI & between(jex2, jex3));
    if (KIND == 3 || (KIND == 0 && P.kind == 3)) {
      // (jex2 jex3|iex2 iex3): alpha lost iex3, gained iex2
      v = ((iex3 == P.e0) == (iex2 == P.f0)) ? P.v0 : P.v1;
    } else {
      int c = (iex2 == jex3) ? iex2 : ((iex3 == jex2) ? iex3 : -1);
      if (c >= 0 && ((A.s >> c) & 1ull) && c != P.e0 && c != P.f0)
        v = xrow[c];
      else
        v = g2(two, jex2, jex3, iex2, iex3);
    }
  } else {
    uint64_t SI = ma ? aI : bI, SJ = ma ? aJ : bJ;
    cnt += popc(SJ & between(P.f0, P.f1)) + popc(SI & between(P.e0, P.e1));
    v = P.v0 - P.v1;
  }
  return (cnt & 1) ? -v : v;
}

// sef_elem_ao() for off-diagonal pairs, same operation order
PSOCI_HD double sef_elem_x(const PairInfo& P, const double* __restrict__ xrow, const ConfView& A, const ConfView& B,
                           uint64_t aI0,
                           uint64_t aI1, uint64_t aJ0, uint64_t aJ1, double multi, double multj,
                           const double* __restrict__ two, bool frac = false) {
  if (frac) return detval_x(P, xrow, A, B, aI0, aJ0, two);
  double ds0, ds1 = 0.0;
  if (B.nopen == 0) {
    ds0 = detval_x(P, xrow, A, B, aI0, aJ0, two);
    if (A.nopen) ds1 = detval_x(P, xrow, A, B, aI1, aJ0, two);
  } else {
    double t = detval_x(P, xrow, A, B, aI0, aJ0, two);
    t += multj * detval_x(P, xrow, A, B, aI0, aJ1, two);
    ds0 = t * kRthlf;
    if (A.nopen) {
      double u = detval_x(P, xrow, A, B, aI1, aJ0, two);
      u += multj * detval_x(P, xrow, A, B, aI1, aJ1, two);
      ds1 = u * kRthlf;
    }
  }
  if (A.nopen == 0) return ds0;
  double t = kRthlf * ds0;
  t += (multi * kRthlf) * ds1;
  return t;
}

PSOCI_HD double sef_elem(const PairInfo& P, const ConfView& A, const ConfView& B, int i, int j, double multi,
                         double multj, const double* __restrict__ two, bool frac = false) {
  if (frac) return detval_ao(P, A, B, deposit(A.pat[i], A.s), deposit(B.pat[j], B.s), two);
  const uint64_t aI0 = deposit(A.pat[A.nopen ? 2 * i : 0], A.s);
  const uint64_t aI1 = A.nopen ? deposit(A.pat[2 * i + 1], A.s) : 0;
  const uint64_t aJ0 = deposit(B.pat[B.nopen ? 2 * j : 0], B.s);
  const uint64_t aJ1 = B.nopen ? deposit(B.pat[2 * j + 1], B.s) : 0;
  return sef_elem_ao(P, A, B, aI0, aI1, aJ0, aJ1, multi, multj, two);
}

}  // namespace psoci
