// Determinant-driven evaluation of one configuration pair's contribution to
// a sigma row: instead of testing every SEF pair (i, j) of the block, each
// determinant of the row configuration enumerates the column determinants it
// actually connects to and adds their elements.
//
// Counting argument.  For determinant a of configuration I and b of J, the
// number of spin-orbitals of a that b does not hold is ex + f, where ex is
// the number of electrons that move between the configurations (moved()) and
// f the number of open shells common to I and J whose spin differs between a
// and b.  Open shells of J that are not open in I (at most four) may take
// either spin at no cost.  Slater–Condon elements vanish beyond two
// replacements, so the connected b are: f = 0 with any admissible spins on
// the new open shells; f = 1 when one electron moves; one alpha-to-beta plus
// one beta-to-alpha flip when none moves (I == J).  The alpha count is kept
// throughout (spin projection is conserved).
//
// SEF assembly.  SEF k of a configuration is (D_2k + mult D_2k+1)/sqrt(2)
// (determinant 2k+1 is the spin flip of 2k; with an odd electron count SEF k
// is determinant k itself, and H_ij = D(i, j)), and every determinant element
// is invariant under flipping both determinants (the phases of detval_ao()
// and detval_x() depend only on spatial occupations).  The reference's
// transformCDCt (PsociHamiltonian.C:1350-1454) reduces to
// H_ij = D(2i, 2j) + mult D(2i, 2j+1) when both configurations have open
// shells; pair_coef() gives the closed-shell cases.  The values agree with
// the stored Hamiltonian's to rounding (an ulp or two), not bit for bit.
#pragma once
#include "elem.hpp"

namespace psoci {

// Coefficients of D(2i, b) in H_ij for b even (ce) and odd (co); with one
// determinant per SEF (frac, odd electron count) H_ij = D(i, j).
struct PairCoef {
  double ce, co;
};

PSOCI_HD PairCoef pair_coef(int nopen_i, int nopen_j, double mult, bool frac = false) {
  PairCoef c;
  if (frac) {
    c.ce = c.co = 1.0;
  } else if (nopen_i) {
    c.ce = nopen_j ? 1.0 : kRthlf * (1.0 + mult);
    c.co = mult;
  } else {
    c.ce = nopen_j ? kRthlf : 1.0;
    c.co = kRthlf * mult;
  }
  return c;
}

// Resumable enumeration of the determinants b of a configuration J != I
// (index within J, alpha open shells ab) within two spin-orbital
// replacements of determinant `aa` of I.  The spins of the new open shells
// run over the subsets of exactly the admissible size (next subset by
// Gosper's rule), and a common open shell is flipped only where such a
// subset exists, so every candidate visited is a connection.  ex: electrons
// moved between the configurations (1 or 2).
struct ConnIter {
  uint64_t base, flips, flip;
  uint32_t norb, npos, pbase, pflip, q, qend;
  int nb, nn;
};

// The open shells of J that are not open in I (at most four: each ends a
// moved electron), packed as orbital (6 bits) and bit of J's pattern
// (4 bits) per entry, and their count.  A property of the configuration
// pair, so the direct sigma keeps it in the pair record.
struct NewShells {
  uint32_t norb, npos;
  int nn;
};

PSOCI_HD NewShells new_shells(uint64_t sI, uint64_t sJ) {
  NewShells w{0u, 0u, 0};
  for (uint64_t r = sJ & ~sI; r; r &= r - 1, ++w.nn) {
    const int o = ctz(r);
    w.norb |= (uint32_t)o << (6 * w.nn);
    w.npos |= (uint32_t)popc(sJ & ((1ull << o) - 1)) << (4 * w.nn);
  }
  return w;
}

PSOCI_HD void conn_begin(ConnIter& it, const ConfView& I, const ConfView& J, uint64_t aa, int ex, const NewShells& w) {
  const uint64_t C = I.s & J.s;
  it.nb = popc(I.d) - popc(J.d) + popc(aa);
  it.base = aa & C;
  it.norb = w.norb;
  it.npos = w.npos;
  it.nn = w.nn;
  it.pbase = 0;
  for (uint64_t r = it.base; r; r &= r - 1) it.pbase |= 1u << popc(J.s & ((1ull << ctz(r)) - 1));
  const int need0 = it.nb - popc(it.base);
  // one electron moved: a common open shell may change spin; alpha-to-beta
  // needs one more alpha among the new shells, beta-to-alpha one fewer
  it.flips = 0;
  if (ex == 1) it.flips = (need0 + 1 <= it.nn ? it.base : 0ull) | (need0 >= 1 ? (C & ~it.base) : 0ull);
  it.flip = 0;
  it.pflip = 0;
  it.qend = 1u << it.nn;
  it.q = (need0 >= 0 && need0 <= it.nn) ? (1u << need0) - 1u : it.qend;
}

// Next connection; false when the enumeration is exhausted.
PSOCI_HD bool conn_next(ConnIter& it, const ConfView& J, const int16_t* __restrict__ didxJ, int& b, uint64_t& ab) {
  for (;;) {
    if (it.q < it.qend) {
      const uint32_t q = it.q;
      ab = it.base ^ it.flip;
      uint32_t p = it.pbase ^ it.pflip;
      for (uint32_t u = q; u; u &= u - 1) {
        const int t = __builtin_ctz(u);
        ab |= 1ull << ((it.norb >> (6 * t)) & 63u);
        p |= 1u << ((it.npos >> (4 * t)) & 15u);
      }
      b = (int)didxJ[p];
      if (!q) {
        it.q = it.qend;  // the empty subset is the only one of its size
      } else {
        const uint32_t c = q & (0u - q), r = q + c;
        it.q = (((r ^ q) >> 2) >> __builtin_ctz(c)) | r;
      }
      return true;
    }
    if (!it.flips) return false;
    it.flip = it.flips & (~it.flips + 1);
    it.flips &= it.flips - 1;
    it.pflip = 1u << popc(J.s & (it.flip - 1));
    const int need = it.nb - popc(it.base ^ it.flip);
    it.q = (need >= 0 && need <= it.nn) ? (1u << need) - 1u : it.qend;
  }
}

// Calls f(b, ab) for every connection of determinant aa of I in J != I;
// w = new_shells(I.s, J.s).
template <int EX = 0, class F>
PSOCI_HD void for_each_conn(const ConfView& I, const ConfView& J, uint64_t aa, const NewShells& w,
                            const int16_t* __restrict__ didxJ, F&& f) {
  ConnIter it;
  conn_begin(it, I, J, aa, EX ? EX : moved(I.d, I.s, J.d, J.s), w);
  int b;
  uint64_t ab;
  while (conn_next(it, J, didxJ, b, ab)) f(b, ab);
}

template <int EX = 0, class F>
PSOCI_HD void for_each_conn(const ConfView& I, const ConfView& J, uint64_t aa, const int16_t* __restrict__ didxJ,
                            F&& f) {
  for_each_conn<EX>(I, J, aa, new_shells(I.s, J.s), didxJ, f);
}

// The diagonal block (J = I): determinant aa itself, then its alpha-beta
// spin exchanges (one alpha and one beta open shell trade spins).
struct DiagIter {
  uint64_t aa, remP, remQ, lp, allQ;
  uint32_t pa, pp;
  bool self;
};

PSOCI_HD void diag_begin(DiagIter& it, const ConfView& I, uint64_t aa) {
  it.aa = aa;
  it.pa = 0;
  for (uint64_t r = aa; r; r &= r - 1) it.pa |= 1u << popc(I.s & ((1ull << ctz(r)) - 1));
  it.self = true;
  it.remP = aa;
  it.allQ = I.s & ~aa;
  it.remQ = 0;
  it.lp = 0;
  it.pp = 0;
}

PSOCI_HD bool diag_next(DiagIter& it, const ConfView& I, const int16_t* __restrict__ didxI, int& b, uint64_t& ab) {
  if (it.self) {
    it.self = false;
    b = (int)didxI[it.pa];
    ab = it.aa;
    return true;
  }
  for (;;) {
    if (it.remQ) {
      const uint64_t lq = it.remQ & (~it.remQ + 1);
      it.remQ &= it.remQ - 1;
      b = (int)didxI[it.pp ^ (1u << popc(I.s & (lq - 1)))];
      ab = it.aa ^ it.lp ^ lq;
      return true;
    }
    if (!it.remP) return false;
    it.lp = it.remP & (~it.remP + 1);
    it.remP &= it.remP - 1;
    it.pp = it.pa ^ (1u << popc(I.s & (it.lp - 1)));
    it.remQ = it.allQ;
  }
}

template <class F>
PSOCI_HD void for_each_conn_diag(const ConfView& I, uint64_t aa, const int16_t* __restrict__ didxI, F&& f) {
  DiagIter it;
  diag_begin(it, I, aa);
  int b;
  uint64_t ab;
  while (diag_next(it, I, didxI, b, ab)) f(b, ab);
}

// Element D(a, b) between determinant aa of I (row) and ab of a different
// configuration J, evaluated in the reference orientation (row = the later
// configuration; rowI says whether that is I); xrow: the pair's row of the
// three-index exchange table (single excitations).
template <int KIND = 0>
PSOCI_HD double conn_value(const PairInfo& P, const double* __restrict__ xrow, bool rowI, const ConfView& I,
                           const ConfView& J, uint64_t aa, uint64_t ab, const double* __restrict__ two) {
  ConfView R, C;
  R.d = rowI ? I.d : J.d;
  R.s = rowI ? I.s : J.s;
  R.ph = rowI ? I.ph : J.ph;
  R.nopen = rowI ? I.nopen : J.nopen;
  C.d = rowI ? J.d : I.d;
  C.s = rowI ? J.s : I.s;
  C.ph = rowI ? J.ph : I.ph;
  C.nopen = rowI ? J.nopen : I.nopen;
  R.pat = C.pat = nullptr;
  return detval_x<false, KIND>(P, xrow, R, C, rowI ? aa : ab, rowI ? ab : aa, two);
}

}  // namespace psoci
