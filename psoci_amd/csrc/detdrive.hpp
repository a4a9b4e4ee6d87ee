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
// (determinant 2k+1 is the spin flip of 2k), and every determinant element
// is invariant under flipping both determinants (the phases of detval_ao()
// and detval_x() depend only on spatial occupations).  The reference's
// transformCDCt (PsociHamiltonian.C:1350-1454) reduces to
// H_ij = D(2i, 2j) + mult D(2i, 2j+1) when both configurations have open
// shells; pair_coef() gives the closed-shell cases.  The values agree with
// the stored Hamiltonian's to rounding (an ulp or two), not bit for bit.
#pragma once
#include "elem.hpp"

namespace psoci {

// Coefficients of D(2i, b) in H_ij for b even (ce) and odd (co).
struct PairCoef {
  double ce, co;
};

PSOCI_HD PairCoef pair_coef(int nopen_i, int nopen_j, double mult) {
  PairCoef c;
  if (nopen_i) {
    c.ce = nopen_j ? 1.0 : kRthlf * (1.0 + mult);
    c.co = mult;
  } else {
    c.ce = nopen_j ? kRthlf : 1.0;
    c.co = kRthlf * mult;
  }
  return c;
}

// Calls f(b, ab) for every determinant b of a configuration J != I (index
// within J, alpha open shells ab) within two spin-orbital replacements of
// determinant `aa` of I.  didxJ: determinant index by alpha pattern for J's
// open-shell count.  The spins of the new open shells run over the subsets
// of exactly the admissible size (next subset by Gosper's rule), and a
// common open shell is flipped only where such a subset exists, so every
// candidate visited is a connection.
template <int EX = 0, class F>
PSOCI_HD void for_each_conn(const ConfView& I, const ConfView& J, uint64_t aa, const int16_t* __restrict__ didxJ,
                            F&& f) {
  const uint64_t C = I.s & J.s, N = J.s & ~I.s;
  // EX: the number of moved electrons when the caller knows it (1 or 2)
  const int ex = EX ? EX : moved(I.d, I.s, J.d, J.s);
  const int nb = popc(I.d) - popc(J.d) + popc(aa);
  const uint64_t base = aa & C;
  // new open shells (at most four: each ends a moved electron), packed as
  // orbital (6 bits) and bit of J's pattern (4 bits) per entry
  uint32_t norb = 0, npos = 0;
  int nn = 0;
  for (uint64_t r = N; r; r &= r - 1, ++nn) {
    const int o = ctz(r);
    norb |= (uint32_t)o << (6 * nn);
    npos |= (uint32_t)popc(J.s & ((1ull << o) - 1)) << (4 * nn);
  }
  uint32_t pbase = 0;
  for (uint64_t r = base; r; r &= r - 1) pbase |= 1u << popc(J.s & ((1ull << ctz(r)) - 1));
  const int need0 = nb - popc(base);
  // one electron moved: a common open shell may change spin; alpha-to-beta
  // needs one more alpha among the new shells, beta-to-alpha one fewer
  uint64_t flips = 0;
  if (ex == 1) flips = (need0 + 1 <= nn ? base : 0ull) | (need0 >= 1 ? (C & ~base) : 0ull);
  uint64_t flip = 0;
  uint32_t pflip = 0;
  int need = need0;
  for (;;) {
    if (need >= 0 && need <= nn) {
      uint32_t q = (1u << need) - 1u;
      const uint32_t qend = 1u << nn;
      while (q < qend) {
        uint64_t ab = base ^ flip;
        uint32_t p = pbase ^ pflip;
        for (uint32_t u = q; u; u &= u - 1) {
          const int t = __builtin_ctz(u);
          ab |= 1ull << ((norb >> (6 * t)) & 63u);
          p |= 1u << ((npos >> (4 * t)) & 15u);
        }
        f((int)didxJ[p], ab);
        if (!q) break;
        const uint32_t c = q & (0u - q), r = q + c;
        q = (((r ^ q) >> 2) >> __builtin_ctz(c)) | r;
      }
    }
    if (!flips) break;
    flip = flips & (~flips + 1);
    flips &= flips - 1;
    pflip = 1u << popc(J.s & (flip - 1));
    need = nb - popc(base ^ flip);
  }
}

// The diagonal block (J = I): determinant aa itself and its alpha-beta spin
// exchanges (one alpha and one beta open shell trade spins).
template <class F>
PSOCI_HD void for_each_conn_diag(const ConfView& I, uint64_t aa, const int16_t* __restrict__ didxI, F&& f) {
  uint32_t pa = 0;
  for (uint64_t r = aa; r; r &= r - 1) pa |= 1u << popc(I.s & ((1ull << ctz(r)) - 1));
  f((int)didxI[pa], aa);
  for (uint64_t p = aa; p; p &= p - 1) {
    const uint64_t lp = p & (~p + 1);
    const uint32_t pp = pa ^ (1u << popc(I.s & (lp - 1)));
    for (uint64_t q = I.s & ~aa; q; q &= q - 1) {
      const uint64_t lq = q & (~q + 1);
      f((int)didxI[pp ^ (1u << popc(I.s & (lq - 1)))], aa ^ lp ^ lq);
    }
  }
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
