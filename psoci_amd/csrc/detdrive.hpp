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

// Calls f(b, ab) for every determinant b of J (index within J, alpha open
// shells ab) within two spin-orbital replacements of determinant `aa` of I.
// didxJ: determinant index by alpha pattern for J's open-shell count.  The
// candidates are walked by one flat loop (a flip set on the common open
// shells times a subset of the new open shells), so f is expanded once.
template <class F>
PSOCI_HD void for_each_conn(const ConfView& I, const ConfView& J, uint64_t aa, const int16_t* __restrict__ didxJ,
                            F&& f) {
  const uint64_t C = I.s & J.s, N = J.s & ~I.s;
  const int ex = moved(I.d, I.s, J.d, J.s);
  const int nb = popc(I.d) - popc(J.d) + popc(aa);
  const uint64_t base = aa & C;
  const int nn = popc(N);  // at most four: each new open shell is a moved electron's end
  // flip sets: {} always; each common open shell when one electron moves;
  // an alpha-beta pair of common open shells when none does
  uint64_t remP = ex == 0 ? base : (ex == 1 ? C : 0ull);
  const uint64_t allQ = ex == 0 ? (C & ~base) : 0ull;
  uint64_t remQ = 0, curP = 0, flip = 0;
  int q = 0;
  for (;;) {
    const uint64_t cb = base ^ flip;
    if (popc((uint64_t)q) == nb - popc(cb)) {
      uint64_t ab = cb, r = N;
      for (int t = 0; t < 4 && r; ++t, r &= r - 1)
        if ((q >> t) & 1) ab |= r & (~r + 1);
      // J's pattern of ab: one bit per open shell of J, ascending
      uint32_t p = 0;
      for (uint64_t u = ab; u; u &= u - 1) p |= 1u << popc(J.s & ((u & (~u + 1)) - 1));
      f((int)didxJ[p], ab);
    }
    // next candidate
    if (++q < (1 << nn)) continue;
    q = 0;
    if (ex == 0) {
      if (!remQ) {
        if (!remP) break;
        curP = remP & (~remP + 1);
        remP &= remP - 1;
        remQ = allQ;
        if (!remQ) break;
      }
      flip = curP | (remQ & (~remQ + 1));
      remQ &= remQ - 1;
    } else {
      if (!remP) break;
      flip = remP & (~remP + 1);
      remP &= remP - 1;
    }
  }
}

// Element D(a, b) between determinant aa of I (row) and ab of a different
// configuration J, evaluated in the reference orientation (row = the later
// configuration; rowI says whether that is I) from the pair's staged values.
PSOCI_HD double conn_value(const PairInfo& P, const PairX& px, bool rowI, const ConfView& I, const ConfView& J,
                           uint64_t aa, uint64_t ab, const double* __restrict__ two) {
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
  return detval_x<false>(P, px, R, C, rowI ? aa : ab, rowI ? ab : aa, two);
}

}  // namespace psoci
