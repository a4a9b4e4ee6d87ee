// Determinant/SEF space on the host: configuration masks, determinant
// patterns and phases.
//
// The reference's EVEN-class generators (PsociDeterminants.C:2848-3633),
// used when a configuration's symmetry class (jsym + 1) / 2 equals the
// state's (generateDeterminants, PsociDeterminants.C:1467-1503),
// order the determinants of a configuration with s open shells as follows:
// first the Ms = 0 determinants, in pairs (2k, 2k+1) where 2k puts the k-th
// (s/2)-subset of the open shells -- subsets taken in colexicographic order,
// which is ascending order of their bit masks -- in the alpha list and
// 2k+1 is its spin flip (k < C(s, s/2)/2); then, for Ms = 2, 4, ... <= s/2,
// the (s/2 + Ms)-subsets in ascending mask order, each followed by its spin
// flip.  Every list is the doubly occupied orbitals followed by the chosen
// open shells in ascending order, so computePhaseAndSort
// (PsociDeterminants.C:4355-4469) gives every determinant of a configuration
// the same phase: the parity of the number of (singly, doubly) orbital pairs
// with the doubly occupied orbital above the singly occupied one.
//
// The ODD-class generators (PsociDeterminants.C:3634-4353), used when the
// classes differ, list the determinants of odd Ms instead: for Ms = 1, 3, ...
// <= s/2 the (s/2 + Ms)-subsets in ascending mask order, each followed by
// its spin flip; a closed shell keeps its one determinant.  The lists have
// the same structure, hence the same phase rule.
//
// With an odd electron count the fractional generators (PsociDeterminants.C:
// 1508-1561, 1680-2443) give one determinant per SEF, in blocks of
// 2 Ms = 1, -3, 5, -7, 9: each block's majority-spin subsets in ascending
// mask order (so for a negative block the alpha masks, their complements,
// descend); when (jsym + 1) / 2 is even, alpha and beta are swapped
// (pushBackToJoutfg with isgn = -1, :2833-2843).
#include <algorithm>
#include <cstdlib>

#include "common.hpp"

namespace psoci {

static void patterns_for(int s, bool odd, std::vector<uint16_t>& out) {
  if (s == 0) {
    out.push_back(0);
    return;
  }
  const uint32_t full = (1u << s) - 1u;
  const int h = s / 2;
  std::vector<uint32_t> subs;
  auto collect = [&](int m) {
    subs.clear();
    for (uint32_t x = 0; x <= full; ++x)
      if (__builtin_popcount(x) == m) subs.push_back(x);
  };
  if (odd) {
    for (int ms = 1; ms <= h; ms += 2) {
      collect(h + ms);
      for (uint32_t x : subs) {
        out.push_back((uint16_t)x);
        out.push_back((uint16_t)(full ^ x));
      }
    }
    return;
  }
  collect(h);
  for (size_t k = 0; k < subs.size() / 2; ++k) {
    out.push_back((uint16_t)subs[k]);
    out.push_back((uint16_t)(full ^ subs[k]));
  }
  for (int ms = 2; ms <= h; ms += 2) {
    collect(h + ms);
    for (uint32_t x : subs) {
      out.push_back((uint16_t)x);
      out.push_back((uint16_t)(full ^ x));
    }
  }
}

static void patterns_frac(int s, bool swap, std::vector<uint16_t>& out) {
  const uint32_t full = (1u << s) - 1u;
  int sign = 1;
  for (int t = 1; t <= s; t += 2, sign = -sign) {
    const int m = (s + t) / 2;  // the block's majority-spin count
    for (uint32_t x = 0; x <= full; ++x) {
      if (__builtin_popcount(x) != m) continue;
      const uint32_t alpha = sign > 0 ? x : (full ^ x);
      out.push_back((uint16_t)(swap ? (full ^ alpha) : alpha));
    }
  }
}

bool class_flag(int nelec, int ksym, int js) {
  return nelec % 2 == 1 ? ((js + 1) / 2) % 2 == 0 : (js + 1) / 2 != (ksym + 1) / 2;
}

void build_space(Space& sp, int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym) {
  PSOCI_CHECK(nbf > 0 && nbf <= 64, PSOCI_ERR_UNSUPPORTED, "nbf must be in 1..64");
  PSOCI_CHECK(ncfg > 0, PSOCI_ERR_ARG, "empty configuration list");
  PSOCI_CHECK(ncfg < (int64_t)1 << 31, PSOCI_ERR_UNSUPPORTED, "more than 2^31 configurations");
  sp = Space();
  sp.nbf = nbf;
  sp.nelec = nelec;
  sp.ksym = ksym;
  sp.ncfg = ncfg;
  sp.jsym0 = jsym[0];
  sp.frac = nelec % 2 == 1;
  // determinant class of a configuration (generateDeterminants,
  // PsociDeterminants.C:1467-1503): N odd -> fractional, alpha and beta
  // swapped when (jsym + 1) / 2 is even; N even -> EVEN when (jsym + 1) / 2
  // equals the state's, else ODD.  Either way a boolean selects the family.
  auto family_flag = [&](int js) { return class_flag(nelec, ksym, js); };
  const bool flag0 = family_flag(sp.jsym0);
  sp.flag0 = flag0;
  sp.odd = !sp.frac && flag0;
  sp.mult = (sp.jsym0 % 2 != ksym % 2) ? -1.0 : 1.0;
  sp.jsym.assign(jsym, jsym + ncfg);
  sp.cls.resize(ncfg);
  sp.cmult.resize(ncfg);
  for (int64_t c = 0; c < ncfg; ++c) {
    if (jsym[c] != sp.jsym0) sp.mixed = true;
    sp.cls[c] = family_flag(jsym[c]) != flag0 ? 1 : 0;
    sp.cmult[c] = (jsym[c] % 2 != ksym % 2) ? -1.0 : 1.0;
  }
  sp.didx.assign(2 * 2047, (int16_t)-1);
  for (int f = 0; f < 2; ++f) {
    const bool flag = f == 0 ? flag0 : !flag0;
    for (int s = 0; s <= 10; ++s) {
      sp.pat_off[12 * f + s] = (int)sp.pat.size();
      if (sp.frac && s % 2 == 1)
        patterns_frac(s, flag, sp.pat);
      else if (!sp.frac && s % 2 == 0)
        patterns_for(s, flag, sp.pat);
    }
    sp.pat_off[12 * f + 11] = (int)sp.pat.size();
    for (int s = 0; s <= 10; ++s) {
      const int nd = sp.pat_off[12 * f + s + 1] - sp.pat_off[12 * f + s];
      for (int d = 0; d < nd; ++d)
        sp.didx[2047 * f + (1 << s) - 1 + sp.pat[sp.pat_off[12 * f + s] + d]] = (int16_t)d;
    }
  }
  sp.dm.resize(ncfg);
  sp.sm.resize(ncfg);
  sp.nopen.resize(ncfg);
  sp.nsef.resize(ncfg);
  sp.ndet.resize(ncfg);
  sp.sef_off.resize(ncfg);
  sp.det_off.resize(ncfg);
  sp.cphase.resize(ncfg);
  int64_t soff = 0, doff = 0;
  for (int64_t c = 0; c < ncfg; ++c) {
    uint64_t d = 0, s = 0;
    int e = 0;
    for (int o = 0; o < nbf; ++o) {
      uint8_t v = occ[c * nbf + o];
      PSOCI_CHECK(v <= 2, PSOCI_ERR_ARG, "occupation values must be 0, 1 or 2 (configuration " + std::to_string(c) + ")");
      if (v == 2) d |= 1ull << o;
      if (v == 1) s |= 1ull << o;
      e += v;
    }
    PSOCI_CHECK(e == nelec, PSOCI_ERR_ARG, "configuration " + std::to_string(c) + " has " + std::to_string(e) +
                                               " electrons, expected " + std::to_string(nelec));
    int n = __builtin_popcountll(s);
    PSOCI_CHECK(n <= 10, PSOCI_ERR_UNSUPPORTED, "more than 10 open shells");
    sp.dm[c] = d;
    sp.sm[c] = s;
    sp.nopen[c] = (int8_t)n;
    int h = (n > 0 ? n - 1 : 0) / 2;
    sp.nsef[c] = (1 << h) * (1 << h);  // findParams: 2^(s-2) pairs, or 2^(s-1) determinants for odd s
    sp.ndet[c] = 1 << (n > 0 ? n - 1 : 0);
    // phase: pairs (singly o, doubly d) with d above o
    int par = 0;
    uint64_t t = s;
    while (t) {
      int o = __builtin_ctzll(t);
      t &= t - 1;
      par += __builtin_popcountll(d >> o >> 1);
    }
    sp.cphase[c] = (uint8_t)(par & 1);
    sp.sef_off[c] = soff;
    sp.det_off[c] = doff;
    soff += sp.nsef[c];
    doff += sp.ndet[c];
    sp.max_open = std::max(sp.max_open, n);
  }
  sp.nsef_total = soff;
  sp.ndet_total = doff;
  PSOCI_CHECK(soff < ((int64_t)1 << 31), PSOCI_ERR_UNSUPPORTED, "more than 2^31 SEFs (int32 columns)");
}

}  // namespace psoci
