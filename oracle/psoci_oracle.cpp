// PSOCI oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference PSOCI arithmetic for the determinant
// Hamiltonian build and the Davidson sigma path.  It exists to check the
// HIP product library (psoci_amd/) and to time the CPU path beside it; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and never as the thing measured or shipped.
//
// Every function follows the reference algorithm step by step, keeping its
// data structures (per-determinant orbital lists, per-orbital occupation
// arrays) and, above all, its floating-point operation order, so that the
// product's bit-packed GPU formulation can be compared with it bit for bit.
// Citations are to psociRestart-Cori/ in the reference tree.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp).
//
// Scope: EVEN-class determinants (even electron count, configuration
// symmetry class equal to the state class, what a C1 space uses), ODD-class
// ones (the classes differ) and the fractional class (odd electron count).
// The spin-orbit block (socig) is not restated; configurations of different
// symmetry never meet in the spaces the oracle is used on.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// JOUTFG (PsociDeterminants.hpp:60-87), restricted to the fields the
// Hamiltonian path reads.
struct Joutfg {
  int index = 0;  // 1-based configuration index (file order)
  int jsym = 1;
  int num_singly = 0, num_doubly = 0;
  int ndeti = 0, nsefi = 0;
  std::vector<int> singly, doubly;
  std::vector<std::vector<int>> alpha_spins, beta_spins;
  std::vector<int> spin_projection, phases, sumsa, sumsb, nalpha, nbeta;
};

struct Space {
  int nbf = 0, nelec = 0, ksym = 1, mxopen = 0;
  long ncfg = 0;
  std::vector<std::vector<int>> occ;  // [ncfg][nbf] 0/1/2
  std::vector<int> jsym;
  std::vector<int> nopen, ndet, nsef;
  std::vector<long> sef_off;  // computeNsefEntryPoints
  long nsef_total = 0;
  std::vector<Joutfg> conf;  // generated once for small spaces; large spaces
                             // regenerate per use, as the reference does
  bool prebuilt = true;
  std::vector<double> valone, valtwo;
  bool have_ints = false;
  // stored Hamiltonian (lower triangle + diagonal, diagonal last in each row)
  std::vector<long> rowptr;
  std::vector<int> col;
  std::vector<double> val;
  std::vector<double> diag;
  std::string err;
};

inline long nndaf(long i) { return i * (i + 1) / 2; }

// PsociIntegrals::canonicalIndex / TwoEIndex (PsociIntegrals.C:892-909),
// 0-based orbitals in, 1-based packed index out.
inline int two_e_index(int iC, int jC) {
  int i = iC + 1, j = jC + 1;
  int mx = std::max(i, j), mn = std::min(i, j);
  return ((mx - 1) * mx) / 2 + mn;
}
inline int canonical_index(int iC, int jC, int kC, int lC) {
  int ij = two_e_index(iC, jC);
  int kl = two_e_index(kC, lC);
  int mx = std::max(ij, kl), mn = std::min(ij, kl);
  return ((mx - 1) * mx) / 2 + mn;
}

struct Ints {
  const double* one;
  const double* two;
  // fetch1eIntegralValue / fetch2eIntegralValue (PsociIntegrals.C:853-873)
  double h(int i, int j) const { return one[two_e_index(i, j) - 1]; }
  double g(int i, int j, int k, int l) const { return two[canonical_index(i, j, k, l) - 1]; }
};

// ---------------------------------------------------------------------------
// Determinant generation (PsociDeterminants.C:1467-1503, 2848-3633)
// ---------------------------------------------------------------------------

// The generators' nested loops "for i_m = m..s, for i_{m-1} = m-1..i_m - 1,
// ..., for i_1 = 1..i_2 - 1" visit m-subsets of the open shells with the
// outermost (largest) index slowest: colexicographic order.  Subsets are
// returned 1-based and ascending, as the loops write them into jdbl.
void colex_subsets(int s, int m, std::vector<std::vector<int>>& out) {
  out.clear();
  std::vector<int> idx(m);
  for (int t = 0; t < m; ++t) idx[t] = t + 1;
  if (m == 0) { out.push_back({}); return; }
  while (true) {
    out.push_back(idx);
    // advance: smallest position that can move up without passing its neighbour
    int t = 0;
    while (t < m) {
      int limit = (t + 1 < m) ? idx[t + 1] - 1 : s;
      if (idx[t] < limit) break;
      ++t;
    }
    if (t == m) break;
    ++idx[t];
    for (int u = 0; u < t; ++u) idx[u] = u + 1;
  }
}

struct GenState {
  Joutfg* c;
  int msbas;
  std::vector<std::vector<int>> ifiga, ifigb;
  std::vector<int> ms;
  std::vector<int> list_of(const std::vector<int>& sub) const {
    // jdbl = doubly orbitals followed by the chosen open shells (jdbl[nalpha-m..])
    std::vector<int> l(c->doubly);
    for (int v : sub) l.push_back(c->singly[v - 1]);
    return l;
  }
  // Ms = 0 loop with folded placement: ndetb = min(ndeta, A - ndeta),
  // ndetc = min(ndeta + 1, A - 1 - ndeta)  (e.g. PsociDeterminants.C:3214-3231)
  void folded(int s, int m, int A) {
    std::vector<std::vector<int>> subs;
    colex_subsets(s, m, subs);
    int ndeta = 1;
    for (auto& sub : subs) {
      int ndetb = std::min(ndeta, A - ndeta);
      int ndetc = std::min(ndeta + 1, A - 1 - ndeta);
      ms[ndetb - 1] = msbas;
      ifiga[ndetb - 1] = list_of(sub);
      ifigb[ndetc - 1] = list_of(sub);
      ndeta += 2;
    }
  }
  // ascending block: alpha at ndeta-1, beta at ndeta, ms = msbas +/- dms
  // (e.g. PsociDeterminants.C:3236-3265)
  void ascending(int s, int m, int start, int dms) {
    std::vector<std::vector<int>> subs;
    colex_subsets(s, m, subs);
    int ndeta = start;
    for (auto& sub : subs) {
      ms[ndeta - 1] = msbas + dms;
      ms[ndeta] = msbas - dms;
      ifiga[ndeta - 1] = list_of(sub);
      ifigb[ndeta] = list_of(sub);
      ndeta += 2;
    }
  }
  // descending block: alpha at ndeta-1, beta at ndeta-2, ndeta -= 2
  // (e.g. PsociDeterminants.C:3267-3282)
  void descending(int s, int m, int start) {
    std::vector<std::vector<int>> subs;
    colex_subsets(s, m, subs);
    int ndeta = start;
    for (auto& sub : subs) {
      ifiga[ndeta - 1] = list_of(sub);
      ifigb[ndeta - 2] = list_of(sub);
      ndeta -= 2;
    }
  }
};

// generateEVENSpinDeterminants and its per-open-shell-count cases
// (PsociDeterminants.C:1563-1610, 2848-3633).
bool even_determinants(Joutfg& c, int msbas, std::string& err) {
  int s = c.num_singly;
  int ndet = c.ndeti;
  GenState g;
  g.c = &c;
  g.msbas = msbas;
  g.ifiga.assign(ndet, {});
  g.ifigb.assign(ndet, {});
  g.ms.assign(ndet, 0);
  std::vector<int> all(s);
  for (int t = 0; t < s; ++t) all[t] = t + 1;
  switch (s) {
    case 0:  // :2848-2904
      g.ms[0] = msbas;
      g.ifiga[0] = c.doubly;
      g.ifigb[0] = c.doubly;
      break;
    case 2:  // :2906-2971 (jdbl[nalpha-1] written past size(), within reserve())
      g.ms[0] = msbas;
      g.ms[1] = msbas;
      g.ifiga[0] = g.list_of({1});
      g.ifigb[1] = g.list_of({1});
      g.ifiga[1] = g.list_of({2});
      g.ifigb[0] = g.list_of({2});
      break;
    case 4:  // :2973-3072
      g.folded(4, 2, 13);
      g.ms[6] = msbas + 2;
      g.ms[7] = msbas - 2;
      g.ifiga[7] = c.doubly;
      g.ifigb[6] = c.doubly;
      g.ifiga[6] = g.list_of(all);
      g.ifigb[7] = g.list_of(all);
      break;
    case 6:  // :3075-3290
      g.folded(6, 3, 41);
      g.ascending(6, 5, 21, 2);
      g.descending(6, 1, 32);
      break;
    case 8:  // :3293-3414
      g.folded(8, 4, 141);
      g.ascending(8, 6, 71, 2);
      g.descending(8, 2, 126);
      g.ms[126] = msbas + 4;
      g.ms[127] = msbas - 4;
      g.ifiga[127] = c.doubly;
      g.ifigb[126] = c.doubly;
      g.ifiga[126] = g.list_of(all);
      g.ifigb[127] = g.list_of(all);
      break;
    case 10:  // :3418-3620
      g.folded(10, 5, 505);
      g.ascending(10, 7, 253, 2);
      g.descending(10, 3, 492);
      g.ascending(10, 9, 493, 4);
      g.descending(10, 1, 512);
      break;
    default:
      err = "EVEN class supports 0,2,4,6,8,10 open shells";
      return false;
  }
  // pushBackToJoutfg (:2780-2827)
  c.alpha_spins = g.ifiga;
  c.beta_spins = g.ifigb;
  c.spin_projection = g.ms;
  c.nalpha.clear();
  c.nbeta.clear();
  for (auto& v : c.alpha_spins) c.nalpha.push_back((int)v.size());
  for (auto& v : c.beta_spins) c.nbeta.push_back((int)v.size());
  return true;
}

// generateODDSpinDeterminants and its per-open-shell-count cases
// (PsociDeterminants.C:1613-1658, 3634-4353): the determinants of odd spin
// projection (Ms = +-1, +-3, ...), each alpha list followed by its flip; a
// closed shell keeps its single Ms = 0 determinant.
bool odd_determinants(Joutfg& c, int msbas, std::string& err) {
  int s = c.num_singly;
  int ndet = c.ndeti;
  GenState g;
  g.c = &c;
  g.msbas = msbas;
  g.ifiga.assign(ndet, {});
  g.ifigb.assign(ndet, {});
  g.ms.assign(ndet, 0);
  std::vector<int> all(s);
  for (int t = 0; t < s; ++t) all[t] = t + 1;
  switch (s) {
    case 0:  // :3634-3693
      g.ms[0] = msbas;
      g.ifiga[0] = c.doubly;
      g.ifigb[0] = c.doubly;
      break;
    case 2:  // :3695-3766 (jdbl written past size(), within reserve())
      g.ms[0] = msbas + 1;
      g.ms[1] = msbas - 1;
      g.ifiga[0] = g.list_of(all);
      g.ifigb[1] = g.list_of(all);
      g.ifiga[1] = c.doubly;
      g.ifigb[0] = c.doubly;
      break;
    case 4:  // :3768-3857
      g.ascending(4, 3, 1, 1);
      g.descending(4, 1, 8);
      break;
    case 6:  // :3859-4009
      g.ascending(6, 4, 1, 1);
      g.descending(6, 2, 30);
      g.ms[30] = msbas + 3;
      g.ms[31] = msbas - 3;
      g.ifiga[31] = c.doubly;
      g.ifigb[30] = c.doubly;
      g.ifiga[30] = g.list_of(all);
      g.ifigb[31] = g.list_of(all);
      break;
    case 8:  // :4011-4158
      g.ascending(8, 5, 1, 1);
      g.descending(8, 3, 112);
      g.ascending(8, 7, 113, 3);
      g.descending(8, 1, 128);
      break;
    case 10:  // :4160-4353
      g.ascending(10, 6, 1, 1);
      g.descending(10, 4, 420);
      g.ascending(10, 8, 421, 3);
      g.descending(10, 2, 510);
      g.ms[510] = msbas + 5;
      g.ms[511] = msbas - 5;
      g.ifiga[511] = c.doubly;
      g.ifigb[510] = c.doubly;
      g.ifiga[510] = g.list_of(all);
      g.ifigb[511] = g.list_of(all);
      break;
    default:
      err = "ODD class supports 0,2,4,6,8,10 open shells";
      return false;
  }
  c.alpha_spins = g.ifiga;
  c.beta_spins = g.ifigb;
  c.spin_projection = g.ms;
  c.nalpha.clear();
  c.nbeta.clear();
  for (auto& v : c.alpha_spins) c.nalpha.push_back((int)v.size());
  for (auto& v : c.beta_spins) c.nbeta.push_back((int)v.size());
  return true;
}

// generateSpinFractionalDeterminants and its per-open-shell-count cases
// (PsociDeterminants.C:1508-1561, 1680-2443), odd electron counts: one
// determinant per double-group function, Ms = 1/2, -3/2, 5/2, ...  Each
// block lists the subsets of its majority spin in colexicographic order at
// ascending determinant indices and the minority spin's (complementary)
// subsets at descending ones; pushBackToJoutfg with isgn = -1 swaps the
// alpha and beta lists (:2833-2843).
bool fractional_determinants(Joutfg& c, int msbas, int isgn, std::string& err) {
  const int s = c.num_singly;
  const int ndet = c.ndeti;
  GenState g;
  g.c = &c;
  g.msbas = msbas;
  g.ifiga.assign(ndet, {});
  g.ifigb.assign(ndet, {});
  g.ms.assign(ndet, 0);
  // majority block: list[ndeta-1] = subset, ms label, ++ndeta
  auto major = [&](bool alpha, int m, int start, int mslab) {
    std::vector<std::vector<int>> subs;
    colex_subsets(s, m, subs);
    int ndeta = start;
    for (auto& sub : subs) {
      g.ms[ndeta - 1] = mslab;
      (alpha ? g.ifiga : g.ifigb)[ndeta - 1] = g.list_of(sub);
      ++ndeta;
    }
  };
  // minority block: list[ndeta-1] = subset, --ndeta
  auto minor = [&](bool alpha, int m, int start) {
    std::vector<std::vector<int>> subs;
    colex_subsets(s, m, subs);
    int ndeta = start;
    for (auto& sub : subs) {
      (alpha ? g.ifiga : g.ifigb)[ndeta - 1] = g.list_of(sub);
      --ndeta;
    }
  };
  std::vector<int> all(s);
  for (int t = 0; t < s; ++t) all[t] = t + 1;
  auto last = [&](int d, bool alpha_full, int mslab) {
    g.ms[d] = mslab;
    (alpha_full ? g.ifiga : g.ifigb)[d] = g.list_of(all);
    (alpha_full ? g.ifigb : g.ifiga)[d] = c.doubly;
  };
  switch (s) {
    case 1:  // :1680-1747
      g.ms[0] = msbas + (1 + isgn) / 2;
      g.ifigb[0] = c.doubly;
      g.ifiga[0] = g.list_of(all);
      break;
    case 3:  // :1749-1858
      major(true, 2, 1, msbas + (1 + isgn) / 2);
      minor(false, 1, 3);
      last(3, false, msbas + (1 - 3 * isgn) / 2);
      break;
    case 5:  // :1860-2009
      major(true, 3, 1, msbas + (1 + isgn) / 2);
      minor(false, 2, 10);
      major(false, 4, 11, msbas + (1 - 3 * isgn) / 2);
      minor(true, 1, 15);
      last(15, true, msbas + (1 + 5 * isgn) / 2);
      break;
    case 7:  // :2011-2193
      major(true, 4, 1, msbas + (1 + isgn) / 2);
      minor(false, 3, 35);
      major(false, 5, 36, msbas + (1 - 3 * isgn) / 2);
      minor(true, 2, 56);
      major(true, 6, 57, msbas + (1 + 5 * isgn) / 2);
      minor(false, 1, 63);
      last(63, false, msbas + (1 - 7 * isgn) / 2);
      break;
    case 9:  // :2195-2443
      major(true, 5, 1, msbas + (1 + isgn) / 2);
      minor(false, 4, 126);
      major(false, 6, 127, msbas + (1 - 3 * isgn) / 2);
      minor(true, 3, 210);
      major(true, 7, 211, msbas + (1 + 5 * isgn) / 2);
      minor(false, 2, 246);
      major(false, 8, 247, msbas + (1 - 7 * isgn) / 2);
      minor(true, 1, 255);
      last(255, true, msbas + (1 + 9 * isgn) / 2);
      break;
    default:
      err = "the fractional class supports 1,3,5,7,9 open shells";
      return false;
  }
  c.alpha_spins = isgn == 1 ? g.ifiga : g.ifigb;
  c.beta_spins = isgn == 1 ? g.ifigb : g.ifiga;
  c.spin_projection = g.ms;
  c.nalpha.clear();
  c.nbeta.clear();
  for (auto& v : c.alpha_spins) c.nalpha.push_back((int)v.size());
  for (auto& v : c.beta_spins) c.nbeta.push_back((int)v.size());
  return true;
}

// generateDeterminants (PsociDeterminants.C:1467-1503): odd electron counts
// use the fractional generators with isgn = -1 when (jsym + 1) / 2 is even;
// otherwise the EVEN generators when the configuration's symmetry class
// (jsym + 1) / 2 equals the state's, else the ODD ones.
bool class_determinants(Joutfg& c, int ksym, int nelec, int msbas, std::string& err) {
  if (nelec % 2 == 1) return fractional_determinants(c, msbas, ((c.jsym + 1) / 2) % 2 == 0 ? -1 : 1, err);
  if ((c.jsym + 1) / 2 == (ksym + 1) / 2) return even_determinants(c, msbas, err);
  return odd_determinants(c, msbas, err);
}

// computePhaseAndSort (PsociDeterminants.C:4355-4469): per determinant, the
// number of already-listed orbitals larger than each new one (generator
// order), alpha plus beta, mod 2; then sort each list; sums of 1-based orbitals.
void phase_and_sort(Joutfg& c, int nbf) {
  std::vector<int> pa, pb;
  c.sumsa.clear();
  c.sumsb.clear();
  c.phases.clear();
  for (int pass = 0; pass < 2; ++pass) {
    auto& lists = pass == 0 ? c.alpha_spins : c.beta_spins;
    auto& sums = pass == 0 ? c.sumsa : c.sumsb;
    auto& ph = pass == 0 ? pa : pb;
    for (auto& l : lists) {
      std::vector<int> buffer(nbf, 0);
      int lsum = 0, isign = 0;
      for (int orb : l) {
        lsum += orb + 1;
        buffer[orb] = 1;
        for (int i = orb + 1; i < nbf; ++i) isign += buffer[i];
      }
      sums.push_back(lsum);
      ph.push_back(isign);
      std::set<int> sorted(l.begin(), l.end());
      std::copy(sorted.begin(), sorted.end(), l.begin());
    }
  }
  for (size_t d = 0; d < pa.size(); ++d) c.phases.push_back((pa[d] + pb[d]) % 2);
}

// ---------------------------------------------------------------------------
// Block arithmetic (PsociHamiltonian.C:98-1516)
// ---------------------------------------------------------------------------

struct Ham {
  int nbf, nelec, ksym;
  Ints in;

  // processDoubleExcitation (:139-186)
  int double_exc(const std::vector<int>& iex, const std::vector<int>& jex, std::vector<double>& vnt) const {
    vnt.clear();
    vnt.push_back(in.g(iex[0], jex[0], iex[1], jex[1]));
    vnt.push_back(in.g(iex[0], jex[1], iex[1], jex[0]));
    if (vnt[0] == 0.0 && vnt[1] == 0.0) return 0;
    return 1;
  }

  // processSingleExcitation, same-symmetry branch (:190-261)
  int single_exc(const std::vector<int>& iocc, const std::vector<int>& jocc, const std::vector<int>& iex,
                 const std::vector<int>& jex, std::vector<double>& vnt) const {
    vnt.clear();
    int iext = iex[0], jext = jex[0];
    double vnt1 = in.h(iext, jext) - in.g(iext, jext, iext, iext);
    for (int i = 0; i < nbf; ++i) {
      if (iocc[i] != 0) {
        double buf = in.g(iext, jext, i, i);
        vnt1 += buf;
        if (iocc[i] == 2) vnt1 += buf;
        if (iocc[i] == 2 && jocc[i] == 2) vnt1 -= in.g(iext, i, i, jext);
      }
    }
    vnt.push_back(vnt1);
    return vnt1 == 0.0 ? 0 : 1;
  }

  // processDiagExcitation (:306-388)
  // Planted mistakes for the tests that show the spin-orbit invariants catch
  // them (orc_set_so_mutation; 0 = the restatement as is).
  static int& so_mutation() {
    static int m = 0;
    return m;
  }

  // processSingleExcitation, spin-orbit branch (:223-238): configurations of
  // different jsym one electron apart.  vnt[0] = valone(iex, jex) with the
  // sign flipped when jex > iex, halved; a zero integral leaves vnt empty
  // (socig is not called) but the block still goes through ciout.
  bool so_single(const std::vector<int>& iex, const std::vector<int>& jex, std::vector<double>& vnt) const {
    vnt.clear();
    int iext = iex[0], jext = jex[0];
    double vnt1 = in.h(iext, jext);
    if (vnt1 == 0.0) return false;
    if (jext > iext && so_mutation() != 1) vnt1 = -vnt1;
    vnt1 = vnt1 / 2;
    vnt.push_back(vnt1);
    return true;
  }

  // socig (:709-858): determinant pairs at most one spin-orbital apart get
  // +-vnt[0]; the sign counts the configuration phases, the orbitals the
  // moved electron passes in its spin list(s), and the Lx / Ly / Lz parity
  // rules from jsym.
  void socig(const Joutfg& ic, const Joutfg& jc, const std::vector<int>& iex, const std::vector<int>& jex,
             const std::vector<double>& vnt, std::vector<double>& detdet) const {
    int ndeti = ic.ndeti, ndetj = jc.ndeti;
    int jsymi = ic.jsym, jsymj = jc.jsym;
    for (int ii = 0; ii < ndeti; ++ii) {
      const std::vector<int>& aorbsi = ic.alpha_spins[ii];
      const std::vector<int>& borbsi = ic.beta_spins[ii];
      int nalphai = ic.nalpha[ii], nbetai = ic.nbeta[ii];
      std::vector<int> ifiga(nbf, 0), ifigb(nbf, 0);
      for (int ka = 0; ka < nalphai; ++ka) ifiga[aorbsi[ka]] = 1;
      for (int ka = 0; ka < nbetai; ++ka) ifigb[borbsi[ka]] = 1;
      for (int jj = 0; jj < ndetj; ++jj) {
        const std::vector<int>& aorbsj = jc.alpha_spins[jj];
        const std::vector<int>& borbsj = jc.beta_spins[jj];
        int nalphaj = jc.nalpha[jj], nbetaj = jc.nbeta[jj];
        int nmarka = 0;
        for (int ka = 0; ka < nalphaj; ++ka) {
          if (ifiga[aorbsj[ka]] != 1) ++nmarka;
          if (nmarka > 1) break;
        }
        int nmarkb = nmarka;
        for (int ka = 0; ka < nbetaj; ++ka) {
          if (ifigb[borbsj[ka]] != 1) ++nmarkb;
          if (nmarkb > 1) break;
        }
        if (nmarka <= 1 && nmarkb <= 1) {
          int ncount = so_mutation() == 6 ? 0 : ic.phases[ii] + jc.phases[jj];
          if (nalphai < nalphaj) {
            for (int i = 0; i < nbetai; ++i)
              if (borbsi[i] < iex[0]) ++ncount;
            for (int i = 0; i < nalphaj; ++i)
              if (aorbsj[i] > jex[0]) ++ncount;
            if (so_mutation() == 2) {
            } else if ((jsymi + jsymj) % 2 == 0) {
              ++ncount;
            } else if ((jsymi % 2) == 1) {
              ++ncount;
            }
          } else if (nalphai > nalphaj) {
            for (int i = 0; i < nalphai; ++i)
              if (aorbsi[i] > iex[0]) ++ncount;
            for (int i = 0; i < nbetaj; ++i)
              if (borbsj[i] < jex[0]) ++ncount;
            if ((jsymi + jsymj) % 2 != 0 && so_mutation() != 3) {
              if ((jsymi % 2) == 1) ++ncount;
            }
          } else {
            if (nmarka == 0 && so_mutation() != 4) ++ncount;
            if ((jsymi % 2) == 1) ++ncount;
            if (std::abs(iex[0] - jex[0]) > 1 && so_mutation() != 5) {
              int kbs = std::max(iex[0], jex[0]) - 1;
              int ka = iex[0] + jex[0] - kbs;
              if (nmarka != 0) {
                for (int ia = 0; ia < nalphai; ++ia) {
                  if (aorbsi[ia] > kbs) break;
                  if (aorbsi[ia] >= ka) ++ncount;
                }
              } else {
                for (int ia = 0; ia < nbetai; ++ia) {
                  if (borbsi[ia] > kbs) break;
                  if (borbsi[ia] >= ka) ++ncount;
                }
              }
            }
          }
          int index = ii * ndetj + jj;
          detdet[index] = (ncount % 2 == 1) ? -vnt[0] : vnt[0];
        }
      }
    }
  }

  int diag_exc(const Joutfg& c, const std::vector<int>& iocc, std::vector<double>& vnt) const {
    int nopeni = c.num_singly;
    if (nopeni > 1) {
      vnt.assign(1 + nopeni * (nopeni - 1) / 2, 0.0);
      int mm = 1;
      for (int i = 1; i < nopeni; ++i)
        for (int j = 0; j < i; ++j) {
          int a = c.singly[i], b = c.singly[j];
          vnt[mm++] = in.g(a, b, b, a);
        }
    } else {
      vnt.assign(1, 0.0);
    }
    double vnt1 = 0.0;
    for (int ibf = 0; ibf < nbf; ++ibf) {
      if (iocc[ibf] == 0) continue;
      if (iocc[ibf] == 1) {
        vnt1 += in.h(ibf, ibf);
      } else {
        vnt1 += 2.0 * in.h(ibf, ibf) + in.g(ibf, ibf, ibf, ibf);
      }
      for (int jbf = 0; jbf < ibf; ++jbf) {
        if (iocc[jbf] == 0) continue;
        int ninj = iocc[ibf] * iocc[jbf];
        if (ninj == 1) {
          vnt1 += in.g(ibf, ibf, jbf, jbf);
        } else {
          double buf = 2.0 * in.g(ibf, ibf, jbf, jbf);
          buf -= in.g(ibf, jbf, ibf, jbf);
          if (ninj == 2) vnt1 += buf;
          else vnt1 += 2.0 * buf;
        }
      }
    }
    vnt[0] = vnt1;
    return 1;
  }

  // generateSEFCIG (:860-1118)
  void sefcig(const Joutfg& ic, const Joutfg& jc, const std::vector<int>& iex, const std::vector<int>& jex,
              const std::vector<double>& vnt, std::vector<double>& detdet) const {
    int ndeti = ic.ndeti, ndetj = jc.ndeti;
    for (int ii = 0; ii < ndeti; ++ii) {
      const std::vector<int>& aorbsi = ic.alpha_spins[ii];
      const std::vector<int>& borbsi = ic.beta_spins[ii];
      int nalphai = ic.nalpha[ii], nbetai = ic.nbeta[ii];
      std::vector<int> ifiga(nbf, 0), ifigb(nbf, 0);
      for (int k = 0; k < nalphai; ++k) ifiga[aorbsi[k]] = 1;
      for (int k = 0; k < nbetai; ++k) ifigb[borbsi[k]] = 1;
      int nopeni = ic.num_singly;
      int msi = ic.spin_projection[ii];
      for (int jj = 0; jj < ndetj; ++jj) {
        int index = ii * ndetj + jj;
        if (msi != jc.spin_projection[jj]) continue;
        const std::vector<int>& aorbsj = jc.alpha_spins[jj];
        const std::vector<int>& borbsj = jc.beta_spins[jj];
        int nalphaj = jc.nalpha[jj], nbetaj = jc.nbeta[jj];
        int nij[4] = {0, 0, 0, 0};
        int nmarka = 0;
        for (int ka = 0; ka < nalphaj; ++ka) {
          if (ifiga[aorbsj[ka]] != 1) {
            ++nmarka;
            if (nmarka > 2) break;
            nij[nmarka - 1] = ka;
          }
        }
        int nmarkb = nmarka;
        for (int ka = 0; ka < nbetaj; ++ka) {
          if (ifigb[borbsj[ka]] != 1) {
            ++nmarkb;
            if (nmarkb > 2) break;
            nij[nmarkb - 1] = ka;
          }
        }
        if (!(nmarkb <= 2 && nmarka <= 2)) continue;
        if (nmarkb == 0) {
          // no excitation: diagonal determinant
          int mm = 0;
          detdet[index] = vnt[mm];
          if (nopeni > 1) {
            for (int iopen = 1; iopen < nopeni; ++iopen) {
              int iorb = ic.singly[iopen];
              for (int jopen = 0; jopen < iopen; ++jopen) {
                int jorb = jc.singly[jopen];
                ++mm;
                if (ifiga[iorb] == ifiga[jorb]) detdet[index] -= vnt[mm];
              }
            }
          }
        } else if (nmarkb == 1) {
          // single excitation
          int ncount = ic.phases[ii] + jc.phases[jj];
          if (std::abs(iex[0] - jex[0]) > 1) {
            int kbs = std::max(iex[0], jex[0]) - 1;
            int ka = iex[0] + jex[0] - kbs;
            if (nmarka != 0) {
              for (int ia = 0; ia < nalphai; ++ia) {
                if (aorbsi[ia] > kbs) break;
                if (aorbsi[ia] >= ka) ++ncount;
              }
            } else {
              for (int ia = 0; ia < nbetai; ++ia) {
                if (borbsi[ia] > kbs) break;
                if (borbsi[ia] >= ka) ++ncount;
              }
            }
          }
          detdet[index] = vnt[0];
          for (int ia = 0; ia < nbf; ++ia) {
            int temp = (nmarka == 0) ? ifigb[ia] : ifiga[ia];
            if (temp != 0 && ia != iex[0] && ia != jex[0] && ifiga[ia] + ifigb[ia] != 2)
              detdet[index] -= in.g(iex[0], ia, ia, jex[0]);
          }
          if (ncount % 2 == 1) detdet[index] = -detdet[index];
        } else if (nmarka == 1) {
          // alpha-beta double excitation
          int ncount = ic.phases[ii] + jc.phases[jj];
          int iex2 = aorbsj[nij[0]];
          int jex2 = borbsj[nij[1]];
          int iex3 = ic.sumsa[ii] - jc.sumsa[jj] + iex2;
          int jex3 = ic.sumsb[ii] - jc.sumsb[jj] + jex2;
          if (std::abs(iex2 - iex3) > 1) {
            int kbs = std::max(iex2, iex3) - 1;
            int ka = iex2 + iex3 - kbs;
            for (int ia = 0; ia < nalphai; ++ia) {
              if (aorbsi[ia] > kbs) break;
              if (aorbsi[ia] >= ka) ++ncount;
            }
          }
          if (std::abs(jex2 - jex3) > 1) {
            int kbs = std::max(jex2, jex3) - 1;
            int ka = jex2 + jex3 - kbs;
            for (int ia = 0; ia < nbetai; ++ia) {
              if (borbsi[ia] > kbs) break;
              if (borbsi[ia] >= ka) ++ncount;
            }
          }
          detdet[index] = in.g(jex2, jex3, iex2, iex3);
          if (ncount % 2 == 1) detdet[index] = -detdet[index];
        } else {
          // same-spin double excitation
          int ncount = ic.phases[ii] + jc.phases[jj] + nij[1] - nij[0] - 1;
          if (iex[1] - iex[0] > 1) {
            if (nmarka != 0) {
              for (int ia = 0; ia < nalphai; ++ia) {
                if (aorbsi[ia] >= iex[1]) break;
                if (aorbsi[ia] > iex[0]) ++ncount;
              }
            } else {
              for (int ia = 0; ia < nbetai; ++ia) {
                if (borbsi[ia] >= iex[1]) break;
                if (borbsi[ia] > iex[0]) ++ncount;
              }
            }
          }
          // a diagonal-path block (equal indices) of two configurations a
          // double excitation apart reads vnt[1], which the diagonal vnt
          // holds only with two or more open shells (the transition build)
          if (vnt.size() < 2) throw std::runtime_error("vnt[1] read past a one-entry diagonal vnt");
          detdet[index] = vnt[0] - vnt[1];
          if (ncount % 2 == 1) detdet[index] = -detdet[index];
        }
      }
    }
  }

  // generateSEFCIGCIDEN (:1123-1269)
  void sefcig_ciden(const Joutfg& ic, const Joutfg& jc, const std::vector<int>& iex, const std::vector<int>& jex,
                    std::vector<double>& detdet) const {
    int ndeti = ic.ndeti, ndetj = jc.ndeti;
    for (int ii = 0; ii < ndeti; ++ii) {
      const std::vector<int>& aorbsi = ic.alpha_spins[ii];
      const std::vector<int>& borbsi = ic.beta_spins[ii];
      int nalphai = ic.nalpha[ii], nbetai = ic.nbeta[ii];
      std::vector<int> ifiga(nbf, 0), ifigb(nbf, 0);
      for (int k = 0; k < nalphai; ++k) ifiga[aorbsi[k]] = 1;
      for (int k = 0; k < nbetai; ++k) ifigb[borbsi[k]] = 1;
      int msi = ic.spin_projection[ii];
      for (int jj = 0; jj < ndetj; ++jj) {
        int index = ii * ndetj + jj;
        if (msi != jc.spin_projection[jj]) continue;
        const std::vector<int>& aorbsj = jc.alpha_spins[jj];
        const std::vector<int>& borbsj = jc.beta_spins[jj];
        int nmarka = 0;
        for (int ka = 0; ka < jc.nalpha[jj]; ++ka)
          if (ifiga[aorbsj[ka]] != 1) {
            ++nmarka;
            if (nmarka > 1) break;
          }
        int nmarkb = nmarka;
        for (int ka = 0; ka < jc.nbeta[jj]; ++ka)
          if (ifigb[borbsj[ka]] != 1) {
            ++nmarkb;
            if (nmarkb > 1) break;
          }
        if (nmarka > 1 || nmarkb > 1) {
          detdet[index] = 0.0;
        } else if (nmarkb == 0) {
          detdet[index] = 1.0;
        } else {
          int ncount = ic.phases[ii] + jc.phases[jj];
          if (std::abs(iex[0] - jex[0]) > 1) {
            int kbs = std::max(iex[0], jex[0]) - 1;
            int ka = iex[0] + jex[0] - kbs;
            if (nmarka != 0) {
              for (int ia = 0; ia < nalphai; ++ia) {
                if (aorbsi[ia] > kbs) break;
                if (aorbsi[ia] >= ka) ++ncount;
              }
            } else {
              for (int ia = 0; ia < nbetai; ++ia) {
                if (borbsi[ia] > kbs) break;
                if (borbsi[ia] >= ka) ++ncount;
              }
            }
          }
          detdet[index] = 1.0;
          if (ncount % 2 == 1) detdet[index] = -detdet[index];
        }
      }
    }
  }

  // ciout + transformCDCt (:1457-1516, :1350-1454); determinantToDoublegroupMaps
  // (:1304-1322) pairs determinant 2k with 2k+1 for SEF k.
  void ciout(const Joutfg& ic, const Joutfg& jc, const std::vector<double>& detdet, std::vector<double>& sefsef) const {
    int ndeti = ic.ndeti, nsefi = ic.nsefi, ndetj = jc.ndeti, nsefj = jc.nsefi;
    int ncpi, ncpj;
    if (nelec % 2 == 1) {
      ncpi = ncpj = 2;
    } else {
      ncpi = (ic.jsym % 2 != ksym % 2) ? 2 : 1;
      ncpj = (jc.jsym % 2 != ksym % 2) ? 2 : 1;
    }
    const double rthlf = std::sqrt(0.5);
    if (nelec % 2 == 1) {
      sefsef = detdet;
      return;
    }
    double multi = (ncpj == 2) ? -1.0 : 1.0;
    std::vector<double> detsef;
    if (jc.num_singly != 0) {
      detsef.assign((size_t)ndeti * nsefj, 0.0);
      for (int ii = 0; ii < ndeti; ++ii)
        for (int jj = 0; jj < nsefj; ++jj) {
          double temp = detdet[ii * ndetj + 2 * jj];
          temp += multi * detdet[ii * ndetj + 2 * jj + 1];
          detsef[ii * nsefj + jj] = temp * rthlf;
        }
    } else {
      detsef = detdet;
    }
    multi = (ncpi == 2) ? -1.0 : 1.0;
    if (ic.num_singly != 0) {
      sefsef.assign((size_t)nsefi * nsefj, 0.0);
      for (int ii = 0; ii < nsefi; ++ii)
        for (int jj = 0; jj < nsefj; ++jj) {
          double temp = rthlf * detsef[(2 * ii) * nsefj + jj];
          temp += multi * rthlf * detsef[(2 * ii + 1) * nsefj + jj];
          sefsef[ii * nsefj + jj] = temp;
        }
    } else {
      sefsef = detsef;
    }
  }

  void occ_of(const Joutfg& c, std::vector<int>& occ) const {
    occ.assign(nbf, 0);
    for (int o : c.doubly) occ[o] = 2;
    for (int o : c.singly) occ[o] = 1;
  }

  // generateSOCIblocks (:401-558).  Returns the block status: 0 means the
  // reference skips the block (packSefsef is not called).
  int block(const Joutfg& ic, const Joutfg& jc, std::vector<double>& sefsef, std::string& err) const {
    sefsef.assign((size_t)ic.nsefi * jc.nsefi, 0.0);
    std::vector<double> detdet((size_t)ic.ndeti * jc.ndeti, 0.0);
    if (std::abs(ic.num_singly - jc.num_singly) > 4) return 0;
    std::vector<int> iocc, jocc, iex, jex;
    occ_of(ic, iocc);
    occ_of(jc, jocc);
    int sumiex = 0, sumjex = 0;
    for (int i = 0; i < nbf; ++i) {
      int jmarr = iocc[i] - jocc[i];
      if (jmarr < 0) {
        if (sumjex - jmarr > 2) return 0;
        for (int j = -1; j >= jmarr; --j) { jex.push_back(i); ++sumjex; }
      } else if (jmarr > 0) {
        if (sumiex + jmarr > 2) return 0;
        for (int j = 1; j <= jmarr; ++j) { iex.push_back(i); ++sumiex; }
      }
    }
    std::vector<double> vnt;
    int sefstatus = 0;
    if (ic.index == jc.index) {
      if (diag_exc(ic, iocc, vnt) == 1) {
        sefcig(ic, jc, iex, jex, vnt, detdet);
        sefstatus = 1;
      }
    } else {
      int kk = (int)iex.size();
      int status = 0;
      if (kk == 2) {
        status = double_exc(iex, jex, vnt);
      } else if (kk == 1 && ic.jsym != jc.jsym) {
        // spin-orbit single: socig fills detdet (when the integral is not
        // zero) and processSingleExcitation returns 2, which skips sefcig
        // and keeps sefstatus = 1, so ciout runs even on a zero detdet
        // (the outer status is shadowed, PsociHamiltonian.C:531-548)
        if (so_single(iex, jex, vnt)) socig(ic, jc, iex, jex, vnt, detdet);
        status = 2;
      } else if (kk == 1) {
        status = single_exc(iocc, jocc, iex, jex, vnt);
      }
      sefstatus = 1;
      if (status == 1) sefcig(ic, jc, iex, jex, vnt, detdet);
      if (status == 0) sefstatus = 0;
    }
    if (sefstatus == 1) ciout(ic, jc, detdet, sefsef);
    return sefstatus;
  }

  // generateSOCIblocksForCIDEN (:572-705)
  int block_ciden(const Joutfg& ic, const Joutfg& jc, std::vector<double>& sefsef) const {
    sefsef.assign((size_t)ic.nsefi * jc.nsefi, 0.0);
    std::vector<double> detdet((size_t)ic.ndeti * jc.ndeti, 0.0);
    if (std::abs(ic.num_singly - jc.num_singly) > 2) return 0;
    if (ic.jsym != jc.jsym) return 0;
    std::vector<int> iex, jex;
    if (ic.index != jc.index) {
      std::vector<int> iocc, jocc;
      occ_of(ic, iocc);
      occ_of(jc, jocc);
      int sumiex = 0, sumjex = 0;
      for (int i = 0; i < nbf; ++i) {
        int jmarr = iocc[i] - jocc[i];
        if (jmarr < 0) {
          if (sumjex - jmarr > 1) return 0;
          for (int j = -1; j >= jmarr; --j) { jex.push_back(i); ++sumjex; }
        } else if (jmarr > 0) {
          if (sumiex + jmarr > 1) return 0;
          for (int j = 1; j <= jmarr; ++j) { iex.push_back(i); ++sumiex; }
        }
      }
    }
    sefcig_ciden(ic, jc, iex, jex, detdet);
    ciout(ic, jc, detdet, sefsef);
    return 1;
  }
};

// computeConfigs(orbmap, JOUTFG&) (PsociDeterminants.C:1083-1191):
// determinants of one configuration, generated on demand.
void make_conf(const Space& sp, long c, Joutfg& j) {
  j = Joutfg();
  j.index = (int)c + 1;
  j.jsym = sp.jsym[c];
  for (int o = 0; o < sp.nbf; ++o) {
    if (sp.occ[c][o] == 1) j.singly.push_back(o);
    if (sp.occ[c][o] == 2) j.doubly.push_back(o);
  }
  j.num_singly = (int)j.singly.size();
  j.num_doubly = (int)j.doubly.size();
  int h = std::max(j.num_singly - 1, 0) / 2;
  j.nsefi = (1 << h) * (1 << h);
  j.ndeti = 1 << std::max(j.num_singly - 1, 0);
  std::string e;
  class_determinants(j, sp.ksym, sp.nelec, sp.mxopen / 2, e);
  phase_and_sort(j, sp.nbf);
}

const Joutfg& conf_of(const Space& sp, long c, Joutfg& tmp) {
  if (sp.prebuilt) return sp.conf[c];
  make_conf(sp, c, tmp);
  return tmp;
}

Ham make_ham(const Space& sp) {
  Ham h;
  h.nbf = sp.nbf;
  h.nelec = sp.nelec;
  h.ksym = sp.ksym;
  h.in.one = sp.valone.data();
  h.in.two = sp.valtwo.data();
  return h;
}

// compareOrbMap (PsociGADeterminants.C:2249-2275): at most two electrons move.
bool orbmap_connected(const std::vector<int>& a, const std::vector<int>& b, int limit) {
  int sumiex = 0, sumjex = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    int jmarr = a[i] - b[i];
    if (jmarr < 0) {
      if (sumjex - jmarr > limit) return false;
      sumjex -= jmarr;
    } else if (jmarr > 0) {
      if (sumiex + jmarr > limit) return false;
      sumiex += jmarr;
    }
  }
  return true;
}

// compareLocalJSubsetOrbMap (PsociGADeterminants.C:2501-2546) with one J chunk:
// ascending J <= I, |nopen_I - nopen_J| <= 4, at most two electrons moved.
void jlist_for(const Space& sp, long I, std::vector<long>& list) {
  list.clear();
  for (long J = 0; J <= I; ++J) {
    if (std::abs(sp.nopen[I] - sp.nopen[J]) > 4) continue;
    if (orbmap_connected(sp.occ[I], sp.occ[J], 2)) list.push_back(J);
  }
}

void set_err(char* err, int errlen, const std::string& s) {
  if (err && errlen > 0) {
    std::snprintf(err, errlen, "%s", s.c_str());
  }
}

// dsygv-equivalent for the small subspace problem A y = lambda S y (itype 1):
// Cholesky S = L L^T, then cyclic Jacobi on L^-1 A L^-T.  Eigenvalues ascending.
bool small_gen_eig(int m, const std::vector<double>& A, const std::vector<double>& S, std::vector<double>& w,
                   std::vector<double>& Y) {
  std::vector<double> L(m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    double d = S[j * m + j];
    for (int k = 0; k < j; ++k) d -= L[j * m + k] * L[j * m + k];
    if (d <= 0) return false;
    L[j * m + j] = std::sqrt(d);
    for (int i = j + 1; i < m; ++i) {
      double s = S[i * m + j];
      for (int k = 0; k < j; ++k) s -= L[i * m + k] * L[j * m + k];
      L[i * m + j] = s / L[j * m + j];
    }
  }
  // C = L^-1 A L^-T
  std::vector<double> T(m * m), C(m * m);
  for (int c = 0; c < m; ++c)
    for (int i = 0; i < m; ++i) {
      double s = A[i * m + c];
      for (int k = 0; k < i; ++k) s -= L[i * m + k] * T[k * m + c];
      T[i * m + c] = s / L[i * m + i];
    }
  for (int r = 0; r < m; ++r)
    for (int i = 0; i < m; ++i) {
      double s = T[r * m + i];
      for (int k = 0; k < i; ++k) s -= L[i * m + k] * C[r * m + k];
      C[r * m + i] = s / L[i * m + i];
    }
  std::vector<double> V(m * m, 0.0);
  for (int i = 0; i < m; ++i) V[i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += C[p * m + q] * C[p * m + q];
    if (off < 1e-30) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        double apq = C[p * m + q];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (C[q * m + q] - C[p * m + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < m; ++k) {
          double akp = C[k * m + p], akq = C[k * m + q];
          C[k * m + p] = c * akp - s * akq;
          C[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          double apk = C[p * m + k], aqk = C[q * m + k];
          C[p * m + k] = c * apk - s * aqk;
          C[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          double vkp = V[k * m + p], vkq = V[k * m + q];
          V[k * m + p] = c * vkp - s * vkq;
          V[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(m);
  for (int i = 0; i < m; ++i) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return C[a * m + a] < C[b * m + b]; });
  w.resize(m);
  Y.assign(m * m, 0.0);  // Y[row * m + col]: col = eigenvector index
  for (int c = 0; c < m; ++c) {
    int e = ord[c];
    w[c] = C[e * m + e];
    // back-transform y = L^-T v
    std::vector<double> v(m);
    for (int i = 0; i < m; ++i) v[i] = V[i * m + e];
    for (int i = m - 1; i >= 0; --i) {
      double s = v[i];
      for (int k = i + 1; k < m; ++k) s -= L[k * m + i] * v[k];
      v[i] = s / L[i * m + i];
    }
    for (int i = 0; i < m; ++i) Y[i * m + c] = v[i];
  }
  return true;
}

}  // namespace

extern "C" {

// Build the determinant/SEF space: findParams + generateDeterminants +
// computePhaseAndSort per configuration (PsociDeterminants.C:749-840,
// 1083-1191) and computeNsefEntryPoints (PsociGAhamiltonian.C:4022-4048).
// occ is [ncfg][nbf] with values 0/1/2 in file order.
// test hook: plant mistake m in the spin-orbit restatement (0: none)
void orc_set_so_mutation(int m) { Ham::so_mutation() = m; }

void* orc_space_new(int nbf, int nelec, int ksym, long ncfg, const unsigned char* occ, const int* jsym, char* err,
                    int errlen) {
  Space* sp = new Space;
  sp->nbf = nbf;
  sp->nelec = nelec;
  sp->ksym = ksym;
  sp->ncfg = ncfg;
  sp->occ.assign(ncfg, std::vector<int>(nbf));
  sp->jsym.assign(jsym, jsym + ncfg);
  sp->nopen.resize(ncfg);
  sp->ndet.resize(ncfg);
  sp->nsef.resize(ncfg);
  sp->sef_off.resize(ncfg);
  sp->conf.resize(ncfg);
  int mxopen = 0;
  for (long c = 0; c < ncfg; ++c) {
    int s = 0, e = 0;
    for (int o = 0; o < nbf; ++o) {
      int v = occ[c * nbf + o];
      sp->occ[c][o] = v;
      if (v == 1) ++s;
      e += v;
    }
    if (e != nelec) {
      set_err(err, errlen, "configuration electron count differs from nelec");
      delete sp;
      return nullptr;
    }
    mxopen = std::max(mxopen, s);
  }
  sp->mxopen = mxopen;
  int msbas = mxopen / 2;
  std::string e2;
  long off = 0;
  sp->prebuilt = ncfg <= 300000;
  if (!sp->prebuilt) sp->conf.clear();
  for (long c = 0; c < ncfg; ++c) {
    Joutfg jtmp;
    Joutfg& j = sp->prebuilt ? sp->conf[c] : jtmp;
    j.index = (int)c + 1;
    j.jsym = sp->jsym[c];
    for (int o = 0; o < nbf; ++o) {
      if (sp->occ[c][o] == 1) j.singly.push_back(o);
      if (sp->occ[c][o] == 2) j.doubly.push_back(o);
    }
    j.num_singly = (int)j.singly.size();
    j.num_doubly = (int)j.doubly.size();
    int h = std::max(j.num_singly - 1, 0) / 2;
    j.nsefi = (1 << h) * (1 << h);
    j.ndeti = 1 << std::max(j.num_singly - 1, 0);
    if (j.num_singly > 10) {
      set_err(err, errlen, "at most 10 open shells");
      delete sp;
      return nullptr;
    }
    if (sp->prebuilt) {
      if (!class_determinants(j, ksym, nelec, msbas, e2)) {
        set_err(err, errlen, e2);
        delete sp;
        return nullptr;
      }
      phase_and_sort(j, nbf);
    }
    sp->nopen[c] = j.num_singly;
    sp->ndet[c] = j.ndeti;
    sp->nsef[c] = j.nsefi;
    sp->sef_off[c] = off;
    off += j.nsefi;
  }
  sp->nsef_total = off;
  return sp;
}

void orc_space_free(void* h) { delete static_cast<Space*>(h); }

long orc_space_nsef_total(void* h) { return static_cast<Space*>(h)->nsef_total; }

void orc_space_info(void* h, int* nopen, int* ndet, int* nsef, long* sef_off) {
  Space* sp = static_cast<Space*>(h);
  for (long c = 0; c < sp->ncfg; ++c) {
    if (nopen) nopen[c] = sp->nopen[c];
    if (ndet) ndet[c] = sp->ndet[c];
    if (nsef) nsef[c] = sp->nsef[c];
    if (sef_off) sef_off[c] = sp->sef_off[c];
  }
}

// Determinant lists of configuration I after computePhaseAndSort: alpha/beta
// as [ndet][nbf] (sorted, -1 padded), Ms label, phase, sums of 1-based orbitals.
// With raw != 0 the lists are returned in generator order instead (the
// reference's lists before sorting); phases/sums are unaffected.
int orc_dets(void* h, long I, int raw, int* alpha, int* beta, int* ms, int* phase, int* suma, int* sumb) {
  Space* sp = static_cast<Space*>(h);
  if (I < 0 || I >= sp->ncfg) return -1;
  Joutfg ctmp;
  Joutfg c = conf_of(*sp, I, ctmp);
  if (raw) {
    int msbas = sp->mxopen / 2;
    std::string e;
    Joutfg r;
    r.index = c.index;
    r.jsym = c.jsym;
    r.singly = c.singly;
    r.doubly = c.doubly;
    r.num_singly = c.num_singly;
    r.num_doubly = c.num_doubly;
    r.ndeti = c.ndeti;
    r.nsefi = c.nsefi;
    class_determinants(r, sp->ksym, sp->nelec, msbas, e);
    c.alpha_spins = r.alpha_spins;
    c.beta_spins = r.beta_spins;
  }
  int nbf = sp->nbf;
  for (int d = 0; d < c.ndeti; ++d) {
    for (int o = 0; o < nbf; ++o) {
      alpha[d * nbf + o] = o < (int)c.alpha_spins[d].size() ? c.alpha_spins[d][o] : -1;
      beta[d * nbf + o] = o < (int)c.beta_spins[d].size() ? c.beta_spins[d][o] : -1;
    }
    ms[d] = c.spin_projection[d] - sp->mxopen / 2;
    phase[d] = c.phases[d];
    suma[d] = c.sumsa[d];
    sumb[d] = c.sumsb[d];
  }
  return c.ndeti;
}

int orc_set_integrals(void* h, const double* valone, const double* valtwo) {
  Space* sp = static_cast<Space*>(h);
  long n1 = nndaf(sp->nbf), n2 = nndaf(n1);
  sp->valone.assign(valone, valone + n1);
  sp->valtwo.assign(valtwo, valtwo + n2);
  sp->have_ints = true;
  return 0;
}

// One configuration-pair block, generateSOCIblocks semantics (no threshold).
int orc_block(void* h, long I, long J, double* sefsef) {
  Space* sp = static_cast<Space*>(h);
  Ham hm = make_ham(*sp);
  std::vector<double> blk;
  Joutfg ti, tj;
  int st = hm.block(conf_of(*sp, I, ti), conf_of(*sp, J, tj), blk, sp->err);
  std::copy(blk.begin(), blk.end(), sefsef);
  return st;
}

// The J list the reference visits for row configuration I.
long orc_jlist(void* h, long I, long* out) {
  Space* sp = static_cast<Space*>(h);
  std::vector<long> l;
  jlist_for(*sp, I, l);
  if (out) std::copy(l.begin(), l.end(), out);
  return (long)l.size();
}

// Stored-H construction: constructGlobalHamiltonianAltOrbMapDirectChunkJmap
// (PsociGAhamiltonian.C:2530-2857) with one J chunk, packSefsef (:4077-4136)
// and the upload's diagonal extraction.  OpenMP over configuration rows
// mirrors the reference's per-rank iconf decomposition.  Returns nnz or < 0.
long orc_build_h(void* h, int nthreads) {
  Space* sp = static_cast<Space*>(h);
  if (!sp->have_ints) return -1;
  Ham hm = make_ham(*sp);
  long ncfg = sp->ncfg;
  std::vector<std::vector<std::vector<double>>> cim(ncfg);
  std::vector<std::vector<std::vector<int>>> ico(ncfg);
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(| : bad)
#endif
  for (long I = 0; I < ncfg; ++I) {
    Joutfg ti;
    const Joutfg& ic = conf_of(*sp, I, ti);
    std::vector<long> list;
    jlist_for(*sp, I, list);
    auto& cimat = cim[I];
    auto& icicol = ico[I];
    cimat.assign(ic.nsefi, {});
    icicol.assign(ic.nsefi, {});
    std::vector<double> sefsef;
    std::string err;
    for (long J : list) {
      Joutfg tj;
      const Joutfg& jc = conf_of(*sp, J, tj);
      int st = hm.block(ic, jc, sefsef, err);
      if (st < 0) { bad = 1; continue; }
      if (!st) continue;
      long isef = sp->sef_off[I], jsef = sp->sef_off[J];
      int nsefi = ic.nsefi, nsefj = jc.nsefi;
      for (int i = 0; i < nsefi; ++i) {
        int jmax = (isef == jsef) ? i : nsefj - 1;
        for (int j = 0; j <= jmax; ++j) {
          double v = sefsef[i * nsefj + j];
          if (std::fabs(v) >= 1.0e-8 || (isef == jsef && i == j)) {
            cimat[i].push_back(v);
            icicol[i].push_back((int)(j + jsef));
          }
        }
      }
    }
  }
  if (bad) return -2;
  long n = sp->nsef_total;
  sp->rowptr.assign(n + 1, 0);
  sp->diag.assign(n, 0.0);
  long nnz = 0;
  for (long I = 0; I < ncfg; ++I)
    for (int i = 0; i < sp->nsef[I]; ++i) {
      sp->rowptr[sp->sef_off[I] + i] = nnz;
      nnz += (long)cim[I][i].size();
    }
  sp->rowptr[n] = nnz;
  sp->col.resize(nnz);
  sp->val.resize(nnz);
  for (long I = 0; I < ncfg; ++I)
    for (int i = 0; i < sp->nsef[I]; ++i) {
      long r = sp->sef_off[I] + i;
      long p = sp->rowptr[r];
      std::copy(cim[I][i].begin(), cim[I][i].end(), sp->val.begin() + p);
      std::copy(ico[I][i].begin(), ico[I][i].end(), sp->col.begin() + p);
      if (!cim[I][i].empty()) sp->diag[r] = cim[I][i].back();
    }
  return nnz;
}

// Transition Hamiltonian between two configuration lists of one orbital
// basis: constructGlobalTransitionHamiltonian2WAY (PsociGAhamiltonian.C:
// 2089-2287) with packSefsef (:4077-4136).  Rows: the SEFs of list 1; for
// each configuration I of list 1 and each J <= I of list 2 (positions in the
// two lists), generateSOCIblocks(list1[I], list2[J]) (its diagonal path
// whenever the 1-based indices agree, I == J, whatever the configurations)
// and packSefsef with list 1's SEF entry points for both the row and the
// column (the driver computes them from list 1 only, driver.Transition.C:
// 579): columns sef_off1[J] + j, the lower triangle j <= i when I == J, and
// |h| >= 1e-8 or the (I == J, i == j) element.  h1 and h2 share integrals
// (h1's are used).  Returns nnz, or < 0 on error (-3: a diagonal-path block
// reading vnt[1] of a configuration with fewer than two open shells; -4: an
// equal-index block whose list-2 configuration has fewer SEFs, where
// packSefsef reads past the block).
long orc_transition_h(void* h1, void* h2, long* rowptr, int* col, double* val, long cap) {
  Space* s1 = static_cast<Space*>(h1);
  Space* s2 = static_cast<Space*>(h2);
  if (!s1->have_ints || s1->nbf != s2->nbf || s1->nelec != s2->nelec || s1->ksym != s2->ksym) return -1;
  Ham hm = make_ham(*s1);
  long n1 = s1->nsef_total, nnz = 0;
  std::vector<std::vector<double>> vrow(n1);
  std::vector<std::vector<int>> crow(n1);
  try {
    for (long I = 0; I < s1->ncfg; ++I) {
      Joutfg ti;
      const Joutfg& ic = conf_of(*s1, I, ti);
      const long jend = std::min(I, s2->ncfg - 1);
      for (long J = 0; J <= jend; ++J) {
        Joutfg tj;
        const Joutfg& jc = conf_of(*s2, J, tj);
        std::vector<double> sefsef;
        std::string err;
        int st = hm.block(ic, jc, sefsef, err);
        if (st < 0) return -2;
        if (!st) continue;
        // packSefsef's lower triangle of an equal-index block reads
        // sefsef[i * nsefj + j] for j <= i: past the block when nsefj < nsefi
        if (I == J && jc.nsefi < ic.nsefi) return -4;
        long isef = s1->sef_off[I], jsef = s1->sef_off[J];
        for (int i = 0; i < ic.nsefi; ++i) {
          int jmax = (isef == jsef) ? i : jc.nsefi - 1;
          for (int j = 0; j <= jmax; ++j) {
            double v = sefsef[i * jc.nsefi + j];
            if (std::fabs(v) >= 1.0e-8 || (isef == jsef && i == j)) {
              vrow[isef + i].push_back(v);
              crow[isef + i].push_back((int)(j + jsef));
            }
          }
        }
      }
    }
  } catch (const std::runtime_error&) {
    return -3;
  }
  for (long r = 0; r < n1; ++r) {
    if (rowptr) rowptr[r] = nnz;
    for (size_t t = 0; t < vrow[r].size(); ++t, ++nnz) {
      if (nnz < cap) {
        if (col) col[nnz] = crow[r][t];
        if (val) val[nnz] = vrow[r][t];
      }
    }
  }
  if (rowptr) rowptr[n1] = nnz;
  return nnz;
}

void orc_get_h(void* h, long* rowptr, int* col, double* val, double* diag) {
  Space* sp = static_cast<Space*>(h);
  if (rowptr) std::copy(sp->rowptr.begin(), sp->rowptr.end(), rowptr);
  if (col) std::copy(sp->col.begin(), sp->col.end(), col);
  if (val) std::copy(sp->val.begin(), sp->val.end(), val);
  if (diag) std::copy(sp->diag.begin(), sp->diag.end(), diag);
}

// Stored sigma, matrixVectorProductsSplitChunkMultipleJchunks semantics
// (PsociGAhamiltonian.C:7171-7225): for each stored row i,
// y_i += sum_j H_ij x_j (diagonal included) and y_j += H_ij x_i for every
// entry but the last (the diagonal).  X, Y are [n][k] row-major.
void orc_sigma_csr(long n, const long* rowptr, const int* col, const double* val, int k, const double* X,
                   double* Y) {
  std::fill(Y, Y + n * k, 0.0);
  for (int v = 0; v < k; ++v) {
    for (long i = 0; i < n; ++i) {
      long lo = rowptr[i], hi = rowptr[i + 1];
      double dtemp = 0.0;
      for (long p = lo; p < hi; ++p) dtemp += val[p] * X[(long)col[p] * k + v];
      Y[i * k + v] += dtemp;
      double xi = X[i * k + v];
      for (long p = lo; p < hi - 1; ++p) Y[(long)col[p] * k + v] += val[p] * xi;
    }
  }
}

void orc_sigma(void* h, int k, const double* X, double* Y) {
  Space* sp = static_cast<Space*>(h);
  orc_sigma_csr(sp->nsef_total, sp->rowptr.data(), sp->col.data(), sp->val.data(), k, X, Y);
}

// Single-vector Davidson with the production driver's schedule
// (driver.PSOCI.splittest.C:895-1340): unit guesses on the first k SEFs
// (PsociGAvectors.C:687-726), MGS append (PsociGAbasis.C:1371-1470),
// incremental subspace + dsygv (PsociGAsubspace.C:273-362, 1088-1171),
// residual of the lowest unconverged root (:874-989), diagonal
// preconditioner with DAVIDSON_MIN = 1e-3 (PsociGAbasis.C:1860-1920),
// coarse 1e-3 then final rtol, collapse to 2k when above 4k vectors or at the
// coarse->final switch (expandAndReplaceVectorsNew, :483-579).
// Returns total iterations; evals[k], evecs [n][k] row-major (may be null).
int orc_davidson(void* h, int k, double rtol, int maxit, int guess, double* evals, double* evecs, double* rnorms_out,
                 int* nsigma_out) {
  Space* sp = static_cast<Space*>(h);
  long n = sp->nsef_total;
  const double* D = sp->diag.data();
  int maxsub = 4 * k, minvecs = 2;
  int cap = k + maxsub + 1;
  std::vector<std::vector<double>> V, HV;
  auto sigma1 = [&](const std::vector<double>& x, std::vector<double>& y) {
    y.assign(n, 0.0);
    orc_sigma_csr(n, sp->rowptr.data(), sp->col.data(), sp->val.data(), 1, x.data(), y.data());
  };
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  };
  auto mgs_append = [&](std::vector<double> v) -> bool {
    for (size_t j = 0; j < V.size(); ++j) {
      double ov = -dot(v, V[j]);
      double nrm = std::sqrt(dot(V[j], V[j]));
      if (!(nrm > 1.0e-6)) return false;
      double proj = ov / nrm;
      for (long i = 0; i < n; ++i) v[i] += proj * V[j][i];
    }
    double nv = std::sqrt(dot(v, v));
    if (!(nv > 0.0)) return false;
    for (long i = 0; i < n; ++i) v[i] /= nv;
    V.push_back(v);
    return true;
  };
  int nsigma = 0;
  // guess codes as psoci_dav_params.guess: 1 unit vectors on the first k
  // SEFs (the reference's guess); 0 on the k lowest diagonal elements
  std::vector<long> start(k);
  for (int r = 0; r < k; ++r) start[r] = r;
  if (guess == 0) {
    std::vector<long> ord(n);
    for (long i = 0; i < n; ++i) ord[i] = i;
    std::partial_sort(ord.begin(), ord.begin() + k, ord.end(),
                      [&](long a, long b) { return D[a] < D[b] || (D[a] == D[b] && a < b); });
    for (int r = 0; r < k; ++r) start[r] = ord[r];
  }
  for (int r = 0; r < k; ++r) {
    std::vector<double> v(n, 0.0);
    v[start[r]] = 1.0;
    if (!mgs_append(v)) return -1;
  }
  for (int r = 0; r < k; ++r) {
    std::vector<double> y;
    sigma1(V[r], y);
    HV.push_back(y);
    ++nsigma;
  }
  std::vector<double> w, Y;
  std::vector<double> rn(k, -1.0);
  double coarse = 1.0e-3, cur = coarse;
  int curroot = 1, count = 0, total = 0;
  bool converged = false;
  auto solve = [&]() -> bool {
    int m = (int)V.size();
    std::vector<double> A(m * m), S(m * m);
    for (int a = 0; a < m; ++a)
      for (int b = 0; b <= a; ++b) {
        A[a * m + b] = A[b * m + a] = dot(V[a], HV[b]);
        S[a * m + b] = S[b * m + a] = dot(V[a], V[b]);
      }
    return small_gen_eig(m, A, S, w, Y);
  };
  auto residual = [&](int root, std::vector<double>& r) -> double {
    int m = (int)V.size();
    r.assign(n, 0.0);
    for (int s = 0; s < m; ++s) {
      double e = Y[s * m + root];
      for (long i = 0; i < n; ++i) r[i] += e * HV[s][i];
      double val = -1.0 * e * w[root];
      for (long i = 0; i < n; ++i) r[i] += val * V[s][i];
    }
    return std::sqrt(dot(r, r));
  };
  auto collapse = [&](int num) {
    int m = (int)V.size();
    std::vector<std::vector<double>> nV(num, std::vector<double>(n, 0.0)), nH(num, std::vector<double>(n, 0.0));
    for (int c = 0; c < num; ++c)
      for (int s = 0; s < m; ++s) {
        double e = Y[s * m + c];
        for (long i = 0; i < n; ++i) {
          nV[c][i] += e * V[s][i];
          nH[c][i] += e * HV[s][i];
        }
      }
    V.swap(nV);
    HV.swap(nH);
  };
  std::vector<double> r;
  while (count < maxit && !converged) {
    ++total;
    ++count;
    if (!solve()) return -2;
    converged = true;
    for (int it = curroot; it <= k; ++it) {
      double rv = residual(it - 1, r);
      rn[it - 1] = rv;
      if (rv > cur) {
        converged = false;
        curroot = it;
        break;
      }
    }
    if (!converged) {
      double shift = w[curroot - 1];
      residual(curroot - 1, r);
      for (long i = 0; i < n; ++i)
        if (std::fabs(D[i] - shift) > 1.0e-3) r[i] /= (D[i] - shift);
      if (!mgs_append(r)) return -3;
      std::vector<double> y;
      sigma1(V.back(), y);
      HV.push_back(y);
      ++nsigma;
    }
    if (converged && cur == coarse) {
      converged = false;
      count = 0;
      cur = rtol;
      curroot = 1;
      if ((int)V.size() >= minvecs * k) {
        solve();
        collapse(minvecs * k);
      }
    }
    if ((int)V.size() > maxsub) {
      solve();
      collapse(minvecs * k);
    }
    (void)cap;
  }
  solve();
  for (int i = 0; i < k; ++i) {
    rn[i] = residual(i, r);
    if (evals) evals[i] = w[i];
  }
  if (evecs) {
    int m = (int)V.size();
    for (long i = 0; i < n; ++i)
      for (int c = 0; c < k; ++c) {
        double s = 0.0;
        for (int t = 0; t < m; ++t) s += Y[t * m + c] * V[t][i];
        evecs[i * k + c] = s;
      }
  }
  if (rnorms_out)
    for (int i = 0; i < k; ++i) rnorms_out[i] = rn[i];
  if (nsigma_out) *nsigma_out = nsigma;
  return converged ? total : -total;
}

// CIDEN 1-RDM of one CI vector c (length nsef_total):
// constructCIdensityAltOrbMapDirectChunkJmap (PsociGAhamiltonian.C:11354-11784)
// with compareLocalJSubsetOrbMapCIDEN (J <= I, at most one electron moved),
// generateSOCIblocksForCIDEN and transformSEFSEFtoMOGET (:10027-10119).
// D is packed [NNDAF(nbf)] with 1-based twoDindex addressing (index-1).
void orc_rdm1(void* h, const double* c, double* D) {
  Space* sp = static_cast<Space*>(h);
  Ham hm = make_ham(*sp);
  long nd = nndaf(sp->nbf);
  std::fill(D, D + nd, 0.0);
  std::vector<double> blk;
  for (long I = 0; I < sp->ncfg; ++I) {
    Joutfg ti;
    const Joutfg& ic = conf_of(*sp, I, ti);
    for (long J = 0; J <= I; ++J) {
      if (!orbmap_connected(sp->occ[I], sp->occ[J], 1)) continue;
      Joutfg tj;
      const Joutfg& jc = conf_of(*sp, J, tj);
      int st = hm.block_ciden(ic, jc, blk);
      if (!st) continue;
      int nsefi = ic.nsefi, nsefj = jc.nsefi;
      const double* vi = c + sp->sef_off[I];
      const double* vj = c + sp->sef_off[J];
      if (I == J) {
        std::vector<std::pair<int, int>> molist;
        for (int o : ic.doubly) molist.push_back({o, 2});
        for (int o : ic.singly) molist.push_back({o, 1});
        for (auto& mo : molist) {
          double value = 0.0;
          for (int a = 0; a < nsefi; ++a) {
            double temp = vj[a];
            for (int b = 0; b < nsefj; ++b) value += blk[a * nsefj + b] * temp * vj[b] * mo.second;
          }
          D[two_e_index(mo.first, mo.first) - 1] += value;
        }
      } else {
        int iex = -1, jex = -1;
        for (int o = 0; o < sp->nbf; ++o) {
          int d = sp->occ[I][o] - sp->occ[J][o];
          if (d > 0) iex = o;
          if (d < 0) jex = o;
        }
        double value = 0.0;
        for (int a = 0; a < nsefi; ++a) {
          double temp = vi[a];
          for (int b = 0; b < nsefj; ++b) value += blk[a * nsefj + b] * temp * vj[b];
        }
        D[two_e_index(iex, jex) - 1] += value;
      }
    }
  }
}

// Direct sigma for selected rows: y_I = sum_J H_IJ x_J with every block
// regenerated (determinants of J recomputed per pair, as the reference's
// direct fetch does: computeConfigsGA per pair, PsociGAhamiltonian.C:2763;
// blocks by generateSOCIblocks in the reference orientation, J <= I, used
// transposed otherwise; matrixVectorUpdate semantics, :13036-13115).
// X is [n][k]; Y receives the rows of the listed configurations, packed in
// list order ([sum nsef][k]).  OpenMP over the listed rows.
int orc_direct_rows(void* h, long nrows, const long* rows, int k, const double* X, double* Y, int nthreads) {
  Space* sp = static_cast<Space*>(h);
  if (!sp->have_ints) return -1;
  Ham hm = make_ham(*sp);
  std::vector<long> yoff(nrows + 1, 0);
  for (long r = 0; r < nrows; ++r) yoff[r + 1] = yoff[r] + sp->nsef[rows[r]];
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
  for (long r = 0; r < nrows; ++r) {
    const long I = rows[r];
    double* y = Y + yoff[r] * k;
    std::fill(y, y + (long)sp->nsef[I] * k, 0.0);
    Joutfg ti;
    const Joutfg& ic = conf_of(*sp, I, ti);
    std::vector<double> blk;
    std::string err;
    for (long J = 0; J < sp->ncfg; ++J) {
      if (std::abs(sp->nopen[I] - sp->nopen[J]) > 4) continue;
      if (!orbmap_connected(sp->occ[I], sp->occ[J], 2)) continue;
      Joutfg tj;
      const Joutfg& jc = conf_of(*sp, J, tj);
      int ni = ic.nsefi, nj = jc.nsefi;
      const double* x = X + sp->sef_off[J] * k;
      if (J <= I) {
        int st = hm.block(ic, jc, blk, err);
        if (st < 0) bad = 1;
        if (st <= 0) continue;
        for (int i = 0; i < ni; ++i)
          for (int j = 0; j < nj; ++j) {
            double hv = blk[i * nj + j];
            for (int v = 0; v < k; ++v) y[i * k + v] += hv * x[j * k + v];
          }
      } else {
        int st = hm.block(jc, ic, blk, err);
        if (st < 0) bad = 1;
        if (st <= 0) continue;
        for (int j = 0; j < nj; ++j)
          for (int i = 0; i < ni; ++i) {
            double hv = blk[j * ni + i];
            for (int v = 0; v < k; ++v) y[i * k + v] += hv * x[j * k + v];
          }
      }
    }
  }
  return bad ? -2 : 0;
}

// CI energy analysis, analyzeEnergyContributions (PsociCIAnalysis.C:42-267),
// the per-configuration sums of :126-142 for every root: the reference reads
// root iroot of the GA vector into vect (dims[0] = maxsef long), takes its
// l_nsef offsets as each configuration's first SEF (the last configuration
// ending at maxsef) and accumulates in SEF order.  C is [n][k] row-major.
void orc_energy_contrib(long ncfg, const long* sef_off, long n, int k, const double* C, const double* diags,
                        const double* evals, double* deltae) {
  std::vector<double> vect(n);
  for (int iroot = 0; iroot < k; ++iroot) {
    for (long i = 0; i < n; ++i) vect[i] = C[i * k + iroot];
    for (long itotfg = 0; itotfg < ncfg; ++itotfg) {
      double acc = 0.0;
      long isefstart = sef_off[itotfg];
      long isefstop = itotfg + 1 < ncfg ? sef_off[itotfg + 1] : n;
      for (long i = isefstart; i < isefstop; ++i) {
        double vect2 = vect[i] * vect[i];
        if (vect2 == 1.0)
          acc = evals[iroot];
        else
          acc -= vect2 * (diags[i] - evals[iroot]) / (1.0 - vect2);
      }
      deltae[itotfg * k + iroot] = acc;
    }
  }
}

// Canonical packed indices, exported so tests can check them against brute
// force permutation symmetry.
int orc_two_e_index(int i, int j) { return two_e_index(i, j); }
int orc_canonical_index(int i, int j, int k, int l) { return canonical_index(i, j, k, l); }

}  // extern "C"
