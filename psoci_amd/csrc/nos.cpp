// Natural orbitals and Mulliken populations from the CI one-particle density:
// PsociNOpopulations::constructNaturalOrbitals and its helpers
// (PsociNOpopulations.C:173-500, 569-729) and the Mulliken analysis
// (constructMullikenPopulationDensAO, clusterBasisLabels, atomicPopulations,
// subshellPopulations, atomicOverlapPopulations: :1084-1206, 1451-1790).
// Host code on nbf x nbf matrices (nbf <= 64): one symmetry block, as in C1.
//
// Conventions kept from the reference:
//  * the density is packed like valone, D[max(p,q)(max(p,q)+1)/2 + min(p,q)],
//    with the off-diagonal counted once per orbital pair (constructDensityMatrix;
//    psoci_rdm's grad_one is D_pp on the diagonal and 2 D_pq off it);
//  * MO coefficients are [MO][AO] (mocoef[k * nbf + j]: AO j of MO k);
//  * natural orbitals come out sorted by occupation, largest first, with
//    row i of the [NO][MO] matrix the i-th natural orbital
//    (sortNOinMObasis; the reference's dsygv returns them as columns);
//  * the AO density is C^T D C, packed (computeMOdensity), built by the same
//    two half transforms in the same summation order;
//  * Mulliken population matrix P_{mu nu} S_{mu nu}; atomic / subshell labels
//    are the basis labels with leading digits and blanks removed, cut at the
//    first '_' for atoms (clusterBasisLabels); groups are numbered in the
//    sorted order of their labels, as the reference's multimap visits them.
//
// The eigen-decomposition is the library's Householder + QR solver; an
// eigenvector's sign is fixed so that its largest-magnitude MO coefficient
// is positive (the reference leaves it to LAPACK).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "smalleig.hpp"

namespace psoci {
namespace {

inline size_t tri(int p, int q) {
  const size_t a = (size_t)std::max(p, q), b = (size_t)std::min(p, q);
  return a * (a + 1) / 2 + b;
}

void natural_orbitals(int nbf, const double* rdm1, const double* mocoef, double* occ, double* no_mo, double* no_ao,
                      double* dens_ao) {
  PSOCI_CHECK(nbf >= 1 && nbf <= 64 && rdm1, PSOCI_ERR_ARG, "bad natural-orbital arguments");
  PSOCI_CHECK(mocoef || (!no_ao && !dens_ao), PSOCI_ERR_ARG, "AO results need the MO coefficients");
  const int n = nbf;
  std::vector<double> d((size_t)n * n), a((size_t)n * n), w(n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) d[(size_t)i * n + j] = rdm1[tri(i, j)];
  a = d;
  PSOCI_CHECK(sym_eig(n, a.data(), w.data()), PSOCI_ERR_NUMERIC, "density matrix diagonalisation failed");
  // descending occupations; equal ones keep the solver's ascending order
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return w[x] > w[y]; });
  std::vector<double> U((size_t)n * n);  // U[i][k]: MO k of NO i
  for (int i = 0; i < n; ++i) {
    const int c = ord[i];
    int big = 0;
    for (int k = 1; k < n; ++k)
      if (std::fabs(a[(size_t)k * n + c]) > std::fabs(a[(size_t)big * n + c])) big = k;
    const double sgn = a[(size_t)big * n + c] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < n; ++k) U[(size_t)i * n + k] = sgn * a[(size_t)k * n + c];
    if (occ) occ[i] = w[c];
  }
  if (no_mo) std::copy(U.begin(), U.end(), no_mo);
  if (no_ao) {
    // transformNOstoAObasis: row i = sum_k U[i][k] mocoef[k][:]
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += mocoef[(size_t)k * n + j] * U[(size_t)i * n + k];
        no_ao[(size_t)i * n + j] = s;
      }
  }
  if (dens_ao) {
    // computeMOdensity: scratch = D C, then densAO[i][j] (j <= i) = sum_k scratch[k][i] C[k][j]
    std::vector<double> sc((size_t)n * n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double t = 0.0;
        for (int k = 0; k < n; ++k) t += d[(size_t)i * n + k] * mocoef[(size_t)k * n + j];
        sc[(size_t)i * n + j] = t;
      }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        double t = 0.0;
        for (int k = 0; k < n; ++k) t += sc[(size_t)k * n + i] * mocoef[(size_t)k * n + j];
        dens_ao[tri(i, j)] = t;
      }
  }
}

void mulliken(int nbf, const double* dens_ao, const double* overlap, double* population, double* total) {
  PSOCI_CHECK(nbf >= 1 && nbf <= 64 && dens_ao && overlap, PSOCI_ERR_ARG, "bad Mulliken arguments");
  const int n = nbf;
  double charge = 0.0;
  std::vector<double> pop((size_t)n * n);
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu < n; ++nu) pop[(size_t)nu * n + mu] = dens_ao[tri(mu, nu)] * overlap[tri(nu, mu)];
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu < n; ++nu) charge += pop[(size_t)nu * n + mu];
  if (population) std::copy(pop.begin(), pop.end(), population);
  if (total) *total = charge;
}

std::string condense(const char* lab, int mode) {
  std::string word(lab, lab + 8);
  if (mode == 2) return word;
  size_t start = 0;
  while (start < word.size() && (std::isdigit((unsigned char)word[start]) || !std::isalnum((unsigned char)word[start])))
    ++start;
  if (mode == 1) return word.substr(start);
  const size_t us = word.find('_');
  return us == std::string::npos || us < start ? word.substr(start) : word.substr(start, us - start);
}

void basis_groups(int nbf, const char* labels, int mode, int32_t* group, int32_t* ngroups) {
  PSOCI_CHECK(nbf >= 1 && nbf <= 64 && labels && group && (mode >= 0 && mode <= 2), PSOCI_ERR_ARG,
              "bad basis-label arguments");
  std::vector<std::string> keys(nbf);
  std::map<std::string, int> ids;
  for (int i = 0; i < nbf; ++i) {
    keys[i] = condense(labels + 8 * (size_t)i, mode);
    ids[keys[i]] = 0;
  }
  int g = 0;
  for (auto& kv : ids) kv.second = g++;
  for (int i = 0; i < nbf; ++i) group[i] = ids[keys[i]];
  if (ngroups) *ngroups = g;
}

void group_populations(int nbf, const double* pop, const int32_t* group, int ngroups, double* gross, double* bond) {
  PSOCI_CHECK(nbf >= 1 && nbf <= 64 && pop && group && ngroups >= 1, PSOCI_ERR_ARG, "bad group-population arguments");
  for (int i = 0; i < nbf; ++i)
    PSOCI_CHECK(group[i] >= 0 && group[i] < ngroups, PSOCI_ERR_ARG, "group index out of range");
  const int n = nbf;
  // the reference's multimap order: by group, then basis-function index
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return group[x] < group[y]; });
  if (gross) {
    for (int g = 0; g < ngroups; ++g) {
      double f = 0.0;
      for (int i : ord) {
        if (group[i] != g) continue;
        for (int j : ord) f += pop[(size_t)j * n + i];
      }
      gross[g] = f;
    }
  }
  if (bond) {
    for (int g = 0; g < ngroups; ++g)
      for (int h = 0; h < ngroups; ++h) {
        double f = 0.0;
        if (h >= g) {
          const double factor = g == h ? 1.0 : 2.0;
          for (int i : ord) {
            if (group[i] != g) continue;
            for (int j : ord)
              if (group[j] == h) f += factor * pop[(size_t)i * n + j];
          }
        }
        bond[(size_t)g * ngroups + h] = f;
      }
  }
}

}  // namespace
}  // namespace psoci

extern "C" {

int psoci_natural_orbitals(int nbf, const double* rdm1, const double* mocoef, double* occ, double* no_mo,
                           double* no_ao, double* dens_ao) {
  return psoci::ctxfree([&] { psoci::natural_orbitals(nbf, rdm1, mocoef, occ, no_mo, no_ao, dens_ao); });
}

int psoci_mulliken(int nbf, const double* dens_ao, const double* overlap, double* population, double* total) {
  return psoci::ctxfree([&] { psoci::mulliken(nbf, dens_ao, overlap, population, total); });
}

int psoci_basis_groups(int nbf, const char* labels, int mode, int32_t* group, int32_t* ngroups) {
  return psoci::ctxfree([&] { psoci::basis_groups(nbf, labels, mode, group, ngroups); });
}

int psoci_group_populations(int nbf, const double* population, const int32_t* group, int ngroups, double* gross,
                            double* bond) {
  return psoci::ctxfree([&] { psoci::group_populations(nbf, population, group, ngroups, gross, bond); });
}

}  // extern "C"
