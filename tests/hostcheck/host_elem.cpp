// Test harness: the product's space builder and element arithmetic compiled
// for the host CPU, so the bit-mask formulation can be compared with the
// oracle on machines without a GPU.  Not part of the product library.
#include <cstring>
#include <vector>

#include "../../psoci_amd/csrc/common.hpp"
#include "../../psoci_amd/csrc/elem.hpp"

using namespace psoci;

extern "C" {

void* hc_space_new(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym) {
  Space* sp = new Space;
  try {
    build_space(*sp, nbf, nelec, ksym, ncfg, occ, jsym);
  } catch (...) {
    delete sp;
    return nullptr;
  }
  return sp;
}

void hc_space_free(void* h) { delete static_cast<Space*>(h); }

static ConfView view(const Space& sp, int64_t c) {
  ConfView v;
  v.d = sp.dm[c];
  v.s = sp.sm[c];
  v.nopen = sp.nopen[c];
  v.ph = sp.cphase[c];
  v.pat = sp.pat.data() + sp.pat_off[v.nopen];
  return v;
}

// block (I, J) without threshold; returns the block status
int hc_block(void* h, const double* one, const double* two, int64_t I, int64_t J, double* out) {
  const Space& sp = *static_cast<Space*>(h);
  PairInfo P;
  pair_setup(P, (int)I, (int)J, sp.dm[I], sp.sm[I], sp.nopen[I], sp.dm[J], sp.sm[J], sp.nopen[J], sp.nbf, one, two);
  ConfView A = view(sp, I), B = view(sp, J);
  int ni = sp.nsef[I], nj = sp.nsef[J];
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j) out[i * nj + j] = P.kind ? sef_elem(P, A, B, i, j, sp.mult, sp.mult, two) : 0.0;
  return P.kind ? 1 : 0;
}

// block (I, J) through the load-free off-diagonal path (sef_elem_x);
// returns the block status, or -1 for a diagonal pair (not covered)
int hc_block_x(void* h, const double* one, const double* two, int64_t I, int64_t J, double* out) {
  const Space& sp = *static_cast<Space*>(h);
  if (I == J) return -1;
  PairInfo P;
  pair_setup(P, (int)I, (int)J, sp.dm[I], sp.sm[I], sp.nopen[I], sp.dm[J], sp.sm[J], sp.nopen[J], sp.nbf, one, two);
  ConfView A = view(sp, I), B = view(sp, J);
  PairX px;
  pair_x(P, A, two, px);
  int ni = sp.nsef[I], nj = sp.nsef[J];
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j) {
      double v = 0.0;
      if (P.kind) {
        uint64_t aI0 = deposit(A.pat[A.nopen ? 2 * i : 0], A.s), aI1 = A.nopen ? deposit(A.pat[2 * i + 1], A.s) : 0;
        uint64_t aJ0 = deposit(B.pat[B.nopen ? 2 * j : 0], B.s), aJ1 = B.nopen ? deposit(B.pat[2 * j + 1], B.s) : 0;
        v = sef_elem_x(P, px, A, B, aI0, aI1, aJ0, aJ1, sp.mult, sp.mult, two);
      }
      out[i * nj + j] = v;
    }
  return P.kind ? 1 : 0;
}

int hc_connected(void* h, int64_t I, int64_t J) {
  const Space& sp = *static_cast<Space*>(h);
  int dn = sp.nopen[I] - sp.nopen[J];
  if (dn > 4 || dn < -4) return 0;
  return moved(sp.dm[I], sp.sm[I], sp.dm[J], sp.sm[J]) <= 2;
}

}
