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

int hc_connected(void* h, int64_t I, int64_t J) {
  const Space& sp = *static_cast<Space*>(h);
  int dn = sp.nopen[I] - sp.nopen[J];
  if (dn > 4 || dn < -4) return 0;
  return moved(sp.dm[I], sp.sm[I], sp.dm[J], sp.sm[J]) <= 2;
}

}
