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
  v.pat = sp.pat.data() + sp.pat_off[12 * sp.cls[c] + v.nopen];
  return v;
}

// block (I, J) without threshold; returns the block status
int hc_block(void* h, const double* one, const double* two, int64_t I, int64_t J, double* out) {
  const Space& sp = *static_cast<Space*>(h);
  PairInfo P;
  pair_setup(P, (int)I, (int)J, sp.dm[I], sp.sm[I], sp.nopen[I], sp.dm[J], sp.sm[J], sp.nopen[J], sp.nbf, one, two,
             sp.jsym[I], sp.jsym[J]);
  ConfView A = view(sp, I), B = view(sp, J);
  int ni = sp.nsef[I], nj = sp.nsef[J];
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j)
      out[i * nj + j] = P.kind ? sef_elem(P, A, B, i, j, sp.cmult[I], sp.cmult[J], two, sp.frac) : 0.0;
  return P.kind ? 1 : 0;
}

// block kind of (I, J) (4: spin-orbit single)
int hc_kind(void* h, const double* one, const double* two, int64_t I, int64_t J) {
  const Space& sp = *static_cast<Space*>(h);
  PairInfo P;
  pair_setup(P, (int)I, (int)J, sp.dm[I], sp.sm[I], sp.nopen[I], sp.dm[J], sp.sm[J], sp.nopen[J], sp.nbf, one, two,
             sp.jsym[I], sp.jsym[J]);
  return P.kind;
}

// block (I, J) through the three-index-table off-diagonal path (sef_elem_x);
// returns the block status, or -1 for a diagonal pair (not covered)
int hc_block_x(void* h, const double* one, const double* two, int64_t I, int64_t J, double* out) {
  const Space& sp = *static_cast<Space*>(h);
  if (I == J) return -1;
  PairInfo P;
  pair_setup(P, (int)I, (int)J, sp.dm[I], sp.sm[I], sp.nopen[I], sp.dm[J], sp.sm[J], sp.nopen[J], sp.nbf, one, two);
  ConfView A = view(sp, I), B = view(sp, J);
  std::vector<double> xt((size_t)sp.nbf * sp.nbf * sp.nbf);
  make_xt(sp.nbf, two, xt.data());
  const double* xrow = xt.data() + ((int64_t)(P.kind == 2 ? P.e0 : 0) * sp.nbf + (P.kind == 2 ? P.f0 : 0)) * sp.nbf;
  int ni = sp.nsef[I], nj = sp.nsef[J];
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j) {
      double v = 0.0;
      if (P.kind) {
        const int si = sp.frac ? i : (A.nopen ? 2 * i : 0), sj = sp.frac ? j : (B.nopen ? 2 * j : 0);
        uint64_t aI0 = deposit(A.pat[si], A.s), aI1 = (!sp.frac && A.nopen) ? deposit(A.pat[2 * i + 1], A.s) : 0;
        uint64_t aJ0 = deposit(B.pat[sj], B.s), aJ1 = (!sp.frac && B.nopen) ? deposit(B.pat[2 * j + 1], B.s) : 0;
        v = sef_elem_x(P, xrow, A, B, aI0, aI1, aJ0, aJ1, sp.mult, sp.mult, two, sp.frac);
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

#include "../../psoci_amd/csrc/detdrive.hpp"

extern "C" {

// sigma row block y_I = sum_J H_IJ x_J (k columns, row-major) two ways:
// mode 0 SEF pair by SEF pair with sef_elem (reference orientation), mode 1
// determinant-driven (for_each_conn); every J within two moved electrons
int hc_row(void* h, const double* one, const double* two, int64_t I, int k, const double* X, double* y, int mode) {
  const Space& sp = *static_cast<Space*>(h);
  const ConfView A = view(sp, I);
  const int ni = sp.nsef[I];
  for (int t = 0; t < ni * k; ++t) y[t] = 0.0;
  std::vector<double> xt((size_t)sp.nbf * sp.nbf * sp.nbf);
  make_xt(sp.nbf, two, xt.data());
  for (int64_t J = 0; J < sp.ncfg; ++J) {
    int dn = sp.nopen[I] - sp.nopen[J];
    if (dn > 4 || dn < -4) continue;
    if (moved(sp.dm[I], sp.sm[I], sp.dm[J], sp.sm[J]) > 2) continue;
    const bool rowI = I >= J;
    const int64_t R = rowI ? I : J, Cc = rowI ? J : I;
    PairInfo P;
    pair_setup(P, (int)R, (int)Cc, sp.dm[R], sp.sm[R], sp.nopen[R], sp.dm[Cc], sp.sm[Cc], sp.nopen[Cc], sp.nbf, one,
               two);
    if (!P.kind) continue;
    const ConfView B = view(sp, J);
    const int nj = sp.nsef[J];
    const double* xJ = X + sp.sef_off[J] * k;
    if (mode == 0) {
      const ConfView RA = view(sp, R), CB = view(sp, Cc);
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
          double v = rowI ? sef_elem(P, RA, CB, i, j, sp.mult, sp.mult, two, sp.frac)
                          : sef_elem(P, RA, CB, j, i, sp.mult, sp.mult, two, sp.frac);
          for (int c = 0; c < k; ++c) y[i * k + c] += v * xJ[j * k + c];
        }
    } else {
      // the direct sigma's record rule: in an ODD-class space a closed shell
      // (Ms = 0) and an open-shell configuration (odd Ms) never connect
      if (sp.odd && ((A.nopen == 0) != (B.nopen == 0))) continue;
      const double* xrow = xt.data() + ((int64_t)(P.kind == 2 ? P.e0 : 0) * sp.nbf + (P.kind == 2 ? P.f0 : 0)) * sp.nbf;
      const PairCoef pc = pair_coef(A.nopen, B.nopen, sp.mult, sp.frac);
      const int16_t* dj = sp.didx.data() + (1 << B.nopen) - 1;
      const int sh = sp.frac ? 0 : 1;
      for (int i = 0; i < ni; ++i) {
        const uint64_t aa = A.nopen ? deposit(A.pat[i << sh], A.s) : 0ull;
        auto add = [&](int b, uint64_t ab) {
          if (b < 0) throw Error(PSOCI_ERR_NUMERIC, "determinant outside the class");
          const double v = P.kind == 1 ? detval_ao(P, A, A, aa, ab, two) : conn_value(P, xrow, rowI, A, B, aa, ab, two);
          const double cf = (b & 1) ? pc.co : pc.ce;
          const double hv = cf * v;
          for (int c = 0; c < k; ++c) y[i * k + c] += hv * xJ[(b >> sh) * k + c];
        };
        if (I == J)
          for_each_conn_diag(A, aa, dj, add);
        else
          for_each_conn(A, B, aa, dj, add);
      }
    }
  }
  return 0;
}

}

#include "../../psoci_amd/csrc/smalleig.hpp"

extern "C" int hc_sym_eig(int m, double* a, double* w) { return sym_eig(m, a, w) ? 1 : 0; }
