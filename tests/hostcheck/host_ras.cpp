// Test harness: the string-driven direct sigma's host plan (ras.cpp) and the
// index arithmetic its GPU kernels use (ras.hpp: ras_d_entry, ras_terms),
// evaluated with plain loops on the host CPU, so the tables, signs and SEF
// map can be compared with the oracle on machines without a GPU.  Not part
// of the product library.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../psoci_amd/csrc/common.hpp"
#include "../../psoci_amd/csrc/ras.hpp"

using namespace psoci;

extern "C" {

// Plan summary: returns 1 when the space qualifies (0 otherwise, with the
// reason in why[256]); info[0..5] = m, external orbital count, sectors,
// blocks, determinants, W elements per column.
int hr_plan(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym, const double* one,
            const double* two, int64_t* info, char* why) {
  Space sp;
  RasPlan P;
  try {
    build_space(sp, nbf, nelec, ksym, ncfg, occ, jsym);
    ras_plan(P, sp, one, two);
  } catch (const Error& e) {
    std::strncpy(why, e.what(), 255);
    return -1;
  }
  std::strncpy(why, P.why.c_str(), 255);
  if (!P.ok) return 0;
  info[0] = P.m;
  info[1] = popc(P.emask);
  info[2] = (int64_t)P.sec.size();
  info[3] = (int64_t)P.blk.size();
  info[4] = P.ndet;
  info[5] = P.wunit;
  return 1;
}

static int hr_sigma_impl(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym,
                         const double* one, const double* two, int k, const double* X, double* Y, int merged,
                         int64_t* nterm);

// Y[n][k] = H X[n][k] by the plan, on the host; returns 0 on success.
int hr_sigma(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym, const double* one,
             const double* two, int k, const double* X, double* Y) {
  return hr_sigma_impl(nbf, nelec, ksym, ncfg, occ, jsym, one, two, k, X, Y, 0, nullptr);
}

// The same with the SEF term lists the GPU gather uses (ras_sef_terms: the
// mirrored terms of an SEF's two determinants merged) and the GPU product's
// chunk skipping; nterm[0] receives the number of terms over all SEFs,
// nterm[1] the determinants' terms unmerged, nterm[2] the (row, chunk)
// pairs the skipping left out.
int hr_sigma_merged(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym,
                    const double* one, const double* two, int k, const double* X, double* Y, int64_t* nterm) {
  return hr_sigma_impl(nbf, nelec, ksym, ncfg, occ, jsym, one, two, k, X, Y, 1, nterm);
}

}  // extern "C"

static int hr_sigma_impl(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym,
                         const double* one, const double* two, int k, const double* X, double* Y, int merged,
                         int64_t* nterm) {
  Space sp;
  RasPlan P;
  try {
    build_space(sp, nbf, nelec, ksym, ncfg, occ, jsym);
    ras_plan(P, sp, one, two);
  } catch (const Error&) {
    return -1;
  }
  if (!P.ok) return 1;
  const RasView v = ras_host_view(P);
  const double r = 0x1.6a09e667f3bcdp-1;
  auto coef = [&](uint8_t f) { return ((f & 2) ? 1.0 : r) * ((f & 1) ? -1.0 : 1.0); };
  std::vector<double> c((size_t)P.ndet * k, 0.0);
  std::vector<char> seen(P.ndet, 0);
  for (int64_t s = 0; s < v.nsef; ++s) {
    const RasSef& e = v.sef[s];
    for (int t = 0; t < e.n; ++t) {
      const RasSectorDev& S = v.sec[e.sec[t]];
      const int ca = ras_cls_of(v, S.na, e.a[t]), cb = ras_cls_of(v, S.nb, e.b[t]);
      const int64_t d = ras_det(v, e.sec[t], ca, e.a[t] - v.lev[S.na].cls_off[ca], cb, e.b[t] - v.lev[S.nb].cls_off[cb]);
      if (d < 0 || d >= P.ndet || seen[d]) return 2;
      seen[d] = 1;
      for (int j = 0; j < k; ++j) c[d * k + j] = coef(e.flag[t]) * X[s * k + j];
    }
  }
  for (int64_t d = 0; d < P.ndet; ++d)
    if (!seen[d]) return 3;
  std::vector<double> W((size_t)P.wunit * k, 0.0);
  std::vector<double> D;
  int64_t skipped = 0;  // (row, chunk) pairs left out by the chunk skipping
  for (int bi = 0; bi < v.nblk; ++bi) {
    const RasBlockDev& B = v.blk[bi];
    const double* G = P.G.data() + B.goff;
    D.assign((size_t)B.nqp * k, 0.0);
    for (int64_t row = 0; row < B.nrows; ++row) {
      // with merged = 1, the GPU product's chunk skipping: in a tile of 8
      // rows (16 columns) whose rows share their first string u, a chunk of
      // 16 pairs whose first-added orbitals u all holds is left out
      uint64_t live[2] = {~0ull, ~0ull};
      if (merged && B.kind != 2) {
        const int64_t RT = 8, r0 = row / RT * RT, r1 = std::min(r0 + RT, B.nrows) - 1;
        int64_t ua, ub, wd;
        if (B.tri) {
          ras_tri_uw(r0, B.nu, ua, wd);
          ras_tri_uw(r1, B.nu, ub, wd);
        } else {
          ua = r0 / B.nw;
          ub = r1 / B.nw;
        }
        if (ua == ub) {
          const uint64_t um = P.mask[B.umoff + B.u0 + ua];
          for (int ch = 0; ch < B.nqp / 16 && ch < 128; ++ch)
            if ((P.qm[B.qmoff + ch] & ~um) == 0) live[ch >> 6] &= ~(1ull << (ch & 63));
        }
      }
      for (int j = 0; j < B.nqp; ++j) {
        const int ch = j / 16;
        if (ch < 128 && !((live[ch >> 6] >> (ch & 63)) & 1)) {
          for (int q = 0; q < k; ++q) D[(size_t)j * k + q] = 0.0;
          if (j % 16 == 0) ++skipped;
          continue;
        }
        int neg = 0;
        const int64_t d = ras_d_entry(v, B, row, j, neg);
        if (d >= P.ndet) return 4;
        for (int q = 0; q < k; ++q) D[(size_t)j * k + q] = d < 0 ? 0.0 : (neg ? -c[d * k + q] : c[d * k + q]);
      }
      for (int i = 0; i < B.nqp; ++i)
        for (int q = 0; q < k; ++q) {
          double acc = 0.0;
          for (int j = 0; j < B.nqp; ++j) acc += G[(int64_t)i * B.nqp + j] * D[(size_t)j * k + q];
          W[(B.wbase + row * B.nqp + i) * k + q] = acc;
        }
    }
  }
  if (merged) {
    int64_t nt = 0, nt1 = 0;
    const double sc[3] = {r, 1.0 / r, 1.0};
    for (int64_t s = 0; s < v.nsef; ++s) {
      const RasSef& e = v.sef[s];
      std::vector<double> sig(k, 0.0);
      int bad = 0;
      const int code = ras_sef_terms(v, e, &bad, [&](int64_t off, int neg) {
        ++nt;
        for (int q = 0; q < k; ++q) sig[q] += neg ? -W[off * k + q] : W[off * k + q];
      });
      if (bad) return 5;
      for (int q = 0; q < k; ++q) Y[s * k + q] = sc[code] * sig[q];
      for (int t = 0; t < e.n; ++t) ras_terms(v, e.sec[t], e.a[t], e.b[t], [&](int, int64_t, int) { ++nt1; });
    }
    if (nterm) {
      nterm[0] = nt;
      nterm[1] = nt1;
      nterm[2] = skipped;
    }
    return 0;
  }
  for (int64_t s = 0; s < v.nsef; ++s) {
    const RasSef& e = v.sef[s];
    for (int q = 0; q < k; ++q) Y[s * k + q] = 0.0;
    for (int t = 0; t < e.n; ++t) {
      const double w = coef(e.flag[t]) * (e.sec[t] > 0 ? 2.0 : 1.0);
      std::vector<double> sig(k, 0.0);
      bool bad = false;
      ras_terms(v, e.sec[t], e.a[t], e.b[t], [&](int bi, int64_t off, int neg) {
        if (bi < 0 || bi >= v.nblk || off < 0 || off >= P.wunit) {
          bad = true;
          return;
        }
        for (int q = 0; q < k; ++q) sig[q] += neg ? -W[off * k + q] : W[off * k + q];
      });
      if (bad) return 5;
      for (int q = 0; q < k; ++q) Y[s * k + q] += w * sig[q];
    }
  }
  return 0;
}

extern "C" {
// Per-block summary of the plan: rows, nq, nqp, kind, sector, classes (9
// int64 per block); returns the block count (at most cap).
int hr_blocks(int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym, const double* one,
              const double* two, int64_t* out, int cap) {
  Space sp;
  RasPlan P;
  try {
    build_space(sp, nbf, nelec, ksym, ncfg, occ, jsym);
    ras_plan(P, sp, one, two);
  } catch (const Error&) {
    return -1;
  }
  if (!P.ok) return 0;
  int nb = 0;
  for (const RasBlockDev& B : P.blk) {
    if (nb >= cap) break;
    int64_t* o = out + 9 * nb++;
    o[0] = B.nrows; o[1] = B.nq; o[2] = B.nqp; o[3] = B.kind; o[4] = B.sector; o[5] = B.cu; o[6] = B.cw;
    o[7] = B.tri; o[8] = B.budget;
  }
  return nb;
}
}
