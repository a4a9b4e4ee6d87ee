// Host plan of the string-driven direct sigma (ras.hpp): detection of a
// complete RAS configuration space, string levels with their addition and
// removal tables, spin sectors, the dense products (blocks) with their pair
// lists and integral matrices, and the SEF -> determinant map.
#include "ras.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "common.hpp"

namespace psoci {

namespace {

// number of 0/1/2 occupations of nel electrons over norb orbitals
double occ_count(int norb, int nel) {
  double t = 0.0;
  for (int d = 0; 2 * d <= nel; ++d) {
    const int s = nel - 2 * d;
    if (d > norb || s > norb - d) continue;
    double a = 1.0, b = 1.0;
    for (int i = 0; i < d; ++i) a = a * (norb - i) / (i + 1);
    for (int i = 0; i < s; ++i) b = b * (norb - d - i) / (i + 1);
    t += a * b;
  }
  return t;
}

struct Binom {
  int64_t c[65][65];
  Binom() {
    std::memset(c, 0, sizeof(c));
    for (int n = 0; n <= 64; ++n) {
      c[n][0] = 1;
      for (int k = 1; k <= n; ++k) c[n][k] = c[n - 1][k - 1] + c[n - 1][k];
    }
  }
  int64_t operator()(int n, int k) const { return (k < 0 || n < 0 || k > n) ? 0 : c[n][k]; }
};
const Binom& binom() {
  static Binom b;
  return b;
}

// The RAS description: the external orbitals E (orbitals ordered by how many
// configurations occupy them; the least occupied form E) and the largest
// external electron count m, such that the list holds exactly every
// occupation with at most m electrons in E.  False when no split fits.
bool detect(const Space& sp, uint64_t& emask, int& m) {
  const int nbf = sp.nbf;
  const int64_t ncfg = sp.ncfg;
  std::vector<std::pair<uint64_t, uint64_t>> keys(ncfg);
  for (int64_t c = 0; c < ncfg; ++c) keys[c] = {sp.dm[c], sp.sm[c]};
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return false;  // repeated configurations
  std::vector<int64_t> cnt(nbf, 0);
  for (int64_t c = 0; c < ncfg; ++c)
    for (uint64_t x = sp.dm[c] | sp.sm[c]; x; x &= x - 1) ++cnt[ctz(x)];
  std::vector<int> order(nbf);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cnt[x] > cnt[y]; });
  for (int t = nbf; t >= 1; --t) {
    uint64_t e = 0;
    for (int i = t; i < nbf; ++i) e |= 1ull << order[i];
    int mm = 0;
    for (int64_t c = 0; c < ncfg; ++c) mm = std::max(mm, 2 * popc(sp.dm[c] & e) + popc(sp.sm[c] & e));
    double expect = 0.0;
    for (int n = 0; n <= mm; ++n) expect += occ_count(t, sp.nelec - n) * occ_count(nbf - t, n);
    if (expect == (double)ncfg) {
      emask = e;
      m = mm;
      return true;
    }
  }
  return false;
}

struct Ranker {
  int nI = 0, nE = 0;
  int pos[64];
  bool ext[64];
  void init(int nbf, uint64_t emask) {
    nI = nE = 0;
    for (int o = 0; o < nbf; ++o) {
      ext[o] = (emask >> o) & 1;
      pos[o] = ext[o] ? nE++ : nI++;
    }
  }
  // (class, rank within the class) of a string
  void rank(uint64_t s, int& c, int64_t& r) const {
    const Binom& B = binom();
    int64_t ri = 0, re = 0;
    int ki = 0, ke = 0;
    for (uint64_t x = s; x; x &= x - 1) {
      const int o = ctz(x);
      if (ext[o]) re += B(pos[o], ++ke);
      else ri += B(pos[o], ++ki);
    }
    c = ke;
    r = ri * B(nE, ke) + re;
  }
};

}  // namespace

void ras_plan(RasPlan& P, const Space& sp, const double* one, const double* two) {
  P = RasPlan();
  const int nbf = sp.nbf, N = sp.nelec;
  auto fail = [&](const char* why) {
    P.ok = false;
    P.why = why;
  };
  if (sp.frac) return fail("odd electron count (one determinant per SEF)");
  if (sp.odd) return fail("ODD determinant class");
  if (sp.mixed) return fail("configurations of more than one jsym");
  if (N < 2 || N > kRasMaxN) return fail("electron count outside 2..32");
  uint64_t emask = 0;
  int m = 0;
  if (!detect(sp, emask, m)) return fail("not a complete RAS configuration space");
  if (m >= kRasMaxCls) return fail("more than 6 external electrons");
  P.nbf = nbf;
  P.N = N;
  P.m = m;
  P.emask = emask;
  P.mult = sp.mult;
  Ranker rk;
  rk.init(nbf, emask);
  const Binom& Bn = binom();

  // sectors n_alpha - n_beta = 0, 4, 8, ... (the EVEN class's Ms = 0, 2, 4, ...)
  std::vector<std::pair<int, int>> secs;
  for (int dn = 0; dn <= N; dn += 4) secs.push_back({(N + dn) / 2, (N - dn) / 2});
  // levels: strings of n electrons with at most m external electrons
  std::vector<bool> need(N + 1, false);
  for (auto& s : secs)
    for (int d = 0; d <= 2; ++d) {
      if (s.first - d >= 0) need[s.first - d] = true;
      if (s.second - d >= 0) need[s.second - d] = true;
    }
  P.lev.assign(kRasMaxN + 1, RasLevelDev{});
  std::vector<int64_t> lev_total(N + 2, 0);
  const int64_t kMaxStrings = (int64_t)1 << 27;
  for (int n = 0; n <= N; ++n) {
    if (!need[n]) continue;
    RasLevelDev& L = P.lev[n];
    L.mask_off = (int64_t)P.mask.size();
    int64_t g = 0;
    for (int a = 0; a < kRasMaxCls; ++a) {
      L.cls_off[a] = g;
      if (a > m || a > n) continue;
      const int64_t ni = Bn(rk.nI, n - a), ne = Bn(rk.nE, a);
      if (ni == 0 || ne == 0) continue;
      if (ni * ne >= kMaxStrings || g + ni * ne >= kMaxStrings) return fail("too many strings");
      // rank order: internal colex rank major, external colex rank minor
      std::vector<int> iorb, eorb;
      for (int o = 0; o < nbf; ++o) (rk.ext[o] ? eorb : iorb).push_back(o);
      auto unrank = [&](int64_t r, int k, const std::vector<int>& orb) {
        uint64_t s = 0;
        for (int i = k; i >= 1; --i) {
          int c = i - 1;
          while (Bn(c + 1, i) <= r) ++c;
          r -= Bn(c, i);
          s |= 1ull << orb[c];
        }
        return s;
      };
      for (int64_t ri = 0; ri < ni; ++ri) {
        const uint64_t si = unrank(ri, n - a, iorb);
        for (int64_t re = 0; re < ne; ++re) {
          P.mask.push_back(si | unrank(re, a, eorb));
          P.cls.push_back((int8_t)a);
        }
      }
      g += ni * ne;
    }
    L.cls_off[kRasMaxCls] = g;
    lev_total[n] = g;
  }
  // addition tables (level n -> n + 1) and removal tables (n -> n - 1)
  for (int n = 0; n <= N; ++n) {
    if (!need[n]) continue;
    RasLevelDev& L = P.lev[n];
    const int64_t cnt = lev_total[n];
    if (n + 1 <= N && need[n + 1]) {
      L.add_off = (int64_t)P.add.size();
      P.add.resize(P.add.size() + cnt * nbf, kRasNone);
      for (int64_t g = 0; g < cnt; ++g) {
        const uint64_t s = P.mask[L.mask_off + g];
        for (int q = 0; q < nbf; ++q) {
          if ((s >> q) & 1) continue;
          const uint64_t t = s | (1ull << q);
          if (popc(t & emask) > m) continue;
          int c;
          int64_t r;
          rk.rank(t, c, r);
          const int neg = popc(s & ((1ull << q) - 1ull)) & 1;
          P.add[L.add_off + g * nbf + q] = ras_pack(neg, c, r);
        }
      }
    } else {
      L.add_off = -1;
    }
    if (n >= 1 && need[n - 1]) {
      L.rem_off = (int64_t)P.rem.size();
      P.rem.resize(P.rem.size() + cnt * n, -1);
      const RasLevelDev& L1 = P.lev[n - 1];
      for (int64_t g = 0; g < cnt; ++g) {
        const uint64_t s = P.mask[L.mask_off + g];
        int t = 0;
        for (uint64_t x = s; x; x &= x - 1, ++t) {
          int c;
          int64_t r;
          rk.rank(s & ~(1ull << ctz(x)), c, r);
          P.rem[L.rem_off + g * n + t] = (int32_t)(L1.cls_off[c] + r);
        }
      }
    } else {
      L.rem_off = -1;
    }
  }
  auto ccount = [&](int n, int c) -> int64_t {
    if (n < 0 || n > N || !need[n] || c < 0 || c >= kRasMaxCls) return 0;
    return P.lev[n].cls_off[c + 1] - P.lev[n].cls_off[c];
  };
  // sectors and their determinant blocks
  int64_t det0 = 0;
  for (auto& s : secs) {
    RasSectorDev S;
    S.na = s.first;
    S.nb = s.second;
    // the SEF vectors make every Ms = 0 determinant vector symmetric under
    // the spin flip when mult = +1 (D_2k+1 is D_2k's flip with the same
    // coefficient; a closed shell is its own flip)
    S.sym = (S.na == S.nb && sp.mult > 0) ? 1 : 0;
    S.pad = 0;
    S.det0 = det0;
    int64_t o = 0;
    for (int a = 0; a < kRasMaxCls; ++a)
      for (int b = 0; b < kRasMaxCls; ++b) {
        S.boff[a][b] = -1;
        if (a + b > m) continue;
        const int64_t sz = ccount(S.na, a) * ccount(S.nb, b);
        if (!sz) continue;
        S.boff[a][b] = o;
        o += sz;
      }
    det0 += o;
    P.sec.push_back(S);
  }
  P.ndet = det0;
  // the D index table names signed vector rows 2 det + negative (and the zero
  // rows 2 ndet, 2 ndet + 1) in 31 bits
  if (P.ndet >= ((int64_t)1 << 30) - 1) return fail("more than 2^30 determinants");
  // pair lists per (shape, budget): shape 0 ordered (q, s), shape 1 q > s;
  // q-major so that consecutive pairs reach determinants of one alpha string
  std::vector<int64_t> qoff(2 * (m + 1)), qn(2 * (m + 1)), qnp(2 * (m + 1));
  P.pidx.assign((size_t)2 * (m + 1) * nbf * nbf, -1);
  for (int shape = 0; shape < 2; ++shape)
    for (int bud = 0; bud <= m; ++bud) {
      const int id = shape * (m + 1) + bud;
      qoff[id] = (int64_t)P.qlist.size();
      int cnt = 0;
      for (int q = 0; q < nbf; ++q)
        for (int s = 0; s < (shape ? q : nbf); ++s) {
          if ((int)((emask >> q) & 1) + (int)((emask >> s) & 1) > bud) continue;
          P.pidx[(size_t)id * nbf * nbf + q * nbf + s] = (int16_t)cnt;
          P.qlist.push_back((int16_t)(q | (s << 8)));
          ++cnt;
        }
      qn[id] = cnt;
      qnp[id] = (cnt + 15) / 16 * 16;  // the products take chunks of 32 pairs and a last one of 16
      for (int64_t t = cnt; t < qnp[id]; ++t) P.qlist.push_back((int16_t)-1);
    }
  // blocks
  const size_t nsec = P.sec.size();
  P.blockid.assign(nsec * 3 * kRasMaxCls * kRasMaxCls, -1);
  const int nb2 = nbf * nbf;
  auto g2d = [&](int p, int q, int r, int s) { return g2(two, p, q, r, s); };
  auto h1d = [&](int p, int q) { return h1(one, p, q); };
  int64_t wunit = 0, dunit = 0;
  double flops = 0.0;
  for (size_t si = 0; si < nsec; ++si) {
    const RasSectorDev& S = P.sec[si];
    if (S.det0 == (si + 1 < nsec ? P.sec[si + 1].det0 : P.ndet)) continue;  // empty sector
    for (int kind = 0; kind < 3; ++kind) {
      int lu, lw;
      if (kind == 0) {
        if (S.na < 1 || S.nb < 1) continue;
        lu = S.na - 1;
        lw = S.nb - 1;
      } else if (kind == 1) {
        if (S.na < 2) continue;
        lu = S.na - 2;
        lw = S.nb;
      } else {
        if (S.nb < 2) continue;
        lu = S.na;
        lw = S.nb - 2;
      }
      const int shape = kind == 0 ? 0 : 1;
      // one-electron integrals: folded into alpha-beta when both spins have
      // electrons, else into the same-spin product of the occupied spin
      const bool fold_ab = S.na >= 1 && S.nb >= 1;
      const bool fold_ss = !fold_ab && ((kind == 1 && S.nb == 0) || (kind == 2 && S.na == 0));
      int64_t gbase_by_budget[kRasMaxCls];
      for (int b = 0; b <= m; ++b) gbase_by_budget[b] = -1;
      // a symmetric sector's mirrored products are not evaluated
      if (S.sym && kind == 2) continue;
      for (int cu = 0; cu <= m; ++cu)
        for (int cw = 0; cw + cu <= m; ++cw) {
          if (S.sym && kind == 0 && cu > cw) continue;
          const int64_t nu = ccount(lu, cu), nw = ccount(lw, cw);
          if (!nu || !nw) continue;
          const int bud = m - cu - cw;
          const int id = shape * (m + 1) + bud;
          RasBlockDev B;
          std::memset(&B, 0, sizeof(B));
          B.kind = kind;
          B.sector = (int)si;
          B.lu = lu;
          B.lw = lw;
          B.cu = cu;
          B.cw = cw;
          B.budget = bud;
          B.nq = (int)qn[id];
          B.nqp = (int)qnp[id];
          B.u0 = P.lev[lu].cls_off[cu];
          B.nu = nu;
          B.w0 = P.lev[lw].cls_off[cw];
          B.nw = nw;
          B.tri = (S.sym && kind == 0 && cu == cw) ? 1 : 0;
          B.nrows = B.tri ? nu * (nu + 1) / 2 : nu * nw;
          B.qoff = qoff[id];
          B.wbase = wunit;
          wunit += B.nrows * B.nqp;
          B.dbase = dunit;
          dunit += (B.nrows + 15) / 16 * 16 * B.nqp;
          B.qmoff = (int64_t)P.qm.size();
          B.umoff = P.lev[lu].mask_off;
          for (int c0 = 0; c0 < B.nqp; c0 += 16) {
            uint64_t bits = 0;
            for (int j = c0; j < c0 + 16 && j < B.nq; ++j) bits |= 1ull << (P.qlist[B.qoff + j] & 0xff);
            P.qm.push_back(bits);
          }
          flops += 2.0 * (double)B.nrows * B.nq * B.nq;
          if (gbase_by_budget[bud] < 0) {
            // the integral matrix of this (sector, kind, budget)
            gbase_by_budget[bud] = (int64_t)P.G.size();
            const int64_t np = B.nqp;
            P.G.resize(P.G.size() + np * np, 0.0);
            double* G = P.G.data() + gbase_by_budget[bud];
            for (int i = 0; i < B.nq; ++i) {
              const int16_t pr = P.qlist[B.qoff + i];
              const int p = pr & 0xff, r = (pr >> 8) & 0xff;
              // symmetric: (pq|rs) = (qp|sr); the upper triangle is
              // computed and mirrored, so G reads the same by rows and
              // columns bit for bit
              for (int j = i; j < B.nq; ++j) {
                const int16_t qs = P.qlist[B.qoff + j];
                const int q = qs & 0xff, s = (qs >> 8) & 0xff;
                double val;
                if (kind == 0) {
                  val = g2d(p, q, r, s);
                  if (fold_ab) {
                    if (r == s) val += h1d(p, q) / S.nb;
                    if (p == q) val += h1d(r, s) / S.na;
                  }
                } else {
                  const int n = kind == 1 ? S.na : S.nb;
                  auto vp = [&](int a, int b, int c, int d) {
                    double x = g2d(a, b, c, d);
                    if (fold_ss) x += ((c == d ? h1d(a, b) : 0.0) + (a == b ? h1d(c, d) : 0.0)) / (n - 1);
                    return x;
                  };
                  val = 0.5 * (vp(p, q, r, s) - vp(r, q, p, s) - vp(p, s, r, q) + vp(r, s, p, q));
                }
                G[(int64_t)i * np + j] = val;
                G[(int64_t)j * np + i] = val;
              }
            }
          }
          B.goff = gbase_by_budget[bud];
          P.blockid[((si * 3 + kind) * kRasMaxCls + cu) * kRasMaxCls + cw] = (int32_t)P.blk.size();
          P.blk.push_back(B);
        }
    }
  }
  (void)nb2;
  if (wunit >= ((int64_t)1 << 31)) return fail("more than 2^31 W elements per column");
  if (dunit >= ((int64_t)1 << 31)) return fail("more than 2^31 D index entries");
  P.wunit = wunit;
  P.dunit = dunit;
  P.gemm_flops_per_col = flops;
  // SEF -> determinants: SEF k of a configuration with open shells is
  // (D_2k + mult D_2k+1)/sqrt(2), each D = (-1)^phase times its canonical
  // determinant; D_2k lies in a sector with n_alpha >= n_beta
  P.sef.resize(sp.nsef_total);
  int64_t k = 0;
  for (int64_t c = 0; c < sp.ncfg; ++c) {
    const uint64_t d = sp.dm[c], s = sp.sm[c];
    const int no = sp.nopen[c];
    const uint16_t* pat = sp.pat.data() + sp.pat_off[no];
    const int neg = sp.cphase[c] & 1;
    auto locate = [&](uint16_t pt, int& sec, int32_t& ga, int32_t& gb) {
      const uint64_t ao = deposit(pt, s);
      const uint64_t am = d | ao, bm = d | (s & ~ao);
      const int dn = popc(am) - popc(bm);
      if (dn < 0) return false;
      sec = dn / 4;
      int ca, cb;
      int64_t ra, rb;
      rk.rank(am, ca, ra);
      rk.rank(bm, cb, rb);
      ga = (int32_t)(P.lev[P.sec[sec].na].cls_off[ca] + ra);
      gb = (int32_t)(P.lev[P.sec[sec].nb].cls_off[cb] + rb);
      return true;
    };
    for (int t = 0; t < sp.nsef[c]; ++t, ++k) {
      RasSef& e = P.sef[k];
      std::memset(&e, 0, sizeof(e));
      int sc;
      int32_t ga, gb;
      if (no == 0) {
        locate(pat[0], sc, ga, gb);
        e.a[0] = ga;
        e.b[0] = gb;
        e.sec[0] = (int8_t)sc;
        e.flag[0] = (uint8_t)(neg | 2);
        e.n = 1;
        continue;
      }
      e.n = 0;
      const bool ok0 = locate(pat[2 * t], sc, ga, gb);
      PSOCI_CHECK(ok0, PSOCI_ERR_STATE, "SEF map: determinant 2k outside the evaluated sectors");
      e.a[0] = ga;
      e.b[0] = gb;
      e.sec[0] = (int8_t)sc;
      e.flag[0] = (uint8_t)neg;
      e.n = 1;
      if (locate(pat[2 * t + 1], sc, ga, gb)) {
        PSOCI_CHECK(sc == 0, PSOCI_ERR_STATE, "SEF map: both determinants of a pair in an Ms > 0 sector");
        e.a[1] = ga;
        e.b[1] = gb;
        e.sec[1] = 0;
        e.flag[1] = (uint8_t)(neg ^ (sp.mult < 0 ? 1 : 0));
        e.n = 2;
      }
    }
  }
  P.ok = true;
}

RasView ras_host_view(const RasPlan& P) {
  RasView v;
  std::memset(&v, 0, sizeof(v));
  v.nbf = P.nbf;
  v.m = P.m;
  v.nsec = (int)P.sec.size();
  v.nblk = (int)P.blk.size();
  v.mask = P.mask.data();
  v.cls = P.cls.data();
  v.add = P.add.data();
  v.rem = P.rem.data();
  for (int n = 0; n <= kRasMaxN && n < (int)P.lev.size(); ++n) v.lev[n] = P.lev[n];
  for (size_t s = 0; s < P.sec.size(); ++s) v.sec[s] = P.sec[s];
  v.blk = P.blk.data();
  v.qlist = P.qlist.data();
  v.pidx = P.pidx.data();
  for (size_t s = 0; s < P.sec.size(); ++s)
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < kRasMaxCls; ++a)
        for (int b = 0; b < kRasMaxCls; ++b)
          v.blockid[s][k][a][b] = P.blockid[((s * 3 + k) * kRasMaxCls + a) * kRasMaxCls + b];
  v.sef = P.sef.data();
  v.nsef = (int64_t)P.sef.size();
  v.ndet = P.ndet;
  v.wunit = P.wunit;
  return v;
}

}  // namespace psoci
