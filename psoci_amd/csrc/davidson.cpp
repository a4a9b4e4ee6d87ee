// Block Davidson on one GPU with an FP32 -> FP64 sigma switch.
//
// Replaces the production driver loop (driver.PSOCI.splittest.C:895-1380)
// built on PsociGAbasis / PsociGAsubspace.  The reference corrects one root
// per iteration (the lowest unconverged, curroot) with a single new vector,
// modified Gram-Schmidt and a GA gop per dot product; here every
// unconverged root of the block gets its correction in the same iteration,
// so one sigma launch multiplies the whole block, and all products over the
// CI dimension are single GPU launches (FP64 MFMA for V^T W).  The numerics
// kept from the reference: Rayleigh-Ritz on V^T H V, residual
// r = sum_j y_j (HV_j - lambda V_j) (PsociGAsubspace.C:874-989), the diagonal
// preconditioner r_i / (D_ii - lambda) applied where |D_ii - lambda| > 1e-3
// (DAVIDSON_MIN, PsociGAbasis.C:1860-1920) -- optionally with Olsen's
// correction against the Ritz vector, which keeps the correction from
// collapsing onto it -- and Ritz collapse that
// transforms HV instead of recomputing it (expandAndReplaceVectorsNew,
// PsociGAbasis.C:483-579).
//
// Mixed precision: sigma runs on the FP32 copy of H with FP32 vectors and
// FP32 accumulation until every residual norm is below switch_tol; then the
// basis collapses to the current Ritz vectors, their sigma is recomputed in
// FP64, and FP64 sigma continues to convergence.  The FP32 phase keeps its
// basis V and HV in FP32 (the sigma it feeds is FP32 anyway); subspace
// products, orthogonalisation, residuals and Rayleigh quotients are FP64
// arithmetic throughout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <type_traits>

#include "ctx.hpp"
#include "linalg.hpp"
#include "smalleig.hpp"

namespace psoci {

namespace {

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double sec() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

}  // namespace

int davidson(Ctx& c, const psoci_dav_params& p, double* evals, double* evecs, psoci_dav_stats* st) {
  Timer total;
  const int64_t n = c.sp.nsef_total;
  const int kw = p.nroots;  // roots wanted
  PSOCI_CHECK(kw >= 1 && kw <= 64 && kw <= n, PSOCI_ERR_ARG, "nroots must be in 1..min(64, n)");
  PSOCI_CHECK(p.extra_roots >= 0 && p.extra_roots <= 64, PSOCI_ERR_ARG, "extra_roots must be in 0..64");
  PSOCI_CHECK(p.guess != 2 || p.extra_roots == 0, PSOCI_ERR_ARG, "guess 2 takes no extra roots");
  // working block: the wanted roots plus guard roots, which speed up the
  // convergence of the highest wanted ones; only the wanted are tested
  int k = (int)std::min<int64_t>((int64_t)kw + p.extra_roots, n);
  const int k0 = k;  // buffer width
  PSOCI_CHECK(p.prec == PSOCI_FP64 || p.prec == PSOCI_FP32 || p.prec == PSOCI_MIXED, PSOCI_ERR_ARG, "bad precision");
  PSOCI_CHECK(p.sigma_mode == PSOCI_SIGMA_STORED || p.sigma_mode == PSOCI_SIGMA_DIRECT, PSOCI_ERR_ARG, "bad sigma mode");
  PSOCI_CHECK(p.tol > 0, PSOCI_ERR_ARG, "tol must be positive");
  PSOCI_CHECK(p.precond == 0 || p.precond == 1, PSOCI_ERR_ARG, "precond must be 0 or 1");
  PSOCI_CHECK(p.graph == 0 || p.graph == 1, PSOCI_ERR_ARG, "graph must be 0 or 1");
  const bool olsen = p.precond == 1;
  PSOCI_CHECK(p.select_rule == 0 || p.select_rule == 1, PSOCI_ERR_ARG, "select_rule must be 0 or 1");
  const bool fill = p.select_rule == 0;
  int mmax = p.max_subspace > 0 ? p.max_subspace : 5 * k;
  mmax = std::max(mmax, 3 * k);
  if (mmax > n) mmax = (int)n;
  int kc = std::min(2 * k, mmax - k > k ? mmax - k : k);
  const int maxit = p.max_iter > 0 ? p.max_iter : 200;
  if (p.sigma_mode == PSOCI_SIGMA_STORED)
    PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
  if (p.sigma_mode == PSOCI_SIGMA_DIRECT && direct_needs_build(c)) build_direct(c);
  if (!c.have_diag) compute_diag(c);

  Linalg L;
  L.stream = c.stream;
  // the basis V and its sigma HV: FP64 for the FP64 phase, FP32 (V32,
  // HV32) for the FP32 phase; a mixed solve holds both
  DevBuf<double> V, HV, R, T;
  // R32: the FP32 phase's residual block (Olsen's correction keeps FP64 R)
  DevBuf<float> V32, HV32, X32, Y32, R32;
  const bool want32 = p.prec != PSOCI_FP64;
  if (p.prec != PSOCI_FP32) {
    V.alloc((size_t)n * mmax);
    HV.alloc((size_t)n * mmax);
  }
  R.alloc((size_t)n * k0);
  T.alloc((size_t)n * std::max(kc, k0));
  if (want32) {
    V32.alloc((size_t)n * mmax);
    HV32.alloc((size_t)n * mmax);
    X32.alloc((size_t)n * k0);
    Y32.alloc((size_t)n * k0);
    if (!olsen) R32.alloc((size_t)n * k0);
  }
  psoci_dav_stats S{};
  S.switch_iter = -1;

  // guess vectors
  PSOCI_CHECK(p.guess >= 0 && p.guess <= 2, PSOCI_ERR_ARG, "guess must be 0, 1 or 2");
  std::vector<int64_t> idx(k);
  if (p.guess == 2) {
    // caller's vectors, below
  } else if (p.guess == 1) {
    std::iota(idx.begin(), idx.end(), 0);
  } else {
    // the k lowest diagonal elements: per-slice candidates on the GPU, the
    // final order of the few candidates on the host
    L.lowest(n, k, c.d_diag.p, idx.data());
  }
  bool phase32 = (p.prec == PSOCI_FP32 || p.prec == PSOCI_MIXED);
  // f(V, HV) on the basis of the current phase
  auto on_basis = [&](auto&& f) {
    if (phase32) f(V32.p, HV32.p);
    else f(V.p, HV.p);
  };
  if (p.guess == 2) {
    PSOCI_CHECK(evecs, PSOCI_ERR_ARG, "guess 2 needs the starting vectors in evecs");
    R.upload(evecs, (size_t)n * k, c.stream);
    if (phase32) L.to_f32(n, k, R.p, k, V32.p, mmax);
    else L.copy_cols(n, k, R.p, k, V.p, mmax);
  } else {
    on_basis([&](auto* Vb, auto*) { L.unit_cols(n, k, idx.data(), Vb, mmax); });
  }

  const double switch_tol = p.switch_tol > 0 ? p.switch_tol : 1e-4;
  // The device work of one step -- sigma of the new block V[:, c0:c0+nb)
  // into HV and its projection V^T HV[:, c0:c0+nb) into L.small -- has no
  // host decision inside, so on the stored sigma (launch- and
  // latency-bound: about 0.1 ms per sigma on 31,650 SEFs) it runs as one
  // captured hipGraph per (c0, nb, precision), instantiated on first use
  // and replayed; every buffer the graph touches is reserved before the
  // first capture (no allocation under capture, no memset nodes).  On the
  // direct sigma (0.5-1 s per launch) the same steps run eagerly.
  const bool graphs = p.graph == 0 && p.sigma_mode == PSOCI_SIGMA_STORED;
  struct Seg {
    int c0, nb;
    bool p32;
    hipGraphExec_t exec;
  };
  std::vector<Seg> segs;
  struct SegsGuard {
    std::vector<Seg>& s;
    ~SegsGuard() {
      for (auto& g : s) (void)hipGraphExecDestroy(g.exec);
    }
  } segs_guard{segs};
  if (graphs) {
    L.reserve_tn(n, mmax, k0);
    L.small.reserve((size_t)mmax * k0);
    const size_t es = phase32 ? 4 : 8;
    c.d_pad.reserve((size_t)2 * n * 32 * std::max<size_t>(es, 8));
  }
  auto device_step = [&](int c0, int nb) {
    if (phase32 && p.sigma_mode != PSOCI_SIGMA_STORED && direct_engine(c) == 2) {
      // the string-driven sigma reads and writes the basis columns in place
      sigma_ras(c, PSOCI_FP32, nb, V32.p + c0, mmax, HV32.p + c0, mmax);
    } else if (phase32) {
      L.copy_cols(n, nb, V32.p + c0, mmax, X32.p, nb);
      if (p.sigma_mode == PSOCI_SIGMA_STORED) sigma_stored(c, PSOCI_FP32, nb, X32.p, nb, Y32.p, nb);
      else sigma_direct(c, PSOCI_FP32, nb, X32.p, nb, Y32.p, nb);
      L.copy_cols(n, nb, Y32.p, nb, HV32.p + c0, mmax);
    } else {
      if (p.sigma_mode == PSOCI_SIGMA_STORED) sigma_stored(c, PSOCI_FP64, nb, V.p + c0, mmax, HV.p + c0, mmax);
      else sigma_direct(c, PSOCI_FP64, nb, V.p + c0, mmax, HV.p + c0, mmax);
    }
    L.small.reserve((size_t)(c0 + nb) * nb);
    on_basis([&](auto* Vb, auto* HVb) { L.gemm_tn(n, c0 + nb, nb, Vb, mmax, HVb + c0, mmax, L.small.p); });
  };
  auto run_step = [&](int c0, int nb) {
    if (!graphs) {
      device_step(c0, nb);
      return;
    }
    Seg* sg = nullptr;
    for (auto& g : segs)
      if (g.c0 == c0 && g.nb == nb && g.p32 == phase32) sg = &g;
    if (!sg) {
      hipGraph_t graph = nullptr;
      PSOCI_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
      try {
        device_step(c0, nb);
      } catch (...) {
        (void)hipStreamEndCapture(c.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      PSOCI_HIP(hipStreamEndCapture(c.stream, &graph));
      hipGraphExec_t exec = nullptr;
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      PSOCI_HIP(e);
      segs.push_back(Seg{c0, nb, phase32, exec});
      sg = &segs.back();
    }
    PSOCI_HIP(hipGraphLaunch(sg->exec, c.stream));
    S.graph_launches += 1;
  };

  std::vector<double> A((size_t)mmax * mmax, 0.0), tmp;
  // sigma of V[:, c0:c0+nb) and the new columns of A = V^T H V
  auto sigma_update = [&](int c0, int nb) {
    Timer ts;
    run_step(c0, nb);
    c.sync();
    S.sigma_seconds += ts.sec();
    S.sigma_calls += 1;
    S.sigma_vectors += nb;
    int m = c0 + nb;
    tmp.resize((size_t)m * nb);
    L.small.download(tmp.data(), (size_t)m * nb, c.stream);
    c.sync();
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < nb; ++j) A[(size_t)i * mmax + c0 + j] = tmp[(size_t)i * nb + j];
    for (int i = c0; i < m; ++i)
      for (int j = c0; j < i; ++j) {
        double s = 0.5 * (A[(size_t)i * mmax + j] + A[(size_t)j * mmax + i]);
        A[(size_t)i * mmax + j] = A[(size_t)j * mmax + i] = s;
      }
    for (int i = 0; i < c0; ++i)
      for (int j = c0; j < m; ++j) A[(size_t)j * mmax + i] = A[(size_t)i * mmax + j];
  };

  int m = k;

  std::vector<double> Y, w, Yk, Yk2, G, P;
  auto solve = [&](int mm) {
    Y.assign((size_t)mm * mm, 0.0);
    w.assign(mm, 0.0);
    for (int i = 0; i < mm; ++i)
      for (int j = 0; j < mm; ++j) Y[(size_t)i * mm + j] = A[(size_t)i * mmax + j];
    Timer te;
    PSOCI_CHECK(sym_eig(mm, Y.data(), w.data()), PSOCI_ERR_NUMERIC, "subspace eigensolver failed");
    S.eig_seconds += te.sec();
  };
  auto cols_of = [&](int mm, int first, int cnt, std::vector<double>& out) {
    out.assign((size_t)mm * cnt, 0.0);
    for (int i = 0; i < mm; ++i)
      for (int j = 0; j < cnt; ++j) out[(size_t)i * cnt + j] = Y[(size_t)i * mm + first + j];
  };
  // collapse V, HV onto Ritz vectors 0..nc-1 of the current solve (Y, w)
  auto collapse = [&](int nc) {
    std::vector<double> Yc;
    cols_of(m, 0, nc, Yc);
    on_basis([&](auto* Vb, auto* HVb) {
      if (nc <= 128) {
        L.gemm_nn_inplace(n, m, nc, Yc.data(), Vb, mmax);
        L.gemm_nn_inplace(n, m, nc, Yc.data(), HVb, mmax);
      } else {
        // via T (FP64): Vb = T rounded to the basis precision
        auto back = [&](auto* dst) {
          if constexpr (std::is_same_v<std::remove_pointer_t<decltype(dst)>, float>)
            L.to_f32(n, nc, T.p, nc, dst, mmax);
          else
            L.copy_cols(n, nc, T.p, nc, dst, mmax);
        };
        L.gemm_nn(n, m, nc, Vb, mmax, Yc.data(), 0.0, T.p, nc);
        back(Vb);
        L.gemm_nn(n, m, nc, HVb, mmax, Yc.data(), 0.0, T.p, nc);
        back(HVb);
      }
    });
    std::fill(A.begin(), A.end(), 0.0);
    for (int i = 0; i < nc; ++i) A[(size_t)i * mmax + i] = w[i];
    m = nc;
  };
  // the symmetric orthonormaliser of a block with Gram matrix gram (nb x
  // nb): M (nb x nk) with M^T gram M = I over the eigenvectors whose
  // eigenvalues are not negligible; returns nk (0: none left or failure)
  auto orth_matrix = [&](int nb, const double* gram, std::vector<double>& M) -> int {
    G.assign(gram, gram + (size_t)nb * nb);
    std::vector<double> lam(nb);
    Timer te;
    const bool ok = sym_eig(nb, G.data(), lam.data());
    S.eig_seconds += te.sec();
    if (!ok) return 0;
    double lmax = lam[nb - 1];
    std::vector<int> keep;
    for (int j = nb - 1; j >= 0; --j)
      if (lam[j] > 1e-14 * lmax && lam[j] > 1e-28) keep.push_back(j);
    int nk = (int)keep.size();
    if (nk == 0) return 0;
    M.resize((size_t)nb * nk);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nk; ++j) M[(size_t)i * nk + j] = G[(size_t)i * nb + keep[j]] / std::sqrt(lam[keep[j]]);
    return nk;
  };
  // symmetric orthonormalisation of columns [c0, c0+nb) of V in place (via T)
  auto sym_orth = [&](int c0, int nb) -> int {
    std::vector<double> Gb((size_t)nb * nb), M;
    on_basis([&](auto* Vb, auto*) { L.gemm_tn_host(n, nb, nb, Vb + c0, mmax, Vb + c0, mmax, Gb.data()); });
    const int nk = orth_matrix(nb, Gb.data(), M);
    if (nk == 0) return 0;
    // nk <= nb <= k <= 128
    on_basis([&](auto* Vb, auto*) { L.gemm_nn_inplace(n, nb, nk, M.data(), Vb + c0, mmax); });
    return nk;
  };
  auto project_out = [&](int c0, int nb) {
    if (c0 == 0) return;
    P.resize((size_t)c0 * nb);
    on_basis([&](auto* Vb, auto*) {
      L.gemm_tn_host(n, c0, nb, Vb, mmax, Vb + c0, mmax, P.data());
      for (auto& x : P) x = -x;
      L.gemm_nn(n, c0, nb, Vb, mmax, P.data(), 1.0, Vb + c0, mmax);
    });
  };

  if (p.guess == 2) {
    // two passes of symmetric orthonormalisation of the caller's vectors
    for (int pass = 0; pass < 2; ++pass)
      PSOCI_CHECK(sym_orth(0, k) == k, PSOCI_ERR_ARG, "starting vectors are linearly dependent");
  }
  sigma_update(0, k);

  std::vector<double> rn(k, 0.0);
  const char* tr = std::getenv("PSOCI_TRACE");
  const bool trace = tr && tr[0] == '1';
  bool converged = false;
  int it = 0;
  for (it = 1; it <= maxit; ++it) {
    solve(m);
    cols_of(m, 0, k, Yk);
    Yk2 = Yk;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < k; ++j) Yk2[(size_t)i * k + j] *= w[j];
    // residuals R = HV Yk - V Yk diag(w); with Olsen's correction also the
    // Ritz vectors X = V Yk (into T, free until the next collapse) in the
    // same pass over V and HV
    // in the FP32 phase (without Olsen's correction) the residual block is
    // FP32: its norms and the preconditioned corrections need no more
    const bool r32 = phase32 && !olsen;
    G.resize((size_t)k * k);
    if (r32) {
      L.gemm_nn2(n, m, k, HV32.p, mmax, Yk.data(), V32.p, mmax, Yk2.data(), R32.p, k);
      L.gemm_tn_host(n, k, k, R32.p, k, R32.p, k, G.data());
    } else {
      on_basis([&](auto* Vb, auto* HVb) {
        L.gemm_nn2(n, m, k, HVb, mmax, Yk.data(), Vb, mmax, Yk2.data(), R.p, k, olsen ? T.p : nullptr, k);
      });
      L.gemm_tn_host(n, k, k, R.p, k, R.p, k, G.data());
    }
    double maxr = 0.0;
    for (int j = 0; j < k; ++j) {
      rn[j] = std::sqrt(std::max(G[(size_t)j * k + j], 0.0));
      if (j < kw) maxr = std::max(maxr, rn[j]);
    }
    S.max_rnorm = maxr;
    if (trace)
      std::fprintf(stderr, "[psoci davidson] it %d m %d %s max|r| %.3e  E0 %.12f E%d %.12f\n", it, m,
                   phase32 ? "fp32" : "fp64", maxr, w[0], kw - 1, w[kw - 1]);
    if (phase32 && p.prec == PSOCI_MIXED && maxr < switch_tol) {
      // switch: collapse to the Ritz vectors and recompute their sigma in FP64
      // guard roots ride in the FP32 phase only: the FP64 phase continues
      // with the wanted roots (its launches cost more per column)
      // the Ritz vectors V32 Y into the FP64 basis; their sigma is
      // recomputed below
      {
        std::vector<double> Yc;
        cols_of(m, 0, kw, Yc);
        L.gemm_nn(n, m, kw, V32.p, mmax, Yc.data(), 0.0, V.p, mmax);
      }
      k = kw;
      kc = std::min(2 * k, mmax - k > k ? mmax - k : k);
      std::fill(A.begin(), A.end(), 0.0);
      for (int i = 0; i < k; ++i) A[(size_t)i * mmax + i] = w[i];
      m = k;
      phase32 = false;
      S.switch_iter = it;
      // the FP64 phase starts from an orthonormal basis to FP64 precision
      for (int pass = 0; pass < 2; ++pass)
        PSOCI_CHECK(sym_orth(0, k) == k, PSOCI_ERR_NUMERIC, "Ritz vectors dependent at the switch");
      sigma_update(0, k);
      continue;
    }
    if (!phase32 && maxr < p.tol) {
      converged = true;
      break;
    }
    if (phase32 && p.prec == PSOCI_FP32 && maxr < p.tol) {
      converged = true;
      break;
    }
    const double sel_tol = (phase32 && p.prec == PSOCI_MIXED) ? switch_tol : p.tol;
    std::vector<int> sel;
    if (!fill) {
      for (int j = 0; j < k; ++j)
        if (rn[j] >= sel_tol) sel.push_back(j);
    } else {
      // the unconverged wanted roots first; the sigma launch pads the block
      // to its kernel width anyway, so the free columns take the corrections
      // of unconverged guard roots, then of the converged wanted roots with
      // the largest residuals
      for (int j = 0; j < kw && j < k; ++j)
        if (rn[j] >= sel_tol) sel.push_back(j);
      int width = 1;  // the kernel width the block is padded to (powers of two up to 32)
      while (width < (int)sel.size()) width <<= 1;
      if (sel.empty() || width > 32) width = (int)sel.size();
      for (int j = kw; j < k && (int)sel.size() < width; ++j)
        if (rn[j] >= sel_tol) sel.push_back(j);
      std::vector<int> rest;
      for (int j = 0; j < kw && j < k; ++j)
        if (rn[j] < sel_tol && rn[j] > 1e-3 * sel_tol) rest.push_back(j);
      std::stable_sort(rest.begin(), rest.end(), [&](int x, int y) { return rn[x] > rn[y]; });
      for (int j : rest) {
        if ((int)sel.size() >= width) break;
        sel.push_back(j);
      }
    }
    int nsel = (int)sel.size();
    if (olsen) L.precondition_olsen(n, k, c.d_diag.p, w.data(), 1.0e-3, R.p, k, T.p, k);
    if (m + nsel > mmax) {
      collapse(kc);
      // the retained Ritz values and vectors are unchanged: w[0..k) still valid
    }
    // a space of fewer than 2k SEFs (mmax = n < kc + k) leaves fewer free
    // columns than selected roots after the collapse: keep the first
    // corrections (the unconverged wanted roots, ascending, come first); none
    // fit when the basis spans the whole space, whose Ritz pairs are then exact
    if (m + nsel > mmax) nsel = mmax - m;
    if (nsel <= 0) break;
    // new block = selected preconditioned residuals, placed at V[:, m:m+nsel)
    if (olsen) {
      std::vector<double> Ssel((size_t)k * nsel, 0.0);
      for (int j = 0; j < nsel; ++j) Ssel[(size_t)sel[j] * nsel + j] = 1.0;
      on_basis([&](auto* Vb, auto*) { L.gemm_nn(n, k, nsel, R.p, k, Ssel.data(), 0.0, Vb + m, mmax); });
    } else {
      if (phase32)
        L.precondition_select(n, k, c.d_diag.p, w.data(), 1.0e-3, R32.p, k, sel.data(), nsel, V32.p + m, mmax);
      else
        L.precondition_select(n, k, c.d_diag.p, w.data(), 1.0e-3, R.p, k, sel.data(), nsel, V.p + m, mmax);
    }
    int nb = nsel;
    {
      // block Gram-Schmidt against V[:, 0:m) and symmetric orthonormalisation
      // of the block, twice.  In the FP32 phase (whose sigma carries 1e-7
      // relative errors anyway) the second pass runs only when the first
      // removed more than half of a column's squared norm (the
      // Daniel-Gragg-Kaufman-Stewart criterion) or left the block's columns
      // nearly dependent (the smallest eigenvalue of their correlation
      // matrix below 1e-2).  The FP64 phase always takes both passes: with
      // one pass under the same test, solves to residual 1e-8 on small
      // spaces stalled where two passes converge
      std::vector<double> G0((size_t)nb * nb), G1((size_t)nb * nb);
      bool again = m > 0, fused = false;
      if (phase32 && m > 0) {
        // one projection gives P = V^T x (rows 0..m) and the block's Gram
        // matrix G0 (rows m..m+nb)
        std::vector<double> Q((size_t)(m + nb) * nb);
        on_basis([&](auto* Vb, auto*) { L.gemm_tn_host(n, m + nb, nb, Vb, mmax, Vb + m, mmax, Q.data()); });
        P.assign(Q.begin(), Q.begin() + (size_t)m * nb);
        std::copy(Q.begin() + (size_t)m * nb, Q.end(), G0.begin());
        // V[:, 0:m) is orthonormal, so after x -= V (V^T x) the block's Gram
        // matrix is G0 - P^T P (P = V^T x, m x nb): exact where the test
        // below accepts one pass (at most half of a column's squared norm
        // removed: cancellation costs at most a factor 2 in relative error)
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j < nb; ++j) {
            double g = G0[(size_t)i * nb + j];
            for (int t = 0; t < m; ++t) g -= P[(size_t)t * nb + i] * P[(size_t)t * nb + j];
            G1[(size_t)i * nb + j] = g;
          }
        again = false;
        for (int j = 0; j < nb; ++j)
          if (!(G1[(size_t)j * nb + j] >= 0.5 * G0[(size_t)j * nb + j]) || !(G1[(size_t)j * nb + j] > 0.0))
            again = true;
        if (!again) {
          std::vector<double> C((size_t)nb * nb), lam(nb);
          for (int i = 0; i < nb; ++i)
            for (int j = 0; j < nb; ++j)
              C[(size_t)i * nb + j] =
                  G1[(size_t)i * nb + j] / std::sqrt(G1[(size_t)i * nb + i] * G1[(size_t)j * nb + j]);
          if (!sym_eig(nb, C.data(), lam.data()) || !(lam[0] >= 1e-2)) again = true;
        }
        if (!again) {
          // one pass: the projection and the orthonormalisation together,
          // x M - V (P M) = [V x] [-P M; M] into V[:, m:m+nk)
          std::vector<double> M;
          const int nk = orth_matrix(nb, G1.data(), M);
          if (nk > 0) {
            std::vector<double> Yf((size_t)(m + nb) * nk, 0.0);
            for (int t = 0; t < m; ++t)
              for (int j = 0; j < nk; ++j) {
                double acc = 0.0;
                for (int i = 0; i < nb; ++i) acc += P[(size_t)t * nb + i] * M[(size_t)i * nk + j];
                Yf[(size_t)t * nk + j] = -acc;
              }
            for (int i = 0; i < nb; ++i)
              for (int j = 0; j < nk; ++j) Yf[(size_t)(m + i) * nk + j] = M[(size_t)i * nk + j];
            on_basis([&](auto* Vb, auto*) { L.gemm_nn_rows(n, m + nb, nk, Vb, mmax, Yf.data(), Vb + m, mmax); });
          }
          nb = nk;
          fused = true;
        } else {
          // the first pass's update from the P at hand, then the second pass
          std::vector<double> Pn(P.size());
          for (size_t e = 0; e < P.size(); ++e) Pn[e] = -P[e];
          on_basis([&](auto* Vb, auto*) { L.gemm_nn(n, m, nb, Vb, mmax, Pn.data(), 1.0, Vb + m, mmax); });
        }
      } else {
        project_out(m, nb);
      }
      if (!fused) {
        if (again) project_out(m, nb);
        nb = sym_orth(m, nb);
        if (nb > 0 && again) {
          project_out(m, nb);
          nb = sym_orth(m, nb);
        }
      }
    }
    if (nb == 0) break;  // no new directions: stagnation
    sigma_update(m, nb);
    m += nb;
  }
  S.iterations = std::min(it, maxit);
  S.converged = converged ? 1 : 0;
  solve(m);
  for (int j = 0; j < kw; ++j) evals[j] = w[j];
  if (evecs) {
    cols_of(m, 0, kw, Yk);
    on_basis([&](auto* Vb, auto*) { L.gemm_nn(n, m, kw, Vb, mmax, Yk.data(), 0.0, R.p, kw); });
    R.download(evecs, (size_t)n * kw, c.stream);
    c.sync();
  }
  S.seconds = total.sec();
  if (st) *st = S;
  return converged ? PSOCI_OK : PSOCI_ERR_CONVERGE;
}

}  // namespace psoci
