// extern "C" boundary of the PSOCI MI355X library (declared in include/psoci.h).
// Every entry point catches psoci::Error and HIP failures and returns a
// status code; the message is kept on the context for psoci_last_error().
#include <cstring>

#include "ctx.hpp"
#include "linalg.hpp"

using psoci::Ctx;
using psoci::Error;

struct psoci_ctx {
  Ctx c;
};

namespace {

template <class F>
int guarded(psoci_ctx* ctx, F&& f) {
  if (!ctx) {
    psoci::set_ctxfree_error("null context");
    return PSOCI_ERR_ARG;
  }
  try {
    PSOCI_HIP(hipSetDevice(ctx->c.device));
    ctx->c.err.clear();
    return f(ctx->c);
  } catch (const Error& e) {
    ctx->c.err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    ctx->c.err = e.what();
    return PSOCI_ERR_ARG;
  }
}

int64_t nndaf(int64_t i) { return i * (i + 1) / 2; }

}  // namespace

extern "C" {

int psoci_ctx_create(int device, psoci_ctx** out) {
  if (!out) return PSOCI_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PSOCI_ERR_HIP;
  if (device < 0 || device >= count) return PSOCI_ERR_ARG;
  psoci_ctx* ctx = new psoci_ctx;
  ctx->c.device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return PSOCI_ERR_HIP;
  }
  // the allocation-free sigma's scratch address on this device (once per device)
  const int rc = psoci::noalloc_init(device);
  if (rc != PSOCI_OK) {
    delete ctx;
    return rc;
  }
  *out = ctx;
  return PSOCI_OK;
}

int psoci_ctx_destroy(psoci_ctx* ctx) {
  if (!ctx) return PSOCI_ERR_ARG;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  delete ctx;
  return PSOCI_OK;
}

const char* psoci_last_error(const psoci_ctx* ctx) {
  return ctx ? ctx->c.err.c_str() : psoci::ctxfree_error().c_str();
}

int psoci_set_integrals_sifs(psoci_ctx* ctx, const char* path) {
  return guarded(ctx, [&](Ctx& c) {
    int nbf = 0;
    psoci::read_sifs(path, &nbf, nullptr, nullptr, nullptr);
    std::vector<double> one(nndaf(nbf)), two(nndaf(nndaf(nbf)));
    double ecore = 0.0;
    psoci::read_sifs(path, &nbf, one.data(), two.data(), &ecore);
    (void)c;
    return psoci_set_integrals(ctx, nbf, one.data(), two.data(), ecore);
  });
}

int psoci_set_integrals(psoci_ctx* ctx, int nbf, const double* valone, const double* valtwo, double ecore) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(nbf > 0 && nbf <= 64 && valone && valtwo, PSOCI_ERR_ARG, "bad integral arguments");
    if (c.have_space) PSOCI_CHECK(nbf == c.sp.nbf, PSOCI_ERR_ARG, "nbf differs from the configuration list");
    c.d_one.upload(valone, nndaf(nbf), c.stream);
    c.d_two.upload(valtwo, nndaf(nndaf(nbf)), c.stream);
    c.sync();
    c.ints_hash = psoci::hash_bytes(valtwo, nndaf(nndaf(nbf)) * sizeof(double),
                                    psoci::hash_bytes(valone, nndaf(nbf) * sizeof(double), (uint64_t)nbf));
    c.nbf_ints = nbf;
    c.ecore = ecore;
    c.have_ints = true;
    c.have_h = c.have_diag = c.have_pairs = c.have_trans = false;
    c.drop_ras();
    return PSOCI_OK;
  });
}

int psoci_set_configs(psoci_ctx* ctx, int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ,
                      const int32_t* jsym) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(occ && jsym, PSOCI_ERR_ARG, "null configuration arrays");
    if (c.have_ints) PSOCI_CHECK(nbf == c.nbf_ints, PSOCI_ERR_ARG, "nbf differs from the integrals");
    c.have_space = false;
    psoci::build_space(c.sp, nbf, nelec, ksym, ncfg, occ, jsym);
    psoci::upload_space(c);
    const int64_t dims[4] = {nbf, nelec, ksym, ncfg};
    c.space_hash = psoci::hash_bytes(jsym, ncfg * sizeof(int32_t),
                                     psoci::hash_bytes(occ, ncfg * nbf, psoci::hash_bytes(dims, sizeof(dims))));
    c.have_space = true;
    c.have_h = c.have_diag = c.have_pairs = c.have_trans = false;
    c.drop_ras();
    return PSOCI_OK;
  });
}

int psoci_set_configs_file(psoci_ctx* ctx, const char* path, int order) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(order == PSOCI_CONFIGS_FILE_ORDER || order == PSOCI_CONFIGS_SORTED, PSOCI_ERR_ARG,
                "order must be PSOCI_CONFIGS_FILE_ORDER or PSOCI_CONFIGS_SORTED");
    psoci::ConfigFile cf;
    psoci::read_config_file(path, order == PSOCI_CONFIGS_SORTED, cf, nullptr);
    (void)c;
    return psoci_set_configs(ctx, cf.nbf, cf.nelec, cf.ksym, (int64_t)cf.jsym.size(), cf.occ.data(),
                             cf.jsym.data());
  });
}

int psoci_space_info(psoci_ctx* ctx, int64_t* nsef_total, int64_t* ndet_total, int32_t* nsef, int64_t* nsef_offset) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space, PSOCI_ERR_STATE, "no configuration list");
    if (nsef_total) *nsef_total = c.sp.nsef_total;
    if (ndet_total) *ndet_total = c.sp.ndet_total;
    if (nsef) std::memcpy(nsef, c.sp.nsef.data(), c.sp.ncfg * sizeof(int32_t));
    if (nsef_offset) std::memcpy(nsef_offset, c.sp.sef_off.data(), c.sp.ncfg * sizeof(int64_t));
    return PSOCI_OK;
  });
}

int psoci_dets(psoci_ctx* ctx, int64_t I, uint64_t* alpha, uint64_t* beta, int32_t* ms, int32_t* phase) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space, PSOCI_ERR_STATE, "no configuration list");
    PSOCI_CHECK(I >= 0 && I < c.sp.ncfg, PSOCI_ERR_ARG, "configuration index out of range");
    const psoci::Space& sp = c.sp;
    int s = sp.nopen[I];
    const uint16_t* pat = sp.pat.data() + sp.pat_off[12 * sp.cls[I] + s];
    for (int d = 0; d < sp.ndet[I]; ++d) {
      uint64_t ao = 0, rest = sp.sm[I];
      uint32_t pbits = pat[d];
      while (rest) {
        uint64_t low = rest & (~rest + 1);
        if (pbits & 1u) ao |= low;
        pbits >>= 1;
        rest ^= low;
      }
      uint64_t a = sp.dm[I] | ao, b = sp.dm[I] | (sp.sm[I] & ~ao);
      if (alpha) alpha[d] = a;
      if (beta) beta[d] = b;
      if (ms) ms[d] = (__builtin_popcountll(a) - __builtin_popcountll(b)) / 2;
      if (phase) phase[d] = sp.cphase[I];
    }
    return PSOCI_OK;
  });
}

int psoci_block(psoci_ctx* ctx, int64_t I, int64_t J, double* sefsef, int* status) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(I >= 0 && I < c.sp.ncfg && J >= 0 && J < c.sp.ncfg && sefsef, PSOCI_ERR_ARG, "bad block arguments");
    psoci::block_on_device(c, I, J, sefsef, status);
    return PSOCI_OK;
  });
}

int psoci_build_h(psoci_ctx* ctx, int64_t* nnz_lower) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    psoci::build_h(c);
    if (nnz_lower) *nnz_lower = c.nnz_lower;
    return PSOCI_OK;
  });
}

int psoci_build_h_timing(psoci_ctx* ctx, double* seconds) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(seconds, PSOCI_ERR_ARG, "seconds must not be NULL");
    PSOCI_CHECK(c.have_h && c.hbuild_seconds[0] >= 0.0, PSOCI_ERR_STATE, "no stored Hamiltonian built here");
    seconds[0] = c.hbuild_seconds[0];
    seconds[1] = c.hbuild_seconds[1];
    return PSOCI_OK;
  });
}

int psoci_h_save(psoci_ctx* ctx, const char* path) {
  return guarded(ctx, [&](Ctx& c) {
    psoci::save_h(c, path);
    return PSOCI_OK;
  });
}

int psoci_h_load(psoci_ctx* ctx, const char* path, int64_t* nnz_lower) {
  return guarded(ctx, [&](Ctx& c) {
    psoci::load_h(c, path);
    if (nnz_lower) *nnz_lower = c.nnz_lower;
    return PSOCI_OK;
  });
}

int psoci_get_h(psoci_ctx* ctx, int64_t* rowptr, int32_t* col, double* val, double* diag) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "no stored Hamiltonian");
    psoci::get_h_lower(c, rowptr, col, val, diag);
    return PSOCI_OK;
  });
}

int psoci_direct_info(psoci_ctx* ctx, int64_t* npairs, int64_t* nelems) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    if (!c.have_pairs) psoci::build_pairs(c);
    if (npairs) *npairs = c.nadj;
    if (nelems) *nelems = c.direct_elems;
    return PSOCI_OK;
  });
}

int psoci_direct_engine(psoci_ctx* ctx, int request, int* engine) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(request >= -1 && request <= 2, PSOCI_ERR_ARG, "request must be -1, 0, 1 or 2");
    const int before = c.direct_engine_req;
    if (request >= 0) {
      c.direct_engine_req = request;
      if (request == 1 && c.have_ras) c.drop_ras();
      if (request == 2 && !c.have_ras) c.ras_tried = false;
    }
    if (engine) {
      PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
      try {
        psoci::build_direct(c);
      } catch (...) {
        c.direct_engine_req = before;  // a request that cannot be met leaves the choice as it was
        throw;
      }
      *engine = psoci::direct_engine(c);
    }
    return PSOCI_OK;
  });
}

int psoci_direct_ras_info(psoci_ctx* ctx, int* m, int* n_external, int* nsectors, int* nblocks, int64_t* ndet,
                          int64_t* w_per_column, double* flops_per_column) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_ras, PSOCI_ERR_STATE, "the string-driven direct sigma is not set up");
    const psoci::RasPlan& P = c.ras;
    if (m) *m = P.m;
    if (n_external) *n_external = psoci::popc(P.emask);
    if (nsectors) *nsectors = (int)P.sec.size();
    if (nblocks) *nblocks = (int)P.blk.size();
    if (ndet) *ndet = P.ndet;
    if (w_per_column) *w_per_column = P.wunit;
    if (flops_per_column) *flops_per_column = P.gemm_flops_per_col;
    return PSOCI_OK;
  });
}

int psoci_direct_ras_timing(psoci_ctx* ctx, int prec, int k, int reps, double* ms, double* flops) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "prec must be PSOCI_FP64 or PSOCI_FP32");
    PSOCI_CHECK(ms, PSOCI_ERR_ARG, "null output");
    psoci::ras_timing(c, prec, k, reps, ms, flops);
    return PSOCI_OK;
  });
}

int psoci_direct_det_pairs(psoci_ctx* ctx, int64_t* det_pairs) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(det_pairs, PSOCI_ERR_ARG, "null output");
    *det_pairs = psoci::direct_det_pairs(c);
    return PSOCI_OK;
  });
}

int psoci_direct_schedule(psoci_ctx* ctx, int* order, int* chunk, double* tune_ms) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    if (!c.have_pairs) psoci::build_pairs(c);
    if (order) *order = c.direct_order[1] ? 1 : 0;
    if (chunk) *chunk = c.direct_chunk[1];
    if (tune_ms) *tune_ms = c.direct_tune_ms[1];
    return PSOCI_OK;
  });
}

int psoci_direct_schedule_prec(psoci_ctx* ctx, int prec, int* order, int* chunk, double* tune_ms) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    if (!c.have_pairs) psoci::build_pairs(c);
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "prec must be PSOCI_FP64 or PSOCI_FP32");
    const int pi = prec == PSOCI_FP32 ? 1 : 0;
    if (order) *order = c.direct_order[pi] ? 1 : 0;
    if (chunk) *chunk = c.direct_chunk[pi];
    if (tune_ms) *tune_ms = c.direct_tune_ms[pi];
    return PSOCI_OK;
  });
}

int psoci_stored_shape(psoci_ctx* ctx, int prec, int* lanes_per_nz, int* rows_per_wave, double* tune_ms) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "prec must be PSOCI_FP64 or PSOCI_FP32");
    const int pi = prec == PSOCI_FP32 ? 1 : 0;
    if (lanes_per_nz) *lanes_per_nz = c.spmm_g[pi];
    if (rows_per_wave) *rows_per_wave = c.spmm_r[pi];
    if (tune_ms) *tune_ms = c.spmm_tune_ms[pi];
    return PSOCI_OK;
  });
}

int psoci_stored_buckets(psoci_ctx* ctx, int prec, int* used, int64_t* rows, int* lanes_per_nz, int* rows_per_wave,
                         double* bucketed_ms) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "prec must be PSOCI_FP64 or PSOCI_FP32");
    const int pi = prec == PSOCI_FP32 ? 1 : 0;
    if (used) *used = c.spmm_bucketed[pi] ? 1 : 0;
    if (bucketed_ms) *bucketed_ms = c.spmm_bucket_ms[pi];
    for (int b = 0; b < psoci::kSpmmBuckets; ++b) {
      if (rows) rows[b] = c.spmm_boff[b + 1] - c.spmm_boff[b];
      if (lanes_per_nz) lanes_per_nz[b] = c.spmm_bg[pi][b];
      if (rows_per_wave) rows_per_wave[b] = c.spmm_br[pi][b];
    }
    return PSOCI_OK;
  });
}

int psoci_diag(psoci_ctx* ctx, double* diag) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(diag, PSOCI_ERR_ARG, "null output");
    if (!c.have_diag) psoci::compute_diag(c);
    c.d_diag.download(diag, c.sp.nsef_total, c.stream);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_sigma(psoci_ctx* ctx, int mode, int prec, int k, const void* X, void* Y, int on_device) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(k >= 1 && X && Y, PSOCI_ERR_ARG, "bad sigma arguments");
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "sigma precision must be FP64 or FP32");
    PSOCI_CHECK(mode == PSOCI_SIGMA_STORED || mode == PSOCI_SIGMA_DIRECT, PSOCI_ERR_ARG, "bad sigma mode");
    if (mode == PSOCI_SIGMA_STORED) PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
    if (mode == PSOCI_SIGMA_DIRECT && psoci::direct_needs_build(c)) psoci::build_direct(c);
    const int64_t n = c.sp.nsef_total;
    const size_t es = prec == PSOCI_FP32 ? 4 : 8;
    const void* dx = X;
    void* dy = Y;
    psoci::DevBuf<char> bx, by;
    if (!on_device) {
      bx.upload(static_cast<const char*>(X), (size_t)n * k * es, c.stream);
      by.alloc((size_t)n * k * es);
      dx = bx.p;
      dy = by.p;
    }
    if (mode == PSOCI_SIGMA_STORED) psoci::sigma_stored(c, prec, k, dx, k, dy, k);
    else psoci::sigma_direct(c, prec, k, dx, k, dy, k);
    if (!on_device) by.download(static_cast<char*>(Y), (size_t)n * k * es, c.stream);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_sigma_noalloc(psoci_ctx* ctx, int mode, int prec, int k, const void* X, void* Y) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(k >= 1 && X && Y, PSOCI_ERR_ARG, "bad sigma arguments");
    PSOCI_CHECK(prec == PSOCI_FP64 || prec == PSOCI_FP32, PSOCI_ERR_ARG, "sigma precision must be FP64 or FP32");
    PSOCI_CHECK(mode == PSOCI_SIGMA_STORED || mode == PSOCI_SIGMA_DIRECT, PSOCI_ERR_ARG, "bad sigma mode");
    // building the operator allocates: it must exist already
    if (mode == PSOCI_SIGMA_STORED) PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "stored sigma needs psoci_build_h first");
    if (mode == PSOCI_SIGMA_DIRECT)
      PSOCI_CHECK(psoci::direct_engine(c), PSOCI_ERR_STATE,
                  "the direct sigma's operator is not built (psoci_direct_engine, psoci_direct_info or psoci_sigma "
                  "builds it)");
    psoci::Staging st;
    psoci::noalloc_scratch(c.device, &st.p, &st.bytes);
    const size_t need = psoci::sigma_staging_bytes(c, mode, prec, k, X, k, Y, k);
    PSOCI_CHECK(need <= st.bytes, PSOCI_ERR_ARG,
                "this block needs " + std::to_string(need) + " bytes of staging and the scratch holds " +
                    std::to_string(st.bytes) + ": pass a 16-byte aligned block of a kernel width (1, 2, 4, 8, 16, 32; "
                    "24 for the direct sigma) or fewer columns");
    if (mode == PSOCI_SIGMA_STORED) psoci::sigma_stored(c, prec, k, X, k, Y, k, st);
    else psoci::sigma_direct(c, prec, k, X, k, Y, k, st);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_build_transition(psoci_ctx* ctx, int64_t ncfg2, const uint8_t* occ2, const int32_t* jsym2, int64_t* nnz) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(ncfg2 >= 1 && occ2 && jsym2, PSOCI_ERR_ARG, "bad second configuration list");
    psoci::Space s2;
    psoci::build_space(s2, c.sp.nbf, c.sp.nelec, c.sp.ksym, ncfg2, occ2, jsym2);
    c.have_trans = false;
    psoci::build_transition(c, s2, nnz);
    return PSOCI_OK;
  });
}

int psoci_get_transition(psoci_ctx* ctx, int64_t* rowptr, int32_t* col, double* val) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_trans, PSOCI_ERR_STATE, "no transition Hamiltonian: psoci_build_transition first");
    const int64_t n = c.sp.nsef_total;
    if (rowptr) c.d_trow.download(rowptr, n + 1, c.stream);
    if (col) c.d_tcol.download(col, c.nnz_trans, c.stream);
    if (val) c.d_tval.download(val, c.nnz_trans, c.stream);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_davidson(psoci_ctx* ctx, const psoci_dav_params* params, double* evals, double* evecs,
                   psoci_dav_stats* stats) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(params && evals, PSOCI_ERR_ARG, "null Davidson arguments");
    return psoci::davidson(c, *params, evals, evecs, stats);
  });
}

int psoci_rdm(psoci_ctx* ctx, int k, const double* C, const double* w, double* grad_one, double* grad_two,
              double* grad_so) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(k >= 1 && C, PSOCI_ERR_ARG, "bad RDM arguments");
    const int64_t n = c.sp.nsef_total;
    const int64_t n1 = nndaf(c.sp.nbf), n2 = nndaf(n1);
    if (!c.have_pairs) psoci::build_pairs(c);
    psoci::DevBuf<double> dC, g1, g2;
    dC.upload(C, (size_t)n * k, c.stream);
    g1.alloc(n1);
    g2.alloc(n2);
    psoci::rdm_grad(c, k, dC.p, w, g1.p, g2.p);
    if (grad_one) g1.download(grad_one, n1, c.stream);
    if (grad_two) g2.download(grad_two, n2, c.stream);
    c.sync();
    if (grad_so) std::memset(grad_so, 0, n1 * sizeof(double));
    return PSOCI_OK;
  });
}

int psoci_energy_contrib(psoci_ctx* ctx, int k, const double* C, const double* evals, double* deltae) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    PSOCI_CHECK(k >= 1 && C && evals && deltae, PSOCI_ERR_ARG, "bad energy-contribution arguments");
    if (!c.have_diag) psoci::compute_diag(c);
    const int64_t n = c.sp.nsef_total, ncfg = c.sp.ncfg;
    psoci::DevBuf<double> dC, dE, dD;
    dC.upload(C, (size_t)n * k, c.stream);
    dE.upload(evals, (size_t)k, c.stream);
    dD.alloc((size_t)ncfg * k);
    psoci::energy_contrib(c, k, dC.p, dE.p, dD.p);
    dD.download(deltae, (size_t)ncfg * k, c.stream);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_project(psoci_ctx* ctx, int64_t n, int m, int b, const double* A, int lda, const double* B, int ldb,
                  double* out) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(n >= 1 && m >= 1 && b >= 1 && m <= lda && b <= ldb && A && B && out, PSOCI_ERR_ARG,
                "bad projection arguments");
    psoci::Linalg L;
    L.stream = c.stream;
    L.gemm_tn_host(n, m, b, A, lda, B, ldb, out);
    return PSOCI_OK;
  });
}

int psoci_combine(psoci_ctx* ctx, int64_t n, int m, int k, const double* A, int lda, const double* Y, double alpha,
                  double* Out, int ldo) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(n >= 1 && m >= 1 && k >= 1 && m <= lda && k <= ldo && A && Y && Out, PSOCI_ERR_ARG,
                "bad combine arguments");
    psoci::Linalg L;
    L.stream = c.stream;
    L.gemm_nn(n, m, k, A, lda, Y, alpha, Out, ldo);
    c.sync();
    return PSOCI_OK;
  });
}

int psoci_device_ptr(psoci_ctx* ctx, int which, void** ptr) {
  return guarded(ctx, [&](Ctx& c) {
    PSOCI_CHECK(ptr, PSOCI_ERR_ARG, "null output");
    PSOCI_CHECK(which == 0 || which == 1, PSOCI_ERR_ARG, "unknown handle");
    if (which == 1) {
      *ptr = (void*)c.stream;
      return PSOCI_OK;
    }
    PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
    if (!c.have_diag) psoci::compute_diag(c);
    *ptr = c.d_diag.p;
    return PSOCI_OK;
  });
}

int psoci_sync(psoci_ctx* ctx) {
  return guarded(ctx, [&](Ctx& c) {
    c.sync();
    return PSOCI_OK;
  });
}

}  // extern "C"
