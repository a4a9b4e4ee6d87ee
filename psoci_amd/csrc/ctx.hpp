// Context state of the PSOCI MI355X library (one GPU).
#pragma once
#include "common.hpp"
#include "elem.hpp"
#include "ras.hpp"

namespace psoci {

constexpr int kSpmmBuckets = 5;  // row-length classes of the stored sigma

// Caller-provided device memory for the padded column staging of a sigma
// call (p null: the context's own buffer, grown on demand).
struct Staging {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  Space sp;
  bool have_space = false;
  int nbf_ints = 0;
  double ecore = 0.0;
  bool have_ints = false;
  // fingerprints of the inputs (configuration list, integrals), written into
  // a stored-Hamiltonian checkpoint and checked when one is loaded
  uint64_t space_hash = 0, ints_hash = 0;

  // space on the device
  DevBuf<uint64_t> d_dm, d_sm;
  DevBuf<int8_t> d_nopen;
  DevBuf<int32_t> d_nsef;
  DevBuf<int64_t> d_sefoff;
  DevBuf<uint8_t> d_cph;
  DevBuf<uint16_t> d_pat;
  DevBuf<int32_t> d_patoff;    // 24: two pattern families
  DevBuf<uint8_t> d_cls;       // per configuration: pattern family
  DevBuf<double> d_cmult;      // per configuration: SEF combination sign
  DevBuf<int32_t> d_jsym;      // per configuration
  DevBuf<double> d_one, d_two;

  // stored Hamiltonian: symmetric full CSR (lower part first in each row)
  bool have_h = false;
  int64_t nnz_lower = 0, nnz_full = 0;
  DevBuf<int64_t> d_rowptr;   // full, n+1
  DevBuf<int32_t> d_col;      // full
  DevBuf<double> d_val;       // full FP64
  DevBuf<float> d_val32;      // full FP32
  DevBuf<int32_t> d_lowcnt;   // per row: entries of the lower part (incl. diagonal)
  // stored sigma shape per precision ([0] FP64, [1] FP32): lanes per
  // nonzero, rows per wavefront, and the tuner's best 16-column time (-1: fixed)
  int spmm_g[2] = {1, 1}, spmm_r[2] = {1, 1};
  double spmm_tune_ms[2] = {-1, -1};
  // row-length buckets (kSpmmBuckets length classes): the rows of each in
  // d_brows[spmm_boff[b] .. spmm_boff[b + 1]), a shape per bucket and
  // precision, and whether the bucketed launches are used (set_buckets and
  // the tuner in sigma.hip)
  DevBuf<int32_t> d_brows;
  int64_t spmm_boff[kSpmmBuckets + 1] = {0};
  int spmm_bg[2][kSpmmBuckets] = {}, spmm_br[2][kSpmmBuckets] = {};
  bool spmm_bucketed[2] = {false, false};
  double hbuild_seconds[2] = {-1.0, -1.0};  // last build_h: build proper, shape tuner
  double spmm_bucket_ms[2] = {-1, -1};
  DevBuf<double> d_diag;
  bool have_diag = false;

  // direct sigma: every connected configuration (both directions, J = I
  // included) per configuration I, ascending, and per-determinant alpha
  // open-shell masks
  bool have_pairs = false;
  int64_t nadj = 0;
  int64_t direct_elems = 0;  // structural SEF elements per direct sigma
  int64_t direct_det_pairs = -1;  // connected determinant pairs per direct sigma (-1: not counted yet)
  DevBuf<int64_t> d_adjoff;   // ncfg + 1: first record of each configuration
  DevBuf<PairRec> d_rec;      // one record per ordered connected pair
  DevBuf<double> d_cdiag;     // ncfg: vnt[0] of each diagonal block
  DevBuf<int32_t> d_ndbl;     // ncfg: double-excitation records (first in each segment)
  DevBuf<int32_t> d_nsgl;     // ncfg: single-excitation records (next)
  DevBuf<int32_t> d_order;    // alternative row order of the direct sigma (items when split)
  DevBuf<int32_t> d_items;    // split spaces: the natural order's work items
  // schedule of the direct sigma per precision ([0] FP64, [1] FP32)
  bool direct_order[2] = {false, false};  // use d_order
  bool direct_split = false;  // rows are work items (I << 2 | q): SEFs [64q, 64q + 64) of I
  int64_t direct_nrows = 0;   // rows (work items) per direct sigma launch
  int direct_chunk[2] = {0, 0};         // XCD run length (workgroups)
  double direct_tune_ms[2] = {-1, -1};  // the tuner's best sample time (-1: fixed by the environment)
  DevBuf<ConfRec> d_conf;     // packed configurations
  DevBuf<double> d_xt;        // XT[e][f][o] = (e o|o f)
  DevBuf<int64_t> d_detoff;   // ncfg
  DevBuf<uint64_t> d_detao;   // ndet_total
  DevBuf<int16_t> d_didx;     // determinant index by open-shell pattern (Space::didx)

  // direct sigma engine: requested (0 auto: the string-driven one where the
  // space qualifies, else the configuration records; 1 configuration
  // records; 2 string-driven only) and the string-driven plan (strsig.hip)
  int direct_engine_req = 0;
  bool have_ras = false, ras_tried = false;
  std::string ras_why;
  RasPlan ras;
  DevBuf<uint64_t> d_rasmask;
  DevBuf<int8_t> d_rascls;
  DevBuf<uint32_t> d_rasadd;
  DevBuf<int32_t> d_rasrem;
  DevBuf<RasBlockDev> d_rasblk;
  DevBuf<int16_t> d_rasq, d_raspidx;
  DevBuf<uint64_t> d_rasqm;  // per block and chunk: bits of the first-added orbitals
  DevBuf<RasSef> d_rassef;
  DevBuf<double> d_rasg64;
  DevBuf<float> d_rasg32;
  DevBuf<RasView> d_rasview;
  DevBuf<char> d_rasc, d_rasw;  // determinant vectors [2 ndet + 2][16 FP32 | ras_kc64 FP64], W [wunit][same]
  int ras_kc64 = 8;             // FP64 columns per chunk (16 when memory allows, else 8)
  DevBuf<int32_t> d_rasdidx;    // [dunit]: signed vector row (2 det + negative) of every D entry
  DevBuf<int64_t> d_rastoff;    // nsef + 1: first term of each SEF
  int64_t ras_nterm = 0;        // terms over all SEFs (after merging mirrored pairs)
  DevBuf<int32_t> d_rasterm;    // signed W offsets (bit 31: negative)
  DevBuf<int8_t> d_rasscl;      // per SEF: 0 1/sqrt(2), 1 sqrt(2), 2 one
  DevBuf<int2> d_rasrows;       // per SEF: signed rows of its determinants' coefficients
  void drop_ras() {
    have_ras = ras_tried = false;
    ras = RasPlan();
    d_rasmask.release();
    d_rascls.release();
    d_rasadd.release();
    d_rasrem.release();
    d_rasblk.release();
    d_rasq.release();
    d_rasqm.release();
    d_raspidx.release();
    d_rassef.release();
    d_rasg64.release();
    d_rasg32.release();
    d_rasview.release();
    d_rasc.release();
    d_rasw.release();
    d_rasdidx.release();
    d_rastoff.release();
    d_rasterm.release();
    d_rasscl.release();
    d_rasrows.release();
  }

  // transition Hamiltonian (rows: this list's SEFs; columns: a second
  // list's SEFs placed at this list's entry points), lower by index
  bool have_trans = false;
  int64_t nnz_trans = 0;
  DevBuf<int64_t> d_trow;
  DevBuf<int32_t> d_tcol;
  DevBuf<double> d_tval;

  // scratch
  DevBuf<char> d_pad;  // zero-padded column staging for the direct sigma
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;

  ~Ctx() {
    if (scan_tmp) (void)hipFree(scan_tmp);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void scan(const int64_t* in, int64_t* out, int64_t n) { exclusive_scan_i64(in, out, n, stream, scan_tmp, scan_tmp_bytes); }
  void sync() { PSOCI_HIP(hipStreamSynchronize(stream)); }
};

// kernels-side entry points (hbuild.hip, sigma.hip, ...)
void upload_space(Ctx& c);
void build_pairs(Ctx& c);               // adjacency + determinant masks (direct sigma, adjoint)
void build_h(Ctx& c);                   // stored Hamiltonian
void tune_stored(Ctx& c);               // stored sigma shape autotuner (after build_h)
void set_buckets(Ctx& c);               // the stored matrix's row-length buckets
void compute_diag(Ctx& c);              // diagonal only
void block_on_device(Ctx& c, int64_t I, int64_t J, double* host_out, int* status);
void sigma_stored(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy, const Staging& st = Staging{});
void sigma_direct(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy, const Staging& st = Staging{});
int64_t direct_det_pairs(Ctx& c);
// string-driven direct sigma (strsig.hip): the plan and its device tables
// (false when the space does not qualify; c.ras_why says why), and the sigma
bool build_ras(Ctx& c);
void sigma_ras(Ctx& c, int prec, int k, const void* X, int ldx, void* Y, int ldy);
// one timed sigma of k columns on internal buffers, averaged over reps: ms[5]
// per launch group (scatter, products by kind, gather), flops[3] per kind
void ras_timing(Ctx& c, int prec, int k, int reps, double* ms, double* flops);
// the direct sigma's operator: the string-driven plan where requested and
// possible, else the configuration pair records (throws when engine 2 is
// requested and the space does not qualify)
void build_direct(Ctx& c);
// the engine a direct sigma uses with what is built now (0: none yet)
inline int direct_engine(const Ctx& c) {
  if (c.direct_engine_req != 1 && c.have_ras) return 2;
  if (c.direct_engine_req != 2 && c.have_pairs) return 1;
  return 0;
}
// whether build_direct would build something (the string plan not tried yet
// where it may apply, or no operator at all)
inline bool direct_needs_build(const Ctx& c) {
  if (c.direct_engine_req != 1 && !c.have_ras && !c.ras_tried) return true;
  return direct_engine(c) == 0;
}  // connected determinant pairs per direct sigma (one counting launch, cached)
// staging bytes a sigma call of k columns at strides ldx, ldy needs (0: none)
size_t sigma_staging_bytes(const Ctx& c, int mode, int prec, int k, const void* X, int ldx, const void* Y, int ldy);
// the allocation-free scratch array of the current device (noalloc.hip):
// set up once per device by psoci_ctx_create, then its address and size
int noalloc_init(int device);
void noalloc_scratch(int device, void** p, size_t* bytes);
void get_h_lower(Ctx& c, int64_t* rowptr, int32_t* col, double* val, double* diag);
// installs a stored Hamiltonian given as the full symmetric CSR (lower part
// first in each row, lowcnt entries) and its diagonal (restart.cpp -> hbuild.hip)
void install_h(Ctx& c, const std::vector<int64_t>& rowptr, const std::vector<int32_t>& col,
               const std::vector<double>& val, const std::vector<int32_t>& lowcnt, const std::vector<double>& diag);
// per-configuration energy contributions of k roots (analysis.hip)
void energy_contrib(Ctx& c, int k, const double* dC, const double* d_evals, double* d_deltae);
void save_h(Ctx& c, const char* path);  // stored-Hamiltonian checkpoint (restart.cpp)
void load_h(Ctx& c, const char* path);
void rdm_grad(Ctx& c, int k, const double* dC, const double* w, double* d_g1, double* d_g2);
void build_transition(Ctx& c, const Space& list2, int64_t* nnz);  // transition.hip

struct DavParams;
int davidson(Ctx& c, const psoci_dav_params& p, double* evals, double* evecs, psoci_dav_stats* st);

inline ConfView conf_view_host(const Space& sp, int64_t c, const uint16_t* pat_dev) {
  ConfView v;
  v.d = sp.dm[c];
  v.s = sp.sm[c];
  v.nopen = sp.nopen[c];
  v.ph = sp.cphase[c];
  v.pat = pat_dev + sp.pat_off[12 * sp.cls[c] + v.nopen];
  return v;
}

}  // namespace psoci
