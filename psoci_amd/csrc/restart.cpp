// Restart files: CI vectors in the reference's binary vector format, and a
// checkpoint of the stored Hamiltonian.
//
// CI vectors (PsociVector::scalable_WriteHeaderBurstBinary /
// WriteBurstCPPversionBinaryScalable / WriteBurstScalableBinaryTrailer,
// PsociVector.C:530-553, 627-643, 685-693, as driven by
// PsociGAvectors::dumpVectorScalable, PsociGAvectors.C:207-325, and read back
// by readVectorScalable + specifyDiagonalGuessVectors, PsociGAvectors.C:
// 539-659, 687-726), little endian, no record framing:
//   int64 nbasis, int64 nroots, int32 width, int32 prec      (header)
//   per root r = 1 .. nroots: int64 r, double c[nbasis]       (the vector)
//   the title's bytes, no terminator, to the end of the file  (trailer)
// width / prec are the reference's ASCII formatting defaults (18, 14; unused
// by the binary reader).  The reference writes the root number from a 4-byte
// int as 8 bytes, so its upper half is whatever followed the int in memory;
// like the reference, the reader skips the field.  Reading keeps
// min(nroots, k) roots, requires nbasis to equal the space dimension (the
// reference aborts on a mismatch), and completes the remaining columns j of
// the k-column block with unit vectors e_j (specifyDiagonalGuessVectors).
//
// Stored Hamiltonian checkpoint (replaces the DRA files of cimat / icicol /
// number / diag of PsociConstructHamiltonian.C:56-700, whose layout is a
// property of Global Arrays; this layout is the library's own):
//   char magic[8] = "PSOCIHCK", int32 version = 1, int32 0,
//   int64 n, int64 ncfg, int64 nnz (lower triangle + diagonal),
//   uint64 configuration fingerprint, uint64 integral fingerprint,
//   int32 lanes[2], int32 rows[2], double tune_ms[2]   (stored-sigma shapes),
//   int64 rowptr[n + 1], int32 col[nnz], double val[nnz], double diag[n],
//   uint64 fingerprint of the four arrays.
// Rows are the reference's (packSefsef, PsociGAhamiltonian.C:4077-4136):
// columns ascending, the diagonal last.  Loading checks the fingerprints
// against the context's inputs and the payload, validates the pattern, then
// rebuilds the symmetric full matrix the stored sigma runs on (row r: its
// lower part, then the transposed elements (i, r), i > r, ascending).  A
// load that fails these checks leaves the context unchanged.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace psoci {
namespace {

constexpr int32_t kWidth = 18, kPrec = 14;  // MAXWIDTH / MAXDECIMAL (PsociVector.hpp:39-40)
const char kDefaultTitle[] = "default";     // PsociVector::setFilename's title

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : path(p ? p : "") {
    PSOCI_CHECK(p && *p, PSOCI_ERR_ARG, "empty file name");
    f = std::fopen(p, mode);
    PSOCI_CHECK(f, PSOCI_ERR_ARG, std::string("cannot open ") + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  void close_checked() {
    const int rc = std::fclose(f);
    f = nullptr;
    PSOCI_CHECK(rc == 0, PSOCI_ERR_ARG, "write failed: " + path);
  }
  template <class T>
  void put(const T* p, size_t n) {
    PSOCI_CHECK(n == 0 || std::fwrite(p, sizeof(T), n, f) == n, PSOCI_ERR_ARG, "write failed: " + path);
  }
  template <class T>
  void get(T* p, size_t n, const char* what) {
    PSOCI_CHECK(n == 0 || std::fread(p, sizeof(T), n, f) == n, PSOCI_ERR_ARG,
                path + ": file ends inside " + what);
  }
  int64_t size() {
    const long here = std::ftell(f);
    std::fseek(f, 0, SEEK_END);
    const long end = std::ftell(f);
    std::fseek(f, here, SEEK_SET);
    return end;
  }
};

struct VecHeader {
  int64_t nbasis = 0, nroots = 0;
  int32_t width = 0, prec = 0;
};

VecHeader read_vec_header(File& in) {
  VecHeader h;
  in.get(&h.nbasis, 1, "the header");
  in.get(&h.nroots, 1, "the header");
  in.get(&h.width, 1, "the header");
  in.get(&h.prec, 1, "the header");
  // bounds first, so that the size below cannot overflow: nroots * nbasis <= 2^40
  PSOCI_CHECK(h.nbasis > 0 && h.nbasis <= (int64_t(1) << 40) && h.nroots >= 0 &&
                  h.nroots <= (int64_t(1) << 40) / h.nbasis,
              PSOCI_ERR_ARG, in.path + ": bad vector header");
  const int64_t need = 24 + h.nroots * (8 + 8 * h.nbasis);
  PSOCI_CHECK(in.size() >= need, PSOCI_ERR_ARG,
              in.path + ": file shorter than its header says (" + std::to_string(h.nroots) + " roots of " +
                  std::to_string(h.nbasis) + ")");
  return h;
}

void civec_write(const char* path, int64_t n, int k, const double* C, const char* title) {
  PSOCI_CHECK(n > 0 && k >= 0 && (C || k == 0), PSOCI_ERR_ARG, "bad vector arguments");
  File out(path, "wb");
  const int64_t head[2] = {n, (int64_t)k};
  const int32_t prec[2] = {kWidth, kPrec};
  out.put(head, 2);
  out.put(prec, 2);
  std::vector<double> col((size_t)n);
  for (int r = 0; r < k; ++r) {
    const int64_t num = r + 1;
    out.put(&num, 1);
    for (int64_t i = 0; i < n; ++i) col[i] = C[i * k + r];
    out.put(col.data(), (size_t)n);
  }
  const std::string t = title ? title : kDefaultTitle;
  out.put(t.data(), t.size());
  out.close_checked();
}

void civec_info(const char* path, int64_t* nbasis, int64_t* nroots, char* title, int64_t cap) {
  File in(path, "rb");
  const VecHeader h = read_vec_header(in);
  if (nbasis) *nbasis = h.nbasis;
  if (nroots) *nroots = h.nroots;
  if (title && cap > 0) {
    // the trailer runs to the end of the file (the reference reads it with
    // getline, so a newline ends it)
    std::fseek(in.f, 24 + h.nroots * (8 + 8 * h.nbasis), SEEK_SET);
    int64_t m = 0;
    int ch;
    while (m + 1 < cap && (ch = std::fgetc(in.f)) != EOF && ch != '\n') title[m++] = (char)ch;
    title[m] = '\0';
  }
}

void civec_read(const char* path, int64_t n, int k, double* C, int* nread) {
  PSOCI_CHECK(n > 0 && k > 0 && C && k <= n, PSOCI_ERR_ARG, "bad vector arguments");
  File in(path, "rb");
  const VecHeader h = read_vec_header(in);
  PSOCI_CHECK(h.nbasis == n, PSOCI_ERR_ARG,
              in.path + ": vectors of length " + std::to_string(h.nbasis) + " do not conform to a space of " +
                  std::to_string(n) + " SEFs");
  const int keep = (int)std::min<int64_t>(h.nroots, k);
  std::vector<double> col((size_t)n);
  for (int r = 0; r < keep; ++r) {
    int64_t num = 0;
    in.get(&num, 1, "a root number");
    in.get(col.data(), (size_t)n, "a vector");
    for (int64_t i = 0; i < n; ++i) C[i * k + r] = col[i];
  }
  for (int r = keep; r < k; ++r)
    for (int64_t i = 0; i < n; ++i) C[i * k + r] = i == r ? 1.0 : 0.0;
  if (nread) *nread = keep;
}

struct HHeader {
  char magic[8];
  int32_t version, pad;
  int64_t n, ncfg, nnz;
  uint64_t space_hash, ints_hash;
  int32_t lanes[2], rows[2];
  double tune_ms[2];
};
static_assert(sizeof(HHeader) == 88, "checkpoint header layout");
// version 2 follows the header with the stored sigma's row-length bucket
// shapes (version 1 files have none: one shape for all rows)
struct HBuckets {
  int32_t used[2];
  int32_t lanes[2][kSpmmBuckets], rows[2][kSpmmBuckets];
};
static_assert(sizeof(HBuckets) == 88, "checkpoint bucket block layout");
const char kMagic[8] = {'P', 'S', 'O', 'C', 'I', 'H', 'C', 'K'};

uint64_t payload_hash(const std::vector<int64_t>& rp, const std::vector<int32_t>& col, const std::vector<double>& val,
                      const std::vector<double>& diag) {
  uint64_t h = hash_bytes(rp.data(), rp.size() * 8);
  h = hash_bytes(col.data(), col.size() * 4, h);
  h = hash_bytes(val.data(), val.size() * 8, h);
  return hash_bytes(diag.data(), diag.size() * 8, h);
}

}  // namespace

void save_h(Ctx& c, const char* path) {
  PSOCI_CHECK(c.have_h, PSOCI_ERR_STATE, "no stored Hamiltonian");
  const int64_t n = c.sp.nsef_total;
  std::vector<int64_t> rp(n + 1);
  get_h_lower(c, rp.data(), nullptr, nullptr, nullptr);
  const int64_t nnz = rp[n];
  std::vector<int32_t> col(nnz);
  std::vector<double> val(nnz), diag(n);
  get_h_lower(c, nullptr, col.data(), val.data(), diag.data());
  HHeader h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = 2;
  h.n = n;
  h.ncfg = c.sp.ncfg;
  h.nnz = nnz;
  h.space_hash = c.space_hash;
  h.ints_hash = c.ints_hash;
  for (int p = 0; p < 2; ++p) {
    h.lanes[p] = c.spmm_g[p];
    h.rows[p] = c.spmm_r[p];
    h.tune_ms[p] = c.spmm_tune_ms[p];
  }
  HBuckets hb{};
  for (int p = 0; p < 2; ++p) {
    hb.used[p] = c.spmm_bucketed[p] ? 1 : 0;
    for (int b = 0; b < kSpmmBuckets; ++b) {
      hb.lanes[p][b] = c.spmm_bg[p][b];
      hb.rows[p][b] = c.spmm_br[p][b];
    }
  }
  const uint64_t sum = payload_hash(rp, col, val, diag);
  File out(path, "wb");
  out.put(&h, 1);
  out.put(&hb, 1);
  out.put(rp.data(), rp.size());
  out.put(col.data(), col.size());
  out.put(val.data(), val.size());
  out.put(diag.data(), diag.size());
  out.put(&sum, 1);
  out.close_checked();
}

void load_h(Ctx& c, const char* path) {
  PSOCI_CHECK(c.have_space && c.have_ints, PSOCI_ERR_STATE, "configurations and integrals are required");
  File in(path, "rb");
  HHeader h{};
  in.get(&h, 1, "the header");
  PSOCI_CHECK(std::memcmp(h.magic, kMagic, 8) == 0 && (h.version == 1 || h.version == 2), PSOCI_ERR_ARG,
              in.path + ": not a stored-Hamiltonian checkpoint");
  HBuckets hb{};
  const int64_t hsize = (int64_t)sizeof(HHeader) + (h.version == 2 ? (int64_t)sizeof(HBuckets) : 0);
  const int64_t n = c.sp.nsef_total;
  PSOCI_CHECK(h.n == n && h.ncfg == c.sp.ncfg && h.space_hash == c.space_hash, PSOCI_ERR_ARG,
              in.path + ": written for another configuration list");
  PSOCI_CHECK(h.ints_hash == c.ints_hash, PSOCI_ERR_ARG, in.path + ": written for other integrals");
  const int64_t nnz = h.nnz;
  // nnz bounded first (n < 2^31, nnz <= 2^40), so the file size below cannot overflow
  PSOCI_CHECK(nnz >= n && nnz <= n * (n + 1) / 2 && nnz <= (int64_t(1) << 40) &&
                  in.size() == hsize + 8 * (n + 1) + 12 * nnz + 8 * n + 8,
              PSOCI_ERR_ARG, in.path + ": file size does not match its header");
  if (h.version == 2) in.get(&hb, 1, "the bucket shapes");
  std::vector<int64_t> rp(n + 1);
  std::vector<int32_t> col(nnz);
  std::vector<double> val(nnz), diag(n);
  uint64_t sum = 0;
  in.get(rp.data(), rp.size(), "rowptr");
  in.get(col.data(), col.size(), "col");
  in.get(val.data(), val.size(), "val");
  in.get(diag.data(), diag.size(), "diag");
  in.get(&sum, 1, "the checksum");
  PSOCI_CHECK(sum == payload_hash(rp, col, val, diag), PSOCI_ERR_ARG, in.path + ": checksum mismatch (corrupted file)");
  // the reference's pattern: columns ascending, below the row, diagonal last
  PSOCI_CHECK(rp[0] == 0 && rp[n] == nnz, PSOCI_ERR_ARG, in.path + ": bad row pointers");
  std::vector<int32_t> lowcnt(n), upcnt(n, 0);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t a = rp[r], b = rp[r + 1];
    PSOCI_CHECK(b > a && b <= nnz && col[b - 1] == r && val[b - 1] == diag[r], PSOCI_ERR_ARG,
                in.path + ": row " + std::to_string(r) + " does not end with its diagonal");
    for (int64_t e = a; e + 1 < b; ++e) {
      PSOCI_CHECK(col[e] >= 0 && col[e] < col[e + 1], PSOCI_ERR_ARG,
                  in.path + ": row " + std::to_string(r) + " columns not ascending");
      ++upcnt[col[e]];
    }
    lowcnt[r] = (int32_t)(b - a);
  }
  // full symmetric matrix
  std::vector<int64_t> frp(n + 1);
  frp[0] = 0;
  for (int64_t r = 0; r < n; ++r) frp[r + 1] = frp[r] + lowcnt[r] + upcnt[r];
  const int64_t fnnz = frp[n];
  std::vector<int32_t> fcol(fnnz);
  std::vector<double> fval(fnnz);
  std::vector<int64_t> cur(n);
  for (int64_t r = 0; r < n; ++r) {
    std::copy(col.begin() + rp[r], col.begin() + rp[r + 1], fcol.begin() + frp[r]);
    std::copy(val.begin() + rp[r], val.begin() + rp[r + 1], fval.begin() + frp[r]);
    cur[r] = frp[r] + lowcnt[r];
  }
  for (int64_t r = 0; r < n; ++r)  // rows ascending: each upper part fills in ascending column order
    for (int64_t e = rp[r]; e + 1 < rp[r + 1]; ++e) {
      const int64_t q = col[e];
      fcol[cur[q]] = (int32_t)r;
      fval[cur[q]++] = val[e];
    }
  install_h(c, frp, fcol, fval, lowcnt, diag);
  auto lanes_ok = [](int g) { return g == 1 || g == 2 || g == 4; };
  auto rows_ok = [](int r) { return r == 1 || r == 2 || r == 4 || r == 8; };
  bool shapes_ok = true;
  for (int p = 0; p < 2; ++p) {
    shapes_ok = shapes_ok && lanes_ok(h.lanes[p]) && rows_ok(h.rows[p]);
    if (h.version == 2 && hb.used[p]) {
      shapes_ok = shapes_ok && hb.used[p] == 1;
      for (int b = 0; b < kSpmmBuckets; ++b) shapes_ok = shapes_ok && lanes_ok(hb.lanes[p][b]) && rows_ok(hb.rows[p][b]);
    }
  }
  if (shapes_ok && !std::getenv("PSOCI_SPMM_SHAPE")) {
    // the writer's tuned stored-sigma shapes (and row-length buckets, which
    // depend on the matrix alone): the same kernels, bit for bit
    set_buckets(c);
    for (int p = 0; p < 2; ++p) {
      c.spmm_g[p] = h.lanes[p];
      c.spmm_r[p] = h.rows[p];
      c.spmm_tune_ms[p] = h.tune_ms[p];
      c.spmm_bucketed[p] = h.version == 2 && hb.used[p] == 1;
      for (int b = 0; b < kSpmmBuckets; ++b) {
        c.spmm_bg[p][b] = h.version == 2 ? hb.lanes[p][b] : 1;
        c.spmm_br[p][b] = h.version == 2 ? hb.rows[p][b] : 1;
      }
    }
  } else {
    tune_stored(c);
  }
}

}  // namespace psoci

extern "C" {

int psoci_civec_write(const char* path, int64_t n, int k, const double* C, const char* title) {
  return psoci::ctxfree([&] { psoci::civec_write(path, n, k, C, title); });
}

int psoci_civec_info(const char* path, int64_t* nbasis, int64_t* nroots, char* title, int64_t title_cap) {
  return psoci::ctxfree([&] { psoci::civec_info(path, nbasis, nroots, title, title_cap); });
}

int psoci_civec_read(const char* path, int64_t n, int k, double* C, int* nread) {
  return psoci::ctxfree([&] { psoci::civec_read(path, n, k, C, nread); });
}

}  // extern "C"
