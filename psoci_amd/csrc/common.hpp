// Shared host-side infrastructure of the PSOCI MI355X library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/psoci.h"

namespace psoci {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PSOCI_HIP(call)                                                                              \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess)                                                                            \
      throw ::psoci::Error(PSOCI_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

#define PSOCI_CHECK(cond, code, msg)           \
  do {                                         \
    if (!(cond)) throw ::psoci::Error(code, msg); \
  } while (0)

// Device buffer owned by a context.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw Error(PSOCI_ERR_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  // grow only (keeps the allocation when large enough)
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) PSOCI_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) PSOCI_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// Host description of the determinant/SEF space.
struct Space {
  int nbf = 0, nelec = 0, ksym = 1, jsym0 = 1;
  int64_t ncfg = 0;
  // class of configuration 0 (the only one unless mixed):
  double mult = 1.0;  // SEF combination sign: -1 when jsym and ksym parities differ
  bool odd = false;   // ODD class (jsym class != ksym class): determinants of odd Ms
  bool frac = false;  // odd electron count: one determinant per SEF (no pairing)
  // configurations of more than one jsym: spin-orbit blocks between them;
  // each configuration has its own determinant class (pattern family 0 =
  // that of configuration 0, 1 = the other) and SEF combination sign
  bool mixed = false;
  bool flag0 = false;  // class flag of configuration 0 (class_flag)
  std::vector<int32_t> jsym;
  std::vector<uint8_t> cls;
  std::vector<double> cmult;
  std::vector<uint64_t> dm, sm;   // doubly / singly occupied orbital masks
  std::vector<int8_t> nopen;
  std::vector<int32_t> nsef, ndet;
  std::vector<int64_t> sef_off, det_off;
  std::vector<uint8_t> cphase;
  int64_t nsef_total = 0, ndet_total = 0;
  std::vector<uint16_t> pat;  // determinant patterns, all open-shell counts, families 0 then 1
  int pat_off[24] = {0};      // pat_off[12 f + s]: first pattern of family f, s open shells
  // didx[2047 f + (1 << s) - 1 + p]: index of the determinant with alpha
  // pattern p among those of a family-f configuration with s open shells
  // (-1: not in the class)
  std::vector<int16_t> didx;
  int max_open = 0;
};

// Builds Space from a configuration list; throws Error on unsupported input.
void build_space(Space& sp, int nbf, int nelec, int ksym, int64_t ncfg, const uint8_t* occ, const int32_t* jsym);
// determinant class of a jsym as a flag: N odd -> alpha and beta swapped
// ((jsym + 1) / 2 even), N even -> ODD ((jsym + 1) / 2 != (ksym + 1) / 2)
bool class_flag(int nelec, int ksym, int jsym);

// A configuration file read by the reference's rules (configs.cpp): file
// order, or sorted by (jsym, occupation) with duplicates dropped.  Throws
// Error(PSOCI_ERR_ARG) naming the line; *err_line (if given) receives it.
struct ConfigFile {
  std::string title;
  int nbf = 0, nelec = 0, ksym = 0, nopen = 0;  // nopen: the largest found
  std::vector<uint8_t> occ;                      // [ncfg][nbf], 0/1/2
  std::vector<int32_t> jsym;                     // [ncfg]
};
void read_config_file(const char* path, bool sorted, ConfigFile& cf, int64_t* err_line);

// SIFS-style integral files (sifs.cpp); errors of the context-free entry
// points are kept per thread for psoci_last_error(NULL).
void read_sifs(const char* path, int* nbf, double* one, double* two, double* ecore);
const std::string& ctxfree_error();
void set_ctxfree_error(const std::string& m);

// Runs a context-free entry point: status code out, message kept per thread.
template <class F>
int ctxfree(F&& f) {
  try {
    set_ctxfree_error("");
    f();
    return PSOCI_OK;
  } catch (const Error& e) {
    set_ctxfree_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_ctxfree_error(e.what());
    return PSOCI_ERR_NOMEM;
  }
}

// 64-bit fingerprint of a byte range (8-byte words, multiply-rotate mixing):
// detects a checkpoint written for other inputs and corrupted payloads; not
// a cryptographic hash.
inline uint64_t hash_bytes(const void* p, size_t nbytes, uint64_t h = 0x9e3779b97f4a7c15ull) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  auto mix = [&h](uint64_t w) {
    w *= 0x87c37b91114253d5ull;
    w = (w << 31) | (w >> 33);
    h ^= w * 0x4cf5ad432745937full;
    h = ((h << 27) | (h >> 37)) * 5 + 0x52dce729;
  };
  size_t i = 0;
  for (; i + 8 <= nbytes; i += 8) {
    uint64_t w;
    std::memcpy(&w, b + i, 8);
    mix(w);
  }
  uint64_t tail = 0;
  std::memcpy(&tail, b + i, nbytes - i);
  mix(tail ^ ((uint64_t)(nbytes - i) << 56));
  h ^= nbytes;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  return h;
}

// Library-wide kernel launch helpers (defined in the .hip files).
void exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s, void*& tmp, size_t& tmp_bytes);

}  // namespace psoci
