// Configuration files: the reference's two readers with their per-line
// validation, PsociConfigs::readConfigsToGA (file order, PsociConfigs.C:
// 336-495, the production path) and PsociConfigs::readConfigs (sorted by
// (jsym, occupation string), duplicates dropped, PsociConfigs.C:143-300).
//
// Format: a title line; "nbf nelec ksym ncfg nopen" (the first five
// whitespace-separated integers of line 2); then one "jsym occupation" line
// per configuration, the occupation a string of nbf characters 0/1/2.  Every
// line after the header is a configuration line, as in the reference's
// getline loop (a blank line is an error there too).  A malformed line stops
// the read with PSOCI_ERR_ARG and its 1-based line number (the title is line
// 1, the header line 2), matching the reference's return value numcfgs + 2.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/psoci.h"
#include "common.hpp"

namespace psoci {
namespace {

[[noreturn]] void bad_line(int64_t line, int64_t* err_line, const std::string& msg) {
  if (err_line) *err_line = line;
  throw Error(PSOCI_ERR_ARG, "configuration file line " + std::to_string(line) + ": " + msg);
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

struct Tok {
  const char *b = nullptr, *e = nullptr;
  size_t size() const { return (size_t)(e - b); }
  std::string str() const { return std::string(b, e); }
};

// the next whitespace-separated token of [p, e), advancing p (empty at the end)
inline Tok next_tok(const char*& p, const char* e) {
  while (p < e && is_space(*p)) ++p;
  Tok t;
  t.b = p;
  while (p < e && !is_space(*p)) ++p;
  t.e = p;
  return t;
}

// a whole token as an int: optional sign, decimal digits, in range
bool to_int(const Tok& t, int* v) {
  const char* p = t.b;
  bool neg = false;
  if (p < t.e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  if (p == t.e) return false;
  long long x = 0;
  for (; p < t.e; ++p) {
    if (*p < '0' || *p > '9') return false;
    x = x * 10 + (*p - '0');
    if (x > (long long)INT_MAX + 1) return false;
  }
  if (neg) x = -x;
  if (x < INT_MIN || x > INT_MAX) return false;
  *v = (int)x;
  return true;
}

}  // namespace

void read_config_file(const char* path, bool sorted, ConfigFile& cf, int64_t* err_line) {
  if (err_line) *err_line = 0;
  PSOCI_CHECK(path, PSOCI_ERR_ARG, "null path");
  FILE* f = std::fopen(path, "rb");
  PSOCI_CHECK(f, PSOCI_ERR_ARG, std::string("cannot open configuration file ") + path);
  std::string text;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  PSOCI_CHECK(ok, PSOCI_ERR_ARG, std::string("read error on ") + path);

  // lines split at '\n'; a final newline ends the last line, it starts none
  std::vector<std::pair<const char*, const char*>> lines;
  const char* p = text.data();
  const char* end = p + text.size();
  while (p < end) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', end - p));
    const char* le = nl ? nl : end;
    lines.emplace_back(p, (le > p && le[-1] == '\r') ? le - 1 : le);
    p = nl ? nl + 1 : end;
  }
  if (lines.empty()) bad_line(1, err_line, "missing title");
  cf.title.assign(lines[0].first, lines[0].second);

  int hdr[5];
  {
    if (lines.size() < 2) bad_line(2, err_line, "expected 5 integers: nbf nelec ksym ncfg nopen");
    const char* p = lines[1].first;
    bool good = true;
    for (int i = 0; good && i < 5; ++i) good = to_int(next_tok(p, lines[1].second), &hdr[i]);
    if (!good) bad_line(2, err_line, "expected 5 integers: nbf nelec ksym ncfg nopen");
  }
  cf.nbf = hdr[0];
  cf.nelec = hdr[1];
  cf.ksym = hdr[2];
  const int64_t ncfg_header = hdr[3];
  const int nopen_header = hdr[4];

  const int nbf = cf.nbf;
  int max_open = 0;
  const int64_t nline = (int64_t)lines.size();
  cf.occ.assign((size_t)std::max<int64_t>(nline - 2, 0) * (size_t)std::max(nbf, 0), 0);
  cf.jsym.clear();
  cf.jsym.reserve((size_t)std::max<int64_t>(nline - 2, 0));
  for (int64_t n = 2; n < nline; ++n) {
    const int64_t ln = n + 1;
    const char* p = lines[n].first;
    const Tok ts = next_tok(p, lines[n].second), to = next_tok(p, lines[n].second);
    if (!to.size()) bad_line(ln, err_line, "expected 'jsym occupation'");
    int sym;
    if (!to_int(ts, &sym)) bad_line(ln, err_line, "bad symmetry '" + ts.str() + "'");
    if ((int64_t)to.size() != nbf)
      bad_line(ln, err_line, "configuration length " + std::to_string(to.size()) + " != nbf " + std::to_string(nbf));
    uint8_t* row = cf.occ.data() + (size_t)(n - 2) * nbf;
    int nop = 0, nel = 0;
    for (int o = 0; o < nbf; ++o) {
      const char ch = to.b[o];
      if (ch < '0' || ch > '2') bad_line(ln, err_line, "occupations must be 0, 1 or 2");
      nop += ch == '1';
      nel += ch - '0';
      row[o] = (uint8_t)(ch - '0');
    }
    if (nop > nopen_header)
      bad_line(ln, err_line, std::to_string(nop) + " open shells, the header allows " + std::to_string(nopen_header));
    if (nel != cf.nelec)
      bad_line(ln, err_line, "electron count " + std::to_string(nel) + " != " + std::to_string(cf.nelec));
    max_open = std::max(max_open, nop);
    cf.jsym.push_back(sym);
  }
  // the reference adjusts the open-shell maximum down to what it found
  cf.nopen = max_open;

  const int64_t nread = (int64_t)cf.jsym.size();
  bool dropped = false;
  if (sorted) {
    // std::set<pair<int, string>> order: jsym first, then the occupation
    // string (its characters compare as the 0/1/2 bytes do); equal pairs kept once
    std::vector<int64_t> ord(nread);
    std::iota(ord.begin(), ord.end(), 0);
    const uint8_t* o = cf.occ.data();
    auto less = [&](int64_t a, int64_t b) {
      if (cf.jsym[a] != cf.jsym[b]) return cf.jsym[a] < cf.jsym[b];
      return std::memcmp(o + a * nbf, o + b * nbf, nbf) < 0;
    };
    std::stable_sort(ord.begin(), ord.end(), less);
    std::vector<uint8_t> occ2;
    std::vector<int32_t> js2;
    occ2.reserve(cf.occ.size());
    js2.reserve(nread);
    for (int64_t i = 0; i < nread; ++i) {
      const int64_t c = ord[i];
      if (i && !less(ord[i - 1], c)) {
        dropped = true;
        continue;
      }
      occ2.insert(occ2.end(), o + c * nbf, o + (c + 1) * nbf);
      js2.push_back(cf.jsym[c]);
    }
    cf.occ.swap(occ2);
    cf.jsym.swap(js2);
  }
  // with duplicates dropped the count cannot be predicted: the sorted
  // reader accepts what is left (PsociConfigs.C:283-296)
  const int64_t nkept = (int64_t)cf.jsym.size();
  if (nkept != ncfg_header && !dropped)
    bad_line(nline, err_line,
             "found " + std::to_string(nkept) + " configurations, header says " + std::to_string(ncfg_header));
}

}  // namespace psoci

extern "C" {

int psoci_configs_read(const char* path, int order, int* nbf, int* nelec, int* ksym, int64_t* ncfg, int* nopen,
                       uint8_t* occ, int32_t* jsym, char* title, int title_cap, int64_t* err_line) {
  return psoci::ctxfree([&] {
    PSOCI_CHECK(order == PSOCI_CONFIGS_FILE_ORDER || order == PSOCI_CONFIGS_SORTED, PSOCI_ERR_ARG,
                "order must be PSOCI_CONFIGS_FILE_ORDER or PSOCI_CONFIGS_SORTED");
    psoci::ConfigFile cf;
    psoci::read_config_file(path, order == PSOCI_CONFIGS_SORTED, cf, err_line);
    if (nbf) *nbf = cf.nbf;
    if (nelec) *nelec = cf.nelec;
    if (ksym) *ksym = cf.ksym;
    if (ncfg) *ncfg = (int64_t)cf.jsym.size();
    if (nopen) *nopen = cf.nopen;
    if (occ) std::memcpy(occ, cf.occ.data(), cf.occ.size());
    if (jsym) std::memcpy(jsym, cf.jsym.data(), cf.jsym.size() * sizeof(int32_t));
    if (title && title_cap > 0) {
      const size_t m = std::min(cf.title.size(), (size_t)title_cap - 1);
      std::memcpy(title, cf.title.data(), m);
      title[m] = '\0';
    }
  });
}

}  // extern "C"
