// SIFS-style integral files: the reference reads the Columbus "moints" file
// through colib's sifrh1 / sifrh2 / sifrd1 / sifrd2 (PsociIntegrals.C:136-330,
// 423-530, 590-727) and keeps only the packed arrays valone / valtwo and the
// first core energy.  colib is not available in source form, so the layout
// below is this library's own (writer and reader designed together); the
// FIELDS and their meaning follow the colib documentation quoted in
// PsociIntegrals.C (header 1: ntitle, nsym, nbft, ninfo, nenrgy, nmap;
// header 2: titles, nbpsy, slabel, info, bfnlab, ietype, energy, imtype,
// map; integral records: num, last, itypea, itypeb, ifmt, ibvtyp, fcore,
// values, labels), and the READER reproduces the reference's processing:
//
//  * 1-e records with itypea = 0 and itypeb != 0 are summed into
//    valone[NNDXF(max) + min - 1] (so T + V + ...; the overlap, itypeb = 0,
//    is skipped); the record's fcore is not kept (PsociIntegrals.C:505-519);
//  * 1-e records with itypea = 2 and itypeb <= 2 (spin-orbit x, y, z)
//    overwrite valone at the same packed index, negated when the first label
//    is the smaller (PsociIntegrals.C:520-532);
//  * 2-e values are stored at canonicalIndexF77(i, j, k, l) - 1, the last
//    write winning (PsociIntegrals.C:716-724);
//  * the core energy is energy[0] (fetchCoreEnergy, PsociIntegrals.C:79-82).
//
// File layout (little endian).  Every record is framed as
// [int32 nbytes][payload][int32 nbytes]:
//   header 1   int32 ntitle, nsym, nbft, ninfo, nenrgy, nmap
//   header 2   char title[ntitle][80]; int32 nbpsy[nsym]; char slabel[nsym][4];
//              int32 info[ninfo]; char bfnlab[nbft][8]; int32 ietype[nenrgy];
//              double energy[nenrgy]; int32 imtype[nmap]; int32 map[nmap][nbft]
//   1-e records, then 2-e records, each
//              int32 num, last, itypea, itypeb, ifmt, ibvtyp; double fcore;
//              double values[num]; int32 labels[num][2 or 4]   (1-based)
// info = {fsplit = 1, l1rec, n1max, l2rec, n2max}: a record holds at most
// n1max (1-e) or n2max (2-e) values.  last = 0: more records of this type;
// 1: last record of this type; 2: last record of the 1-e (or 2-e) part.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/psoci.h"
#include "common.hpp"

namespace psoci {
namespace {

thread_local std::string g_err;

constexpr int kTitle = 80, kSlabel = 4, kBfnlab = 8;
constexpr int kN1max = 4096, kN2max = 8192;

inline int64_t nndaf(int64_t i) { return i * (i + 1) / 2; }
inline int64_t nndxf(int64_t i) { return i * (i - 1) / 2; }
inline int64_t canon_f77(int64_t i, int64_t j, int64_t k, int64_t l) {
  const int64_t ij = nndxf(std::max(i, j)) + std::min(i, j);
  const int64_t kl = nndxf(std::max(k, l)) + std::min(k, l);
  return nndxf(std::max(ij, kl)) + std::min(ij, kl);
}

struct Writer {
  FILE* f = nullptr;
  std::vector<char> rec;
  template <class T>
  void put(const T* p, size_t n) {
    const char* c = reinterpret_cast<const char*>(p);
    rec.insert(rec.end(), c, c + n * sizeof(T));
  }
  void put_str(const std::string& s, int width) {
    std::string t = s.substr(0, width);
    t.resize(width, ' ');
    rec.insert(rec.end(), t.begin(), t.end());
  }
  void flush() {
    const int32_t nb = (int32_t)rec.size();
    PSOCI_CHECK(std::fwrite(&nb, 4, 1, f) == 1 && (rec.empty() || std::fwrite(rec.data(), 1, rec.size(), f) == rec.size()) &&
                    std::fwrite(&nb, 4, 1, f) == 1,
                PSOCI_ERR_ARG, "write failed");
    rec.clear();
  }
};

struct Reader {
  FILE* f = nullptr;
  std::vector<char> rec;
  size_t pos = 0;
  std::string path;
  void next() {
    int32_t nb = 0, nb2 = 0;
    PSOCI_CHECK(std::fread(&nb, 4, 1, f) == 1, PSOCI_ERR_ARG, path + ": unexpected end of file");
    PSOCI_CHECK(nb >= 0 && nb < (1 << 30), PSOCI_ERR_ARG, path + ": bad record length");
    rec.resize(nb);
    PSOCI_CHECK(nb == 0 || std::fread(rec.data(), 1, nb, f) == (size_t)nb, PSOCI_ERR_ARG, path + ": truncated record");
    PSOCI_CHECK(std::fread(&nb2, 4, 1, f) == 1 && nb2 == nb, PSOCI_ERR_ARG, path + ": record framing mismatch");
    pos = 0;
  }
  // (pos <= rec.size() always; the checks divide rather than multiply, so a
  // count read from a damaged file cannot wrap them)
  template <class T>
  void get(T* p, size_t n) {
    PSOCI_CHECK(n <= (rec.size() - pos) / sizeof(T), PSOCI_ERR_ARG, path + ": record shorter than its fields");
    if (n) std::memcpy(p, rec.data() + pos, n * sizeof(T));
    pos += n * sizeof(T);
  }
  void skip(size_t nbytes) {
    PSOCI_CHECK(nbytes <= rec.size() - pos, PSOCI_ERR_ARG, path + ": record shorter than its fields");
    pos += nbytes;
  }
};

// one block of integral records of a single (itypea, itypeb), values with
// their labels; `final_part` marks the last block of the 1-e or 2-e part
void write_block(Writer& w, int nipv, int nmax, int itypea, int itypeb, const std::vector<double>& vals,
                 const std::vector<int32_t>& labs, bool final_part) {
  const size_t n = vals.size();
  size_t done = 0;
  do {
    const int num = (int)std::min<size_t>(nmax, n - done);
    const bool lastrec = done + num >= n;
    const int32_t hdr[6] = {num, lastrec ? (final_part ? 2 : 1) : 0, itypea, itypeb, 0, 0};
    const double fcore = 0.0;
    w.put(hdr, 6);
    w.put(&fcore, 1);
    w.put(vals.data() + done, num);
    w.put(labs.data() + done * nipv, (size_t)num * nipv);
    w.flush();
    done += num;
  } while (done < n);
}

// labels: [nbf][8] basis-function labels (NULL: "mo1", "mo2", ...); two NULL
// (AO overlap files): an empty 2-e part
void write_file(const char* path, int nbf, const double* s1, const double* t1, const double* v1, const double* so,
                const double* two, double ecore, const char* title, const char* labels = nullptr) {
  PSOCI_CHECK(path && nbf > 0 && nbf <= 64 && (two || (s1 && !t1 && !v1 && !so)), PSOCI_ERR_ARG,
              "bad SIFS write arguments");
  PSOCI_CHECK(t1 || v1 || !two, PSOCI_ERR_ARG, "at least one one-electron integral type is needed");
  Writer w;
  w.f = std::fopen(path, "wb");
  PSOCI_CHECK(w.f, PSOCI_ERR_ARG, std::string("cannot create ") + path);
  try {
    const int32_t h1[6] = {1, 1, nbf, 5, 1, 0};
    w.put(h1, 6);
    w.flush();
    w.put_str(title ? title : "psoci_amd integrals", kTitle);
    const int32_t nbpsy = nbf;
    w.put(&nbpsy, 1);
    w.put_str("a", kSlabel);
    const int32_t info[5] = {1, kN1max, kN1max, kN2max, kN2max};
    w.put(info, 5);
    for (int i = 0; i < nbf; ++i)
      w.put_str(labels ? std::string(labels + (size_t)kBfnlab * i, kBfnlab) : "mo" + std::to_string(i + 1), kBfnlab);
    const int32_t ietype = -1;  // the first core energy is the nuclear repulsion
    w.put(&ietype, 1);
    w.put(&ecore, 1);
    w.flush();
    // 1-e part: S (skipped by the reader), T, V, then the spin-orbit parts
    std::vector<std::pair<int, const double*>> blocks;
    if (s1) blocks.push_back({0, s1});
    if (t1) blocks.push_back({1, t1});
    if (v1) blocks.push_back({2, v1});
    const int n1 = (int)nndaf(nbf);
    int nso = so ? 3 : 0;
    for (size_t b = 0; b < blocks.size(); ++b) {
      std::vector<double> vals;
      std::vector<int32_t> labs;
      for (int i = 1; i <= nbf; ++i)
        for (int j = 1; j <= i; ++j) {
          vals.push_back(blocks[b].second[nndxf(i) + j - 1]);
          labs.push_back(i);
          labs.push_back(j);
        }
      write_block(w, 2, kN1max, 0, blocks[b].first, vals, labs, nso == 0 && b + 1 == blocks.size());
    }
    for (int x = 0; x < nso; ++x) {
      // antisymmetric operator: stored with the larger label first
      std::vector<double> vals;
      std::vector<int32_t> labs;
      for (int i = 1; i <= nbf; ++i)
        for (int j = 1; j < i; ++j) {
          vals.push_back(so[(int64_t)x * n1 + nndxf(i) + j - 1]);
          labs.push_back(i);
          labs.push_back(j);
        }
      write_block(w, 2, kN1max, 2, x, vals, labs, x + 1 == nso);
    }
    // 2-e part: every canonical (ij|kl), i >= j, k >= l, ij >= kl
    std::vector<double> vals;
    std::vector<int32_t> labs;
    for (int i = 1; two && i <= nbf; ++i)
      for (int j = 1; j <= i; ++j)
        for (int k = 1; k <= i; ++k)
          for (int l = 1; l <= k; ++l) {
            if (nndxf(k) + l > nndxf(i) + j) continue;
            vals.push_back(two[canon_f77(i, j, k, l) - 1]);
            labs.insert(labs.end(), {i, j, k, l});
          }
    write_block(w, 4, kN2max, 3, 0, vals, labs, true);
  } catch (...) {
    std::fclose(w.f);
    throw;
  }
  PSOCI_CHECK(std::fclose(w.f) == 0, PSOCI_ERR_ARG, std::string("cannot write ") + path);
}

// AO-basis file (the reference's "aoints" for the Mulliken analysis,
// PsociIntegrals::fetchAndProcessOverlapIntegrals / fetchOverlapInts /
// fetchBasisLabels, PsociIntegrals.C:90-93, 967-1110): the 1-e records with
// itypea = 0 and itypeb = 0 summed into the packed overlap, the basis labels
// as stored (8 characters each); the 2-e part is not read.
void read_ao_file(const char* path, int* nbf_out, double* overlap, char* labels) {
  PSOCI_CHECK(path, PSOCI_ERR_ARG, "null path");
  Reader r;
  r.path = path;
  r.f = std::fopen(path, "rb");
  PSOCI_CHECK(r.f, PSOCI_ERR_ARG, std::string("cannot open ") + path);
  try {
    int32_t h1[6];
    r.next();
    r.get(h1, 6);
    const int ntitle = h1[0], nsym = h1[1], nbft = h1[2], ninfo = h1[3], nenrgy = h1[4], nmap = h1[5];
    PSOCI_CHECK(ntitle >= 0 && nsym >= 1 && nbft >= 1 && nbft <= 64 && ninfo >= 5 && nenrgy >= 1 && nmap >= 0,
                PSOCI_ERR_ARG, r.path + ": bad header 1");
    if (nbf_out) *nbf_out = nbft;
    if (overlap || labels) {
      r.next();
      r.skip((size_t)ntitle * kTitle);
      std::vector<int32_t> nbpsy(nsym), info(ninfo);
      r.get(nbpsy.data(), nsym);
      r.skip((size_t)nsym * kSlabel);
      r.get(info.data(), ninfo);
      std::vector<char> lab((size_t)nbft * kBfnlab);
      r.get(lab.data(), lab.size());
      if (labels) std::memcpy(labels, lab.data(), lab.size());
      const int n1max = info[2];
      PSOCI_CHECK(n1max > 0, PSOCI_ERR_ARG, r.path + ": bad info record");
      std::vector<double> s(nndaf(nbft), 0.0), vals;
      std::vector<int32_t> labs;
      int32_t hdr[6];
      do {
        r.next();
        r.get(hdr, 6);
        double fcore;
        r.get(&fcore, 1);
        PSOCI_CHECK(hdr[0] >= 0 && hdr[0] <= n1max, PSOCI_ERR_ARG, r.path + ": record exceeds its declared capacity");
        vals.resize(hdr[0]);
        labs.resize((size_t)hdr[0] * 2);
        r.get(vals.data(), vals.size());
        r.get(labs.data(), labs.size());
        for (int32_t i = 0; i < hdr[0]; ++i) {
          const int64_t a = labs[2 * i], b = labs[2 * i + 1];
          PSOCI_CHECK(a >= 1 && a <= nbft && b >= 1 && b <= nbft, PSOCI_ERR_ARG, r.path + ": orbital label out of range");
          if (hdr[2] == 0 && hdr[3] == 0) s[nndxf(std::max(a, b)) + std::min(a, b) - 1] += vals[i];
        }
      } while (hdr[1] != 2);
      if (overlap) std::memcpy(overlap, s.data(), s.size() * sizeof(double));
    }
  } catch (...) {
    std::fclose(r.f);
    throw;
  }
  std::fclose(r.f);
}

void read_file(const char* path, int* nbf_out, double* one, double* two, double* ecore) {
  PSOCI_CHECK(path, PSOCI_ERR_ARG, "null path");
  Reader r;
  r.path = path;
  r.f = std::fopen(path, "rb");
  PSOCI_CHECK(r.f, PSOCI_ERR_ARG, std::string("cannot open ") + path);
  try {
    int32_t h1[6];
    r.next();
    r.get(h1, 6);
    const int ntitle = h1[0], nsym = h1[1], nbft = h1[2], ninfo = h1[3], nenrgy = h1[4], nmap = h1[5];
    PSOCI_CHECK(ntitle >= 0 && nsym >= 1 && nbft >= 1 && nbft <= 64 && ninfo >= 5 && nenrgy >= 1 && nmap >= 0,
                PSOCI_ERR_ARG, r.path + ": bad header 1");
    if (nbf_out) *nbf_out = nbft;
    if (!one && !two && !ecore) {
      std::fclose(r.f);
      return;
    }
    r.next();
    r.skip((size_t)ntitle * kTitle);
    std::vector<int32_t> nbpsy(nsym), info(ninfo), ietype(nenrgy);
    r.get(nbpsy.data(), nsym);
    r.skip((size_t)nsym * kSlabel);
    r.get(info.data(), ninfo);
    r.skip((size_t)nbft * kBfnlab);
    r.get(ietype.data(), nenrgy);
    std::vector<double> energy(nenrgy);
    r.get(energy.data(), nenrgy);
    const int n1max = info[2], n2max = info[4];
    PSOCI_CHECK(n1max > 0 && n2max > 0, PSOCI_ERR_ARG, r.path + ": bad info record");
    std::vector<double> v1(nndaf(nbft), 0.0), v2(nndaf(nndaf(nbft)), 0.0);
    auto record = [&](int nipv, int nmax, int32_t* hdr, std::vector<double>& vals, std::vector<int32_t>& labs) {
      r.next();
      r.get(hdr, 6);
      double fcore;
      r.get(&fcore, 1);  // read and discarded, as the reference does
      PSOCI_CHECK(hdr[0] >= 0 && hdr[0] <= nmax, PSOCI_ERR_ARG, r.path + ": record exceeds its declared capacity");
      vals.resize(hdr[0]);
      labs.resize((size_t)hdr[0] * nipv);
      r.get(vals.data(), vals.size());
      r.get(labs.data(), labs.size());
      for (int32_t lab : labs)
        PSOCI_CHECK(lab >= 1 && lab <= nbft, PSOCI_ERR_ARG, r.path + ": orbital label out of range");
    };
    std::vector<double> vals;
    std::vector<int32_t> labs;
    int32_t hdr[6];
    do {
      record(2, n1max, hdr, vals, labs);
      const int itypea = hdr[2], itypeb = hdr[3];
      for (int32_t i = 0; i < hdr[0]; ++i) {
        const int64_t a = labs[2 * i], b = labs[2 * i + 1];
        if (itypea == 0 && itypeb != 0) {
          v1[nndxf(std::max(a, b)) + std::min(a, b) - 1] += vals[i];
        } else if (itypea == 2 && itypeb <= 2) {
          if (a < b)
            v1[nndxf(b) + a - 1] = -vals[i];
          else
            v1[nndxf(a) + b - 1] = vals[i];
        }
      }
    } while (hdr[1] != 2);
    do {
      record(4, n2max, hdr, vals, labs);
      for (int32_t i = 0; i < hdr[0]; ++i)
        v2[canon_f77(labs[4 * i], labs[4 * i + 1], labs[4 * i + 2], labs[4 * i + 3]) - 1] = vals[i];
    } while (hdr[1] != 2);
    if (one) std::memcpy(one, v1.data(), v1.size() * sizeof(double));
    if (two) std::memcpy(two, v2.data(), v2.size() * sizeof(double));
    if (ecore) *ecore = energy[0];
  } catch (...) {
    std::fclose(r.f);
    throw;
  }
  std::fclose(r.f);
}

}  // namespace

const std::string& ctxfree_error() { return g_err; }
void set_ctxfree_error(const std::string& m) { g_err = m; }

void read_sifs(const char* path, int* nbf, double* one, double* two, double* ecore) {
  read_file(path, nbf, one, two, ecore);
}

}  // namespace psoci

extern "C" {

int psoci_sifs_write(const char* path, int nbf, const double* s1, const double* t1, const double* v1,
                     const double* so, const double* valtwo, double ecore, const char* title) {
  return psoci::ctxfree([&] { psoci::write_file(path, nbf, s1, t1, v1, so, valtwo, ecore, title); });
}

int psoci_sifs_read(const char* path, int* nbf, double* valone, double* valtwo, double* ecore) {
  return psoci::ctxfree([&] { psoci::read_file(path, nbf, valone, valtwo, ecore); });
}

int psoci_sifs_write_ao(const char* path, int nbf, const double* overlap, const char* labels, const char* title) {
  return psoci::ctxfree([&] {
    PSOCI_CHECK(overlap, PSOCI_ERR_ARG, "null overlap");
    psoci::write_file(path, nbf, overlap, nullptr, nullptr, nullptr, nullptr, 0.0, title, labels);
  });
}

int psoci_sifs_read_ao(const char* path, int* nbf, double* overlap, char* labels) {
  return psoci::ctxfree([&] { psoci::read_ao_file(path, nbf, overlap, labels); });
}

}  // extern "C"
