// gx_hitran.cpp -- HITRAN .par reader (C ABI in include/gx_hitran.h): the reference's
// HITRANStateReader and its driver's filter, over text in memory, records in parallel.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gx_hitran.h"
#include "gx_common.h"
#include "gx_decimal.h"

#pragma STDC FP_CONTRACT OFF

using namespace gx;

namespace {

// Field widths of a HITRAN record (HITRANStateReader.cpp:4, 15).
constexpr int kWidth[19] = {2, 1, 12, 10, 10, 5, 5, 10, 4, 8, 15, 15, 15, 15, 6, 12, 1, 7, 7};
constexpr int kPieceMax = 1023;   // fgets(buffer, 1024): at most 1023 characters per piece

constexpr int field_start(int f) {
  int s = 0;
  for (int i = 0; i < f; ++i) s += kWidth[i];
  return s;
}

// Bytes of the reader's buffer any field it reads can touch: up to the end of field 17.
constexpr int kReach = field_start(17) + kWidth[17];

unsigned threads_for(int nthreads, size_t work);

// The pieces fgets returns, in order: start offset and length (newline included).  Lines
// (up to and including a newline) are independent, and a line longer than 1023 characters
// is cut every 1023, so the text is scanned for newlines in parallel slices.
void split_pieces(const char* t, size_t len, unsigned nt, std::vector<size_t>& start, std::vector<int>& plen) {
  nt = (unsigned)std::max<size_t>(1, std::min<size_t>(nt, len / (1 << 20) + 1));
  std::vector<std::vector<size_t>> ends(nt);                 // offsets just past each newline
  std::vector<std::thread> th;
  auto scan = [&](unsigned k) {
    const size_t a = len * k / nt, z = len * (k + 1) / nt;
    for (size_t p = a; p < z;) {
      const void* nl = std::memchr(t + p, '\n', z - p);
      if (!nl) break;
      const size_t e = (size_t)((const char*)nl - t) + 1;
      ends[k].push_back(e);
      p = e;
    }
  };
  for (unsigned k = 1; k < nt; ++k) th.emplace_back(scan, k);
  scan(0);
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& v : ends) total += v.size();
  start.reserve(total + 1);
  plen.reserve(total + 1);
  auto emit = [&](size_t a, size_t z) {                      // one line [a, z) as fgets pieces
    while (a < z) {
      const size_t n = std::min(z - a, (size_t)kPieceMax);
      start.push_back(a);
      plen.push_back((int)n);
      a += n;
    }
  };
  size_t p = 0;
  for (auto& v : ends)
    for (size_t e : v) {
      emit(p, e);
      p = e;
    }
  emit(p, len);                                              // the last line without a newline
}

// The reader's buffer as record r sees it, over the first kReach bytes: the piece, its
// terminating zero, then what earlier pieces left behind (zero before the first).
struct BufferView {
  const char* t;
  const std::vector<size_t>& start;
  const std::vector<int>& plen;
  void fill(size_t r, char* b) const {
    const int n = plen[r];
    const int m = std::min(n, kReach);
    std::memcpy(b, t + start[r], (size_t)m);
    if (n >= kReach) return;
    b[n] = '\0';
    // positions n+1 .. kReach-1: the latest earlier piece whose bytes (or terminator) reach there
    int need = n + 1;
    for (size_t q = r; need < kReach && q-- > 0;) {
      const int nq = plen[q];
      if (nq < need) continue;
      const int hi = std::min(nq + 1, kReach);          // piece q wrote [0, nq] (its zero at nq)
      for (int p = need; p < hi; ++p) b[p] = p < nq ? t[start[q] + p] : '\0';
      need = hi;
    }
    for (int p = need; p < kReach; ++p) b[p] = '\0';
  }
};

const uint64_t kPow5Words[] = {GX_POW5_WORDS};
const int kPow5Shifts[] = {GX_POW5_SHIFTS};

// GetField (HITRANStateReader.cpp:21-40): the field's bytes in a zeroed 13-byte array, cut at
// 12 characters, read with atof.  A field that is one clean decimal number (after leading
// white space, up to the cut or white space) is converted by the exact decimal routine
// (strtod's correctly rounded value); any other field goes to atof itself.
double field_f(const char* b, int f) {
  char c[13] = {0};
  std::memcpy(c, b + field_start(f), (size_t)kWidth[f]);
  c[12] = '\0';
  const char* p = c;
  while (dec_is_space((unsigned char)*p)) ++p;
  const char* end = p;
  while (*end) ++end;
  double v;
  if (p < end && dec_parse_double(p, end, kPow5Words, kPow5Shifts, &v) == GX_TOK_OK) return v;
  return std::atof(c);
}

int field_i(const char* b, int f) {
  char c[13] = {0};
  std::memcpy(c, b + field_start(f), (size_t)kWidth[f]);
  c[12] = '\0';
  return std::atoi(c);
}

unsigned threads_for(int nthreads, size_t work) {
  unsigned n = nthreads > 0 ? (unsigned)nthreads : std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::max<size_t>(1, std::min<size_t>(n, work / 4096 + 1));
}

template <class F>
void parallel_for(size_t n, unsigned nt, F f) {
  std::vector<std::thread> th;
  auto work = [&](unsigned k) {
    const size_t a = n * k / nt, z = n * (k + 1) / nt;
    for (size_t i = a; i < z; ++i) f(i);
  };
  for (unsigned k = 1; k < nt; ++k) th.emplace_back(work, k);
  work(0);
  for (auto& x : th) x.join();
}

// Fortran Fw.d: the value in w characters with d decimals; a leading "0." loses its zero
// when the field is otherwise too wide, as Fortran writes it.
void fortran_f(char* dst, int w, int d, double v) {
  char s[64];
  int n = std::snprintf(s, sizeof s, "%.*f", d, v);
  if (n > w) {
    char* z = std::strstr(s, "0.");
    if (z && (z == s || (z == s + 1 && s[0] == '-'))) {
      std::memmove(z, z + 1, std::strlen(z + 1) + 1);
      --n;
    }
  }
  if (n > w) {
    std::memset(dst, '*', (size_t)w);
    return;
  }
  std::memset(dst, ' ', (size_t)(w - n));
  std::memcpy(dst + (w - n), s, (size_t)n);
}

}  // namespace

extern "C" {

int gx_hitran_read(const char* text, size_t len, const gx_hitran_opts* o, size_t cap, double* elow, double* nu,
                   double* aif, int* gns, double* gamma, double* nexp, size_t* nlines, size_t* nrecords, int* mol_id,
                   int* iso) {
  if ((len && !text) || !o || !nlines || (cap && (!elow || !nu || !aif || !gns)))
    return fail(GX_INVALID_ARG, "gx_hitran_read: null pointer");
  std::vector<size_t> start;
  std::vector<int> plen;
  split_pieces(text, len, threads_for(o->nthreads, len / 64), start, plen);
  const size_t nrec = start.size();
  const BufferView view{text, start, plen};
  if (mol_id || iso) {
    // the constructor's first fgets (HITRANStateReader.cpp:49-58); an empty text reads nothing
    char b[kReach];
    if (nrec) view.fill(0, b);
    else std::memset(b, 0, sizeof b);
    if (mol_id) *mol_id = field_i(b, 0);
    if (iso) *iso = field_i(b, 1);
  }
  const unsigned nt = threads_for(o->nthreads, nrec);
  // nu of every record, then the read loop's filter in record order
  std::vector<double> nus(nrec);
  parallel_for(nrec, nt, [&](size_t r) {
    char b[kReach];
    view.fill(r, b);
    nus[r] = field_f(b, 2);
  });
  std::vector<size_t> keep;
  size_t read = nrec;
  for (size_t r = 0; r < nrec; ++r) {
    if (nus[r] < o->lo) continue;           // VoigtProfile.cpp:65-66
    if (nus[r] > o->hi) {                   // :67-68, the record counts as read
      read = r + 1;
      break;
    }
    keep.push_back(r);
  }
  *nlines = keep.size();
  if (nrecords) *nrecords = read;
  if (!cap) return GX_OK;
  if (cap < keep.size()) return fail(GX_INVALID_ARG, "gx_hitran_read: output arrays too small");
  // ReadNextState's values for the kept records (HITRANStateReader.cpp:207-247)
  const double planck = 6.6260693e-27, vellgt = 2.99792458e10, boltz = 1.380658e-16;
  const double pi = std::acos(-1.0);
  const double beta = planck * vellgt / (boltz * 296.0);
  parallel_for(keep.size(), threads_for(o->nthreads, keep.size()), [&](size_t k) {
    char b[kReach];
    view.fill(keep[k], b);
    const double v = nus[keep[k]], e = field_f(b, 7);
    double a = field_f(b, 4);
    const double gam_air = field_f(b, 5), self = field_f(b, 6), n = field_f(b, 8);
    double g = field_f(b, 17);
    const double intens = field_f(b, 3);
    const double gam = (gam_air * o->mixture + self * (1.0 - o->mixture)) * o->pressure;
    if (a == 0.0) {
      a = intens * 8.0 * pi * vellgt * o->ref_partition * v * v / (std::exp(-beta * e) * (1.0 - std::exp(-beta * v)));
      g = 1.0;
    }
    elow[k] = e;
    nu[k] = v;
    aif[k] = a;
    gns[k] = (int)g;                        // h_gns[Ntrans] = t_gns (VoigtProfile.cpp:72)
    if (gamma) gamma[k] = gam;
    if (nexp) nexp[k] = n;
  });
  return GX_OK;
}

int gx_format_hitran(size_t n, int mol_id, int iso, const double* nu, const double* intens, const double* aif,
                     const double* gamma_air, const double* gamma_self, const double* elow, const double* nexp,
                     const double* gup, const double* glow, char* out, size_t cap, size_t* written) {
  if (!written || (n && (!nu || !intens || !aif || !gamma_air || !gamma_self || !elow || !nexp || !gup || !glow)))
    return fail(GX_INVALID_ARG, "gx_format_hitran: null pointer");
  *written = 161 * n;
  if (!out) return GX_OK;
  if (cap < 161 * n) return fail(GX_INVALID_ARG, "gx_format_hitran: output buffer too small");
  parallel_for(n, threads_for(0, n), [&](size_t i) {
    char* r = out + 161 * i;
    std::memset(r, ' ', 160);
    char s[64];
    std::snprintf(s, sizeof s, "%2d%1d", mol_id % 100, iso % 10);
    std::memcpy(r, s, 3);
    fortran_f(r + field_start(2), 12, 6, nu[i]);
    std::snprintf(s, sizeof s, "%10.3E", intens[i]);
    std::memcpy(r + field_start(3), s, 10);
    std::snprintf(s, sizeof s, "%10.3E", aif[i]);
    std::memcpy(r + field_start(4), s, 10);
    fortran_f(r + field_start(5), 5, 4, gamma_air[i]);
    fortran_f(r + field_start(6), 5, 3, gamma_self[i]);
    fortran_f(r + field_start(7), 10, 4, elow[i]);
    fortran_f(r + field_start(8), 4, 2, nexp[i]);
    fortran_f(r + field_start(9), 8, 6, 0.0);
    fortran_f(r + field_start(17), 7, 1, gup[i]);
    fortran_f(r + field_start(18), 7, 1, glow[i]);
    r[160] = '\n';
  });
  return GX_OK;
}

}  // extern "C"
