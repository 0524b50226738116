// gx_ingest.cpp -- host side of the ExoMol reader (C ABI in include/gx_ingest.h).
//
// .states and .broad text are small (one line per state or per broadening key) and are
// read on the host with the reference's own C library calls, so their values are the
// reference's by construction.  .trans text (the bulk: one record per line, 1e8-1e10
// lines) is tokenised and converted on the GPU (gx_ingest_kernels.hip); the rare tokens the
// device parser hands back are read here with an istream, as the reference reads all of
// them (ExomolStateReader.cpp:145).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <istream>
#include <limits>
#include <map>
#include <sstream>
#include <streambuf>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gx_ingest.h"
#include "gx_common.h"
#include "gx_decimal.h"
#include "gx_ingest_internal.h"

using namespace gx;

namespace {

constexpr size_t kPieceBytes = size_t(256) << 20;   // text per device pass

// fgets(buf, 1024, f) pieces of the text (exomol_functions.cpp:15,28): at most 1023
// characters, ending after a newline when one comes first.
void states_pieces(const char* text, size_t len, std::vector<size_t>* starts) {
  size_t s = 0;
  while (s < len) {
    const void* nl = std::memchr(text + s, '\n', std::min<size_t>(len - s, 1023));
    const size_t piece = nl ? (size_t)((const char*)nl - (text + s)) + 1 : std::min<size_t>(len - s, 1023);
    starts->push_back(s);
    s += piece;
  }
  starts->push_back(len);
}

size_t states_count(const char* text, size_t len) {
  size_t n = 0, s = 0;
  while (s < len) {
    const void* nl = std::memchr(text + s, '\n', std::min<size_t>(len - s, 1023));
    s += nl ? (size_t)((const char*)nl - (text + s)) + 1 : std::min<size_t>(len - s, 1023);
    ++n;
  }
  return n;
}

// A std::istream over borrowed bytes (no copy).
struct MemBuf : std::streambuf {
  MemBuf(const char* p, size_t n) {
    char* b = const_cast<char*>(p);
    setg(b, b, b + n);
  }
  size_t consumed() const { return (size_t)(gptr() - eback()); }
};

int hip_check(hipError_t e, const char* what) { return e == hipSuccess ? GX_OK : hip_fail(e, what); }

template <typename T>
int dev_alloc(T** p, size_t n) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (n == 0) return GX_OK;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e == hipErrorOutOfMemory) return fail(GX_OUT_OF_MEMORY, "ingest: out of device memory");
  return hip_check(e, "ingest: hipMalloc");
}

template <typename T>
void dev_free(T*& p) {
  if (p) hipFree(p);
  p = nullptr;
}

}  // namespace

struct gx_ingest {
  int device = 0;
  hipStream_t s = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // states, on the host for records read here and in HBM for the kernels
  size_t nstates = 0;
  std::vector<double> hE;
  std::vector<int> hG, hJ;
  double* dE = nullptr;
  int* dG = nullptr;
  int* dJ = nullptr;

  // broadening: (Ji, Jf) table of value indices, values (gamma * P, n)
  bool use_broad = false;
  int jlo = 0, jn = 0;
  std::vector<int> hBtab;
  std::vector<double2> hBval;
  int* dBtab = nullptr;
  double2* dBval = nullptr;
  double2* dCarry = nullptr;     // two slots: the reader's gamma/n before and after a pass
  int carry_slot = 0;

  double lo = -std::numeric_limits<double>::infinity();
  double hi = std::numeric_limits<double>::infinity();

  // current file
  bool file_open = false;
  bool stopped = false;
  std::string carry_text;        // unread tail: an unfinished record and what follows it

  // per-pass scratch
  unsigned char* dText = nullptr;
  size_t text_cap = 0;
  unsigned* dTokCounts = nullptr;
  unsigned* dTokPos = nullptr;
  size_t tok_cap = 0;
  RecordArrays rec;
  size_t rec_cap = 0;
  int* dHostList = nullptr;
  unsigned* dHostPos = nullptr;
  FeedCtl* dCtl = nullptr;
  unsigned* dKeepCounts = nullptr;
  int* dMaxF = nullptr;
  unsigned long long* dTotal = nullptr;

  // kept lines
  LinesOut out;
  size_t nout = 0, out_cap = 0;
  size_t records_read = 0;

  float last_ms = 0.0f;
  size_t last_bytes = 0, last_records = 0;

  int ensure_scratch(size_t len);
  int ensure_out(size_t n);
  int pass(const std::string& head, const char* tail, size_t tail_len, bool final, size_t* consumed);
  int read_host_record(const std::string& head, const char* tail, size_t pos, size_t end_pos, int r, bool* ok,
                       bool* garbage_after);
  char at(const std::string& head, const char* tail, size_t k) const {
    return k < head.size() ? head[k] : tail[k - head.size()];
  }
  void release();
};

int gx_ingest::ensure_scratch(size_t len) {
  int rc;
  if (len + 16 > text_cap) {
    const size_t cap = std::max(len + 16, std::min(kPieceBytes + 4096, 2 * text_cap));
    if ((rc = dev_alloc(&dText, cap))) return rc;
    text_cap = cap;
    const size_t nblk = (cap + kTokTile - 1) / kTokTile;
    if ((rc = dev_alloc(&dTokCounts, nblk + 1))) return rc;
  }
  const size_t maxtok = len / 2 + 2;          // a token and a separator per two bytes at most
  if (maxtok > tok_cap) {
    if ((rc = dev_alloc(&dTokPos, maxtok))) return rc;
    tok_cap = maxtok;
  }
  const size_t maxrec = maxtok / 3 + 1;
  if (maxrec > rec_cap) {
    if ((rc = dev_alloc(&rec.nu, maxrec)) || (rc = dev_alloc(&rec.elow, maxrec)) ||
        (rc = dev_alloc(&rec.aif, maxrec)) || (rc = dev_alloc(&rec.gns, maxrec)) ||
        (rc = dev_alloc(&rec.bidx, maxrec)) || (rc = dev_alloc(&rec.flag, maxrec)) ||
        (rc = dev_alloc(&dHostList, maxrec)) || (rc = dev_alloc(&dHostPos, maxrec)))
      return rc;
    const size_t nblk = (maxrec + kRecBlock - 1) / kRecBlock;
    if ((rc = dev_alloc(&dKeepCounts, nblk + 1)) || (rc = dev_alloc(&dMaxF, nblk + 1))) return rc;
    rec_cap = maxrec;
  }
  if (!dCtl && ((rc = dev_alloc(&dCtl, 1)) || (rc = dev_alloc(&dTotal, 2)))) return rc;
  return GX_OK;
}

int gx_ingest::ensure_out(size_t n) {
  if (n <= out_cap) return GX_OK;
  const size_t cap = std::max(n, out_cap + out_cap / 2 + 1024);
  LinesOut o;
  int rc;
  if ((rc = dev_alloc(&o.elow, cap)) || (rc = dev_alloc(&o.nu, cap)) || (rc = dev_alloc(&o.aif, cap)) ||
      (rc = dev_alloc(&o.gns, cap)) || (rc = dev_alloc(&o.gamma, cap)) || (rc = dev_alloc(&o.nexp, cap))) {
    dev_free(o.elow), dev_free(o.nu), dev_free(o.aif), dev_free(o.gns), dev_free(o.gamma), dev_free(o.nexp);
    return rc;
  }
  if (nout) {
    hipError_t e;
    if ((e = hipMemcpyAsync(o.elow, out.elow, nout * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(o.nu, out.nu, nout * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(o.aif, out.aif, nout * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(o.gns, out.gns, nout * 4, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(o.gamma, out.gamma, nout * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(o.nexp, out.nexp, nout * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipStreamSynchronize(s)))
      return hip_fail(e, "ingest: growing the line arrays");
  }
  dev_free(out.elow), dev_free(out.nu), dev_free(out.aif), dev_free(out.gns), dev_free(out.gamma),
      dev_free(out.nexp);
  out = o;
  out_cap = cap;
  return GX_OK;
}

void gx_ingest::release() {
  if (device >= 0) hipSetDevice(device);
  dev_free(dE), dev_free(dG), dev_free(dJ), dev_free(dBtab), dev_free(dBval), dev_free(dCarry);
  dev_free(dText), dev_free(dTokCounts), dev_free(dTokPos);
  dev_free(rec.nu), dev_free(rec.elow), dev_free(rec.aif), dev_free(rec.gns), dev_free(rec.bidx),
      dev_free(rec.flag);
  dev_free(dHostList), dev_free(dHostPos), dev_free(dCtl), dev_free(dKeepCounts), dev_free(dMaxF),
      dev_free(dTotal);
  dev_free(out.elow), dev_free(out.nu), dev_free(out.aif), dev_free(out.gns), dev_free(out.gamma),
      dev_free(out.nexp);
  if (ev0) hipEventDestroy(ev0);
  if (ev1) hipEventDestroy(ev1);
  if (s) hipStreamDestroy(s);
  ev0 = ev1 = nullptr;
  s = nullptr;
}

// Record r read the reference's way: `stream >> id_f >> id_i >> aif` from its first
// token (ExomolStateReader.cpp:145) over the bytes [pos, end_pos) of the pass, which hold
// the record's three tokens and the spaces after them.  *ok: the three reads succeeded;
// *garbage_after: the stream stopped inside the third token, so its next read (the next
// record's id_f) fails there.
int gx_ingest::read_host_record(const std::string& head, const char* tail, size_t pos, size_t end_pos, int r,
                                bool* ok, bool* garbage_after) {
  std::string buf;
  buf.reserve(end_pos - pos);
  for (size_t k = pos; k < end_pos; ++k) buf.push_back(at(head, tail, k));
  // the three tokens the device saw
  size_t ts[3] = {0, 0, 0}, te[3] = {0, 0, 0};
  {
    size_t k = 0;
    for (int t = 0; t < 3; ++t) {
      while (k < buf.size() && dec_is_space((unsigned char)buf[k])) ++k;
      ts[t] = k;
      while (k < buf.size() && !dec_is_space((unsigned char)buf[k])) ++k;
      te[t] = k;
    }
  }
  MemBuf mb(buf.data(), buf.size());
  std::istream is(&mb);
  int idf = 0, idi = 0;
  double a = 0.0;
  *ok = false;
  *garbage_after = false;
  if (!(is >> idf >> idi >> a)) return GX_OK;
  *ok = true;
  const size_t p = mb.consumed();
  if (p != te[2]) {
    // the stream split a token: either it stopped inside the third token and its next
    // int read fails there (reading ends after this record, as in the reference), or
    // the records it reads from here on are not the token triples
    const bool int_next = p < buf.size() &&
                          (dec_is_digit((unsigned char)buf[p]) ||
                           ((buf[p] == '+' || buf[p] == '-') && p + 1 < buf.size() && dec_is_digit((unsigned char)buf[p + 1])));
    if (p <= ts[2] || int_next)
      return fail(GX_INVALID_ARG, "gx_ingest_feed: .trans record " + std::to_string(r) +
                                      " of this pass has tokens that run together (the stream would read other "
                                      "records than the lines); not supported");
    *garbage_after = true;
  }
  const int f = idf - 1, i = idi - 1;
  if (f < 0 || (size_t)f >= nstates || i < 0 || (size_t)i >= nstates)
    return fail(GX_INVALID_ARG, "ingest: a transition refers to a state id outside 1.." + std::to_string(nstates) +
                                    " (record " + std::to_string(r) + " of this pass)");
  const double ei = hE[i];
  const double nu = hE[f] - ei;
  int b = -1;
  if (use_broad) {
    const int ji = hJ[i] - jlo, jf = hJ[f] - jlo;
    if ((unsigned)ji < (unsigned)jn && (unsigned)jf < (unsigned)jn) b = hBtab[(size_t)ji * jn + jf];
  }
  const int g = hG[f];
  const unsigned char fl = 0;
  hipError_t e;
  if ((e = hipMemcpyAsync(rec.nu + r, &nu, 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(rec.elow + r, &ei, 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(rec.aif + r, &a, 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(rec.gns + r, &g, 4, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(rec.bidx + r, &b, 4, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(rec.flag + r, &fl, 1, hipMemcpyHostToDevice, s)) || (e = hipStreamSynchronize(s)))
    return hip_fail(e, "ingest: storing a record read on the host");
  return GX_OK;
}

// One device pass over head + tail[0, tail_len) (cut after a space unless final).
// *consumed = bytes of that text fully read; the rest starts an unfinished record.
int gx_ingest::pass(const std::string& head, const char* tail, size_t tail_len, bool final, size_t* consumed) {
  const size_t len = head.size() + tail_len;
  *consumed = len;
  if (len == 0) return GX_OK;
  int rc;
  if ((rc = ensure_scratch(len))) return rc;
  hipError_t e;
  if (!head.empty() && (e = hipMemcpyAsync(dText, head.data(), head.size(), hipMemcpyHostToDevice, s)))
    return hip_fail(e, "ingest: text upload");
  if (tail_len && (e = hipMemcpyAsync(dText + head.size(), tail, tail_len, hipMemcpyHostToDevice, s)))
    return hip_fail(e, "ingest: text upload");
  const int nblk = (int)((len + kTokTile - 1) / kTokTile);
  unsigned long long ntok = 0;
  if ((e = hipEventRecord(ev0, s)) || (e = launch_tok_count(dText, len, dTokCounts, s)) ||
      (e = launch_scan_sum(dTokCounts, nblk, dTotal, s)) || (e = launch_tok_emit(dText, len, dTokCounts, dTokPos, s)) ||
      (e = hipMemcpyAsync(&ntok, dTotal, 8, hipMemcpyDeviceToHost, s)) || (e = hipStreamSynchronize(s)))
    return hip_fail(e, "ingest: tokenising");
  float ms_tok = 0.0f;
  if ((e = hipEventRecord(ev1, s)) || (e = hipEventSynchronize(ev1)) || (e = hipEventElapsedTime(&ms_tok, ev0, ev1)))
    return hip_fail(e, "ingest: timing");
  const int nrec = (int)(ntok / 3);
  const int rest = (int)(ntok % 3);
  int fail_limit = nrec;            // first record the stream cannot read
  if (rest) {
    if (final) {
      stopped = true;               // a record cut short by the end of file ends the loop
    } else {
      unsigned p = 0;
      if ((e = hipMemcpy(&p, dTokPos + 3 * (size_t)nrec, 4, hipMemcpyDeviceToHost)))
        return hip_fail(e, "ingest: tokenising");
      *consumed = p;
    }
  }
  IngestView v;
  v.E = dE;
  v.g = dG;
  v.J = dJ;
  v.nstates = (int)nstates;
  v.btab = dBtab;
  v.jlo = jlo;
  v.jn = jn;
  v.use_broad = use_broad ? 1 : 0;
  v.lo = lo;
  v.hi = hi;
  FeedCtl ctl;
  if ((e = hipEventRecord(ev0, s)) || (e = launch_reset_ctl(dCtl, s)) ||
      (e = launch_records(dText, len, dTokPos, nrec, v, rec, dCtl, dHostList, s)) ||
      (e = hipMemcpyAsync(&ctl, dCtl, sizeof ctl, hipMemcpyDeviceToHost, s)) || (e = hipStreamSynchronize(s)))
    return hip_fail(e, "ingest: reading records");
  float ms_rec = 0.0f;
  if ((e = hipEventRecord(ev1, s)) || (e = hipEventSynchronize(ev1)) || (e = hipEventElapsedTime(&ms_rec, ev0, ev1)))
    return hip_fail(e, "ingest: timing");

  // records the device left to the stream, in order
  int over = ctl.min_over;
  bool read_fails = false;          // the stream fails at or after a record of this pass
  if (ctl.nhost > 0) {
    std::vector<int> hl(ctl.nhost);
    if ((e = hipMemcpy(hl.data(), dHostList, 4 * (size_t)ctl.nhost, hipMemcpyDeviceToHost)))
      return hip_fail(e, "ingest: host records");
    std::sort(hl.begin(), hl.end());
    if ((e = hipMemcpy(dHostList, hl.data(), 4 * (size_t)ctl.nhost, hipMemcpyHostToDevice)) ||
        (e = launch_gather_pos(dTokPos, dHostList, ctl.nhost, dHostPos, s)) || (e = hipStreamSynchronize(s)))
      return hip_fail(e, "ingest: host records");
    std::vector<unsigned> pos(ctl.nhost);
    if ((e = hipMemcpy(pos.data(), dHostPos, 4 * (size_t)ctl.nhost, hipMemcpyDeviceToHost)))
      return hip_fail(e, "ingest: host records");
    for (int k = 0; k < ctl.nhost; ++k) {
      const int r = hl[k];
      if (r >= fail_limit || r > over) break;      // not read: the loop ended before it
      unsigned endp = (unsigned)len;
      if (r + 1 < nrec && (e = hipMemcpy(&endp, dTokPos + 3 * (size_t)(r + 1), 4, hipMemcpyDeviceToHost)))
        return hip_fail(e, "ingest: host records");
      if (r + 1 >= nrec) endp = (unsigned)(rest ? *consumed : len);
      bool ok = false, garbage = false;
      if ((rc = read_host_record(head, tail, pos[k], endp, r, &ok, &garbage))) return rc;
      if (!ok) {
        fail_limit = r;
        read_fails = true;
        break;
      }
      double nu = 0.0;
      if ((e = hipMemcpy(&nu, rec.nu + r, 8, hipMemcpyDeviceToHost))) return hip_fail(e, "ingest: host records");
      if (nu > hi) over = std::min(over, r);
      if (garbage) {
        fail_limit = r + 1;
        read_fails = true;
        break;
      }
    }
  }
  int nread, brk = -1;
  if (over < fail_limit) {
    nread = over + 1;                // the line above hi is read, then the loop breaks
    brk = over;
    stopped = true;
  } else {
    nread = fail_limit;
    if (read_fails) stopped = true;
  }
  if (ctl.min_badid < nread)
    return fail(GX_INVALID_ARG, "ingest: a transition refers to a state id outside 1.." + std::to_string(nstates) +
                                    " (record " + std::to_string(ctl.min_badid) + " of this pass)");
  if (stopped) *consumed = len;

  float ms_keep = 0.0f;
  if (nread > 0) {
    if ((rc = ensure_out(nout + (size_t)nread))) return rc;
    unsigned long long kept = 0;
    if ((e = hipEventRecord(ev0, s)) ||
        (e = launch_keep(rec, nread, brk, lo, dBval, dCarry, carry_slot, out, nout, dKeepCounts, dMaxF, dTotal, s)) ||
        (e = hipEventRecord(ev1, s)) || (e = hipMemcpyAsync(&kept, dTotal, 8, hipMemcpyDeviceToHost, s)) ||
        (e = hipStreamSynchronize(s)) || (e = hipEventElapsedTime(&ms_keep, ev0, ev1)))
      return hip_fail(e, "ingest: keeping lines");
    carry_slot = 1 - carry_slot;
    nout += kept;
    records_read += (size_t)nread;
  }
  last_ms += ms_tok + ms_rec + ms_keep;
  last_bytes += len;
  last_records += (size_t)nread;
  return GX_OK;
}

extern "C" {

int gx_partition(const double* energy, const int* gtot, size_t nstates, double temperature, double* q) {
#pragma clang fp contract(off)
  if (!q || (nstates && (!energy || !gtot))) return fail(GX_INVALID_ARG, "gx_partition: null pointer");
  if (!(temperature > 0)) return fail(GX_INVALID_ARG, "gx_partition: temperature must be > 0");
  // exomol_functions.cpp:55-66: local constants (equal to defines.h), one serial sum
  const double planck = 6.6260693e-27, vellgt = 2.99792458e10, boltz = 1.380658e-16;
  const double beta = planck * vellgt / (boltz * temperature);
  double partition = 0.0;
  for (size_t i = 0; i < nstates; ++i) partition += gtot[i] * exp(-beta * energy[i]);
  *q = partition;
  return GX_OK;
}

int gx_states_count(const char* text, size_t len, size_t* nstates) {
  if (!nstates || (len && !text)) return fail(GX_INVALID_ARG, "gx_states_count: null pointer");
  *nstates = states_count(text, len);
  return GX_OK;
}

int gx_states_parse(const char* text, size_t len, size_t nstates, int nthreads, int* id, double* energy, int* gtot,
                    int* J) {
  if ((len && !text) || (nstates && (!energy || !gtot))) return fail(GX_INVALID_ARG, "gx_states_parse: null pointer");
  std::vector<size_t> starts;
  states_pieces(text, len, &starts);
  const size_t n = starts.size() - 1;
  if (n != nstates)
    return fail(GX_INVALID_ARG, "gx_states_parse: the text holds " + std::to_string(n) + " state records, not " +
                                    std::to_string(nstates));
  unsigned nt = nthreads > 0 ? (unsigned)nthreads : std::max(1u, std::thread::hardware_concurrency());
  nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, n / 4096));
  auto work = [&](size_t a, size_t z) {
    char buf[1024];
    for (size_t k = a; k < z; ++k) {
      const size_t m = starts[k + 1] - starts[k];
      std::memcpy(buf, text + starts[k], m);
      buf[m] = '\0';
      // exomol_functions.cpp:29-35
      char* p;
      const int sid = (int)(strtol(buf, &p, 0) - 1);
      const double e = strtod(p, &p);
      const int g = (int)strtol(p, &p, 0);
      const int j = (int)strtol(p, &p, 0);
      if (id) id[k] = sid;
      energy[k] = e;
      gtot[k] = g;
      if (J) J[k] = j;
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, n * t / nt, n * (t + 1) / nt);
  work(0, n / nt);
  for (auto& x : th) x.join();
  return GX_OK;
}

int gx_ingest_create(int device, size_t nstates, const double* energy, const int* gtot, const int* J,
                     gx_ingest** out) {
  if (!out || (nstates && (!energy || !gtot))) return fail(GX_INVALID_ARG, "gx_ingest_create: null pointer");
  if (nstates > (size_t)0x7fffffff) return fail(GX_INVALID_ARG, "gx_ingest_create: more than 2^31-1 states");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GX_NO_DEVICE, "gx_ingest_create: no HIP device");
  if (device < 0 || device >= ndev) return fail(GX_INVALID_ARG, "gx_ingest_create: bad device index");
  hipError_t e;
  if ((e = hipSetDevice(device))) return hip_fail(e, "gx_ingest_create: hipSetDevice");
  gx_ingest* g = new gx_ingest();
  g->device = device;
  g->nstates = nstates;
  g->hE.assign(energy, energy + nstates);
  g->hG.assign(gtot, gtot + nstates);
  if (J) g->hJ.assign(J, J + nstates);
  else g->hJ.assign(nstates, 0);
  int rc;
  if ((e = hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking)) || (e = hipEventCreate(&g->ev0)) ||
      (e = hipEventCreate(&g->ev1))) {
    g->release();
    delete g;
    return hip_fail(e, "gx_ingest_create");
  }
  if ((rc = dev_alloc(&g->dE, std::max<size_t>(nstates, 1))) || (rc = dev_alloc(&g->dG, std::max<size_t>(nstates, 1))) ||
      (rc = dev_alloc(&g->dJ, std::max<size_t>(nstates, 1))) || (rc = dev_alloc(&g->dCarry, 2))) {
    g->release();
    delete g;
    return rc;
  }
  const double2 zero2[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (nstates && ((e = hipMemcpy(g->dE, energy, nstates * 8, hipMemcpyHostToDevice)) ||
                  (e = hipMemcpy(g->dG, gtot, nstates * 4, hipMemcpyHostToDevice)) ||
                  (e = hipMemcpy(g->dJ, g->hJ.data(), nstates * 4, hipMemcpyHostToDevice)))) {
    g->release();
    delete g;
    return hip_fail(e, "gx_ingest_create: states upload");
  }
  if ((e = hipMemcpy(g->dCarry, zero2, sizeof zero2, hipMemcpyHostToDevice))) {
    g->release();
    delete g;
    return hip_fail(e, "gx_ingest_create");
  }
  *out = g;
  return GX_OK;
}

void gx_ingest_destroy(gx_ingest* ing) {
  if (!ing) return;
  ing->release();
  delete ing;
}

int gx_ingest_set_broadening(gx_ingest* g, const char* text, size_t len, double pressure) {
  if (!g || (len && !text)) return fail(GX_INVALID_ARG, "gx_ingest_set_broadening: null pointer");
  hipSetDevice(g->device);
  dev_free(g->dBtab);
  dev_free(g->dBval);
  g->hBtab.clear();
  g->hBval.clear();
  g->use_broad = false;
  if (len == 0) return GX_OK;
  // InitializeBroadener (ExomolStateReader.cpp:52-79): per line, split on spaces;
  // key = code + Ji + Jf with code "a1" only for an "a1" line (Jf = 5th field), else
  // "a0" and Jf = -1; a later line with the same key replaces the earlier one
  std::map<std::string, std::pair<double, double>> keys;
  size_t s = 0, line_no = 0;
  while (s < len) {
    const void* nl = std::memchr(text + s, '\n', len - s);
    const size_t e = nl ? (size_t)((const char*)nl - text) : len;
    std::string line(text + s, e - s);
    s = e + 1;
    ++line_no;
    std::istringstream is(line);
    std::vector<std::string> f;
    for (std::string w; is >> w;) f.push_back(w);
    if (f.empty()) continue;   // the reference indexes split_line[3] of a blank line (undefined); skipped here
    if (f.size() < 4 || (f[0] == "a1" && f.size() < 5))
      return fail(GX_INVALID_ARG, "gx_ingest_set_broadening: line " + std::to_string(line_no) +
                                      " has too few fields for its code");
    std::string code = "a0";
    const int ji = atoi(f[3].c_str());
    int jf = -1;
    const double gam = atof(f[1].c_str());
    const double n = atof(f[2].c_str());
    if (f[0] == "a1") {
      code = "a1";
      jf = atoi(f[4].c_str());
    }
    std::ostringstream key;
    key << code << ji << jf;
    keys[key.str()] = {gam, n};
  }
  g->use_broad = true;
  if (g->nstates == 0) return GX_OK;
  const auto mm = std::minmax_element(g->hJ.begin(), g->hJ.end());
  const long long span = (long long)*mm.second - *mm.first + 1;
  if (span > 4096) return fail(GX_INVALID_ARG, "gx_ingest_set_broadening: state J values span more than 4096");
  g->jlo = *mm.first;
  g->jn = (int)span;
  // GetGammaN (ExomolStateReader.cpp:12-29): "a1" + Ji + Jf, then, since key.clear()
  // resets only the stream state, "a1" + Ji + Jf + "a0" + Ji + "-1"; gamma *= pressure
  std::map<std::string, int> index;
  g->hBtab.assign((size_t)g->jn * g->jn, -1);
  for (int a = 0; a < g->jn; ++a) {
    for (int b = 0; b < g->jn; ++b) {
      const int ji = g->jlo + a, jf = g->jlo + b;
      std::ostringstream k1;
      k1 << "a1" << ji << jf;
      std::string key = k1.str();
      auto it = keys.find(key);
      if (it == keys.end()) {
        std::ostringstream k2;
        k2 << key << "a0" << ji << -1;
        key = k2.str();
        it = keys.find(key);
      }
      if (it == keys.end()) continue;
      auto ix = index.find(key);
      if (ix == index.end()) {
        ix = index.emplace(key, (int)g->hBval.size()).first;
        g->hBval.push_back(make_double2(it->second.first * pressure, it->second.second));
      }
      g->hBtab[(size_t)a * g->jn + b] = ix->second;
    }
  }
  int rc;
  if ((rc = dev_alloc(&g->dBtab, g->hBtab.size())) || (rc = dev_alloc(&g->dBval, std::max<size_t>(g->hBval.size(), 1))))
    return rc;
  hipError_t e;
  if ((e = hipMemcpy(g->dBtab, g->hBtab.data(), 4 * g->hBtab.size(), hipMemcpyHostToDevice)) ||
      (!g->hBval.empty() &&
       (e = hipMemcpy(g->dBval, g->hBval.data(), sizeof(double2) * g->hBval.size(), hipMemcpyHostToDevice))))
    return hip_fail(e, "gx_ingest_set_broadening: upload");
  return GX_OK;
}

int gx_ingest_set_range(gx_ingest* g, double lo, double hi) {
  if (!g) return fail(GX_INVALID_ARG, "gx_ingest_set_range: null reader");
  g->lo = lo;
  g->hi = hi;
  return GX_OK;
}

int gx_ingest_begin_file(gx_ingest* g) {
  if (!g) return fail(GX_INVALID_ARG, "gx_ingest_begin_file: null reader");
  g->file_open = true;
  g->stopped = false;
  g->carry_text.clear();
  return GX_OK;
}

int gx_ingest_feed(gx_ingest* g, const char* text, size_t len, int eof) {
  if (!g || (len && !text)) return fail(GX_INVALID_ARG, "gx_ingest_feed: null pointer");
  if (!g->file_open) return fail(GX_INVALID_ARG, "gx_ingest_feed: call gx_ingest_begin_file first");
  hipSetDevice(g->device);
  g->last_ms = 0.0f;
  g->last_bytes = 0;
  g->last_records = 0;
  if (g->stopped) return GX_OK;          // the reader left this file's loop already
  std::string pending;
  pending.swap(g->carry_text);
  size_t pos = 0;
  for (;;) {
    if (pending.size() >= kPieceBytes)
      return fail(GX_INVALID_ARG, "gx_ingest_feed: a record longer than 256 MiB");
    const size_t take = std::min(len - pos, kPieceBytes - pending.size());
    const bool all_in = pos + take == len;
    const bool final = all_in && eof;
    const size_t plen = pending.size() + take;
    size_t cut = plen;
    if (!final) {
      size_t k = plen;
      while (k > 0 && !dec_is_space((unsigned char)g->at(pending, text + pos, k - 1))) --k;
      if (k == 0) {
        if (!all_in) return fail(GX_INVALID_ARG, "gx_ingest_feed: a token longer than 256 MiB");
        pending.append(text + pos, take);
        g->carry_text.swap(pending);
        return GX_OK;
      }
      cut = k;
    }
    size_t consumed = 0;
    int rc;
    if (cut <= pending.size()) {
      const std::string head = pending.substr(0, cut);
      rc = g->pass(head, nullptr, 0, final, &consumed);
    } else {
      rc = g->pass(pending, text + pos, cut - pending.size(), final, &consumed);
    }
    if (rc) return rc;
    if (g->stopped) {
      g->carry_text.clear();
      return GX_OK;
    }
    std::string next;
    for (size_t k = consumed; k < plen; ++k) next.push_back(g->at(pending, text + pos, k));
    pending.swap(next);
    pos += take;
    if (all_in) {
      if (final) {
        if (!pending.empty()) {                    // only spaces can remain after a final pass
          g->carry_text.clear();
        }
      } else {
        g->carry_text.swap(pending);
      }
      return GX_OK;
    }
  }
}

int gx_ingest_count(gx_ingest* g, size_t* nlines, size_t* nrecords) {
  if (!g) return fail(GX_INVALID_ARG, "gx_ingest_count: null reader");
  if (nlines) *nlines = g->nout;
  if (nrecords) *nrecords = g->records_read;
  return GX_OK;
}

int gx_ingest_download(gx_ingest* g, size_t n, double* elow, double* nu, double* aif, int* gns, double* gamma,
                       double* nexp) {
  if (!g || (n && (!elow || !nu || !aif || !gns))) return fail(GX_INVALID_ARG, "gx_ingest_download: null pointer");
  if (n > g->nout) return fail(GX_INVALID_ARG, "gx_ingest_download: more lines asked for than kept");
  if (n == 0) return GX_OK;
  hipSetDevice(g->device);
  hipError_t e;
  if ((e = hipMemcpy(elow, g->out.elow, n * 8, hipMemcpyDeviceToHost)) ||
      (e = hipMemcpy(nu, g->out.nu, n * 8, hipMemcpyDeviceToHost)) ||
      (e = hipMemcpy(aif, g->out.aif, n * 8, hipMemcpyDeviceToHost)) ||
      (e = hipMemcpy(gns, g->out.gns, n * 4, hipMemcpyDeviceToHost)) ||
      (gamma && (e = hipMemcpy(gamma, g->out.gamma, n * 8, hipMemcpyDeviceToHost))) ||
      (nexp && (e = hipMemcpy(nexp, g->out.nexp, n * 8, hipMemcpyDeviceToHost))))
    return hip_fail(e, "gx_ingest_download");
  return GX_OK;
}

int gx_ingest_clear(gx_ingest* g) {
  if (!g) return fail(GX_INVALID_ARG, "gx_ingest_clear: null reader");
  g->nout = 0;
  g->records_read = 0;
  return GX_OK;
}

int gx_ingest_last_feed(gx_ingest* g, float* kernel_ms, size_t* text_bytes, size_t* records) {
  if (!g) return fail(GX_INVALID_ARG, "gx_ingest_last_feed: null reader");
  if (kernel_ms) *kernel_ms = g->last_ms;
  if (text_bytes) *text_bytes = g->last_bytes;
  if (records) *records = g->last_records;
  return GX_OK;
}

int gx_format_trans(size_t n, const int* id_f, const int* id_i, const double* aif, char* out, size_t cap,
                    size_t* written) {
  if ((n && (!id_f || !id_i || !aif || !out)) || !written) return fail(GX_INVALID_ARG, "gx_format_trans: null pointer");
  const unsigned nt = (unsigned)std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()),
                                                 std::max<size_t>(1, n / 65536));
  // each thread formats its share into its own buffer, then the shares are joined in order
  std::vector<std::string> part(nt);
  auto work = [&](unsigned t) {
    const size_t a = n * t / nt, z = n * (t + 1) / nt;
    std::string& o = part[t];
    o.reserve((z - a) * 38);
    char buf[96];
    for (size_t k = a; k < z; ++k) {
      const int m = std::snprintf(buf, sizeof buf, "%12d %12d %10.4E\n", id_f[k], id_i[k], aif[k]);
      o.append(buf, (size_t)m);
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& p : part) total += p.size();
  *written = total;
  if (total > cap) return fail(GX_INVALID_ARG, "gx_format_trans: output buffer too small");
  size_t o = 0;
  for (auto& p : part) {
    std::memcpy(out + o, p.data(), p.size());
    o += p.size();
  }
  return GX_OK;
}

}  // extern "C"
