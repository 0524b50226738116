// gx_ingest_internal.h -- device-side view and kernel launchers of the ExoMol .trans
// reader (host side: gx_ingest.cpp, kernels: gx_ingest_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gx {

constexpr int kTokBlock = 256;               // threads per tokenising block
constexpr int kTokTile = kTokBlock * 16;     // text bytes per tokenising block
constexpr int kRecBlock = 256;               // records per compaction block

// States (positional: state id k is entry k - 1) and the broadening lookup of one reader.
struct IngestView {
  const double* E = nullptr;      // state energies
  const int* g = nullptr;         // state g_tot
  const int* J = nullptr;         // state J
  int nstates = 0;
  const int* btab = nullptr;      // (Ji - jlo) * jn + (Jf - jlo) -> value index or -1
  int jlo = 0, jn = 0;
  int use_broad = 0;
  double lo = 0.0, hi = 0.0;      // keep nu >= lo, stop at the first nu > hi
};

// Per-record results of one piece of text.
struct RecordArrays {
  double* nu = nullptr;
  double* elow = nullptr;
  double* aif = nullptr;
  int* gns = nullptr;
  int* bidx = nullptr;            // broadening value index, -1: no key matched
  unsigned char* flag = nullptr;  // 0 read here, 1 left to the host, 2 state id out of range
};

// Scalars the record kernel reduces into (reset per piece).
struct FeedCtl {
  int min_over;                   // first record with nu > hi
  int min_badid;                  // first record with a state id outside 1..nstates
  int nhost;                      // records left to the host
  int pad;
};

// Kept lines (the reader's output, in read order).
struct LinesOut {
  double* elow = nullptr;
  double* nu = nullptr;
  double* aif = nullptr;
  int* gns = nullptr;
  double* gamma = nullptr;
  double* nexp = nullptr;
};

// Tokenising: block counts, then (after the scan) token start offsets.  text must hold
// len bytes plus 16 readable bytes of padding.
hipError_t launch_tok_count(const unsigned char* text, size_t len, unsigned* counts, hipStream_t s);
hipError_t launch_scan_sum(unsigned* v, int n, unsigned long long* total, hipStream_t s);
hipError_t launch_tok_emit(const unsigned char* text, size_t len, const unsigned* offsets, unsigned* tokpos,
                           hipStream_t s);
hipError_t launch_reset_ctl(FeedCtl* ctl, hipStream_t s);
hipError_t launch_records(const unsigned char* text, size_t len, const unsigned* tokpos, int nrec,
                          const IngestView& v, const RecordArrays& r, FeedCtl* ctl, int* host_list,
                          hipStream_t s);
hipError_t launch_gather_pos(const unsigned* tokpos, const int* recs, int n, unsigned* out, hipStream_t s);
// Ordered compaction of records [0, nread) (the break record brk excluded, -1: none):
// kept lines appended at out + base, gamma/n carried from the last record with a
// broadening match (carry[in_slot] before the first), the reader's values after the last
// record written to carry[1 - in_slot]; *nkept = the number kept by this call.
hipError_t launch_keep(const RecordArrays& r, int nread, int brk, double lo, const double2* bval, double2* carry,
                       int in_slot, LinesOut out, size_t base, unsigned* counts, int* maxf,
                       unsigned long long* nkept, hipStream_t s);

}  // namespace gx
