// gx_server.cpp -- continuous batching of cross-section requests on one GPU.
//
// One scheduler thread owns every device call.  Clients submit requests
// (T, P, Q, profile, grid) from any thread; the scheduler
//   1. retires finished batches (event query, never a blocking wait on a
//      running batch), copies their grids to host memory and frees budget;
//   2. admits queued requests, oldest first, while an in-flight slot is free
//      and the batch fits the HBM budget, pulling every compatible queued
//      request (same grid, profile, precision, cutoff) into the same batch;
//   3. launches the batch on the slot's own stream, so batches overlap on
//      the device and new requests start while others are still running.
// A request only touches the slice of the resident, nu-sorted line list
// that can reach its grid.  Mixed-precision batches are one launch of the
// table kernel over all requests of the batch, with one workgroup per work
// item (not persistent), so the hardware interleaves the workgroups of
// batches on other streams and a lone batch still fills the chip; the work
// split per request depends on the request's own grid only, so its result
// is the same alone and in any mix.  FP64 and Doppler requests run the plain
// per-point kernels.  Slot buffers grow and are reused (no per-batch
// hipMalloc/hipFree on the critical path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/gx.h"
#include "gx_common.h"

using namespace gx;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
  uint64_t id = 0;
  gx_request r{};
  int state = GX_REQ_QUEUED;
  std::atomic<bool> cancel{false};   // set by gx_server_cancel, read by the scheduler without the lock
  Clock::time_point t_submit, t_done;
  std::vector<double> host;
  long long evals = 0;
  float kernel_ms = 0;
  std::string error;
};

// Device buffers of one in-flight slot; they only grow.
struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  bool busy = false;
  std::vector<Req*> reqs;
  std::vector<PointConsts> h_pcs;
  double* freq = nullptr;
  size_t freq_cap = 0;
  double* out = nullptr;
  size_t out_cap = 0;
  int* ib = nullptr;
  int* ie = nullptr;
  size_t win_cap = 0;
  unsigned long long* evals = nullptr;
  size_t evals_cap = 0;        // counters
  Derived der;
  Ranges rng;
  MixedRunner run;
  Clock::time_point t_launch;
  // where each request's grid sits in out, and whether the batch ran the batched FP64 path
  // (one eval counter set per request) or the mixed path (one set for the batch)
  std::vector<size_t> out_off;
  bool f64 = false;
  // batched FP64 path: tasks, work items and the block -> items table (the host copies stay
  // until the batch retires, since the uploads read them), their device copies, item partials
  std::vector<F64Task> h_t64;
  std::vector<F64Item> h_items;
  std::vector<int2> h_bitems;
  F64Task* d_t64 = nullptr;
  size_t t64_cap = 0;
  F64Item* d_items = nullptr;
  size_t items_cap = 0;
  int2* d_bitems = nullptr;
  size_t bitems_cap = 0;
  double* partial = nullptr;
  size_t partial_cap = 0;      // doubles

  size_t footprint() const {
    return freq_cap * 8 + out_cap * 8 + win_cap * 8 + der.cap * 32 +
           der.mcap * ((size_t)der.mpoints * kMixedPointBytes + (size_t)der.mtemps * kMixedTempBytes) +
           (size_t)rng.cap * sizeof(int2) + run.bytes() + evals_cap * 8 + t64_cap * sizeof(F64Task) +
           items_cap * sizeof(F64Item) + bitems_cap * sizeof(int2) + partial_cap * 8;
  }
  void release_buffers() {
    (void)hipFree(freq); (void)hipFree(out); (void)hipFree(ib); (void)hipFree(ie);
    freq = out = nullptr;
    ib = ie = nullptr;
    freq_cap = out_cap = win_cap = 0;
    (void)hipFree(d_t64); (void)hipFree(d_items); (void)hipFree(d_bitems); (void)hipFree(partial);
    d_t64 = nullptr;
    d_items = nullptr;
    d_bitems = nullptr;
    partial = nullptr;
    t64_cap = items_cap = bitems_cap = partial_cap = 0;
    der.release();
    rng.release();
    run.release();
  }
};

bool fp64_req(const gx_request& r) { return r.precision == GX_PREC_FP64; }

// Requests that can share a launch: any two FP64 requests (the batched FP64 path gives each
// its own grid, constants and line slice); mixed-precision requests on one grid and cutoff.
bool compatible(const gx_request& a, const gx_request& b) {
  if (fp64_req(a) && fp64_req(b)) return true;
  return a.profile == b.profile && a.precision == b.precision && a.start_nu == b.start_nu && a.end_nu == b.end_nu &&
         a.npoints == b.npoints && reference_cutoff(a.profile, a.pressure, a.cutoff) ==
         reference_cutoff(b.profile, b.pressure, b.cutoff) && a.doppler_widths == b.doppler_widths;
}

bool mixed_req(const gx_request& r) { return r.profile == GX_VOIGT && r.precision == GX_PREC_MIXED; }

template <class T>
int grow(T*& p, size_t& cap, size_t n, const char* what) {
  if (n <= cap) return GX_OK;
  (void)hipFree(p);
  p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1));
  if (e) return hip_fail(e, what);
  cap = std::max<size_t>(n, 1);
  return GX_OK;
}

}  // namespace

struct gx_server {
  gx_lines* L = nullptr;
  gx_server_opts o{};
  std::vector<std::unique_ptr<Slot>> slots;
  size_t budget = 0;
  int cus = 1;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread worker;

  // statistics (under mu)
  long long submitted = 0, completed = 0, cancelled = 0, failed = 0, batches = 0, batched = 0, evals = 0, shared = 0;
  std::vector<double> latencies_ms;
  bool started = false;
  Clock::time_point t_first, t_last;
  size_t peak_bytes = 0;
};

namespace {

size_t f64_batch_need(const gx_server* S, const Slot& sl, const std::vector<Req*>& batch);

size_t batch_need(const gx_server* S, const Slot& sl, const std::vector<Req*>& batch, size_t nlines) {
  const gx_request& r = batch[0]->r;
  if (fp64_req(r)) return f64_batch_need(S, sl, batch);
  const size_t M = (size_t)r.npoints, B = batch.size();
  size_t need = std::max(sl.freq_cap, M) * 8 + std::max(sl.out_cap, M * B) * 8 + std::max(sl.win_cap, nlines) * 8;
  if (mixed_req(r)) {
    need += std::max(sl.der.mcap * ((size_t)sl.der.mpoints * kMixedPointBytes + (size_t)sl.der.mtemps * kMixedTempBytes),
                     nlines * B * (kMixedPointBytes + kMixedTempBytes));
    const size_t nb = (M + 1023) / 1024 + 1;
    need += nb * B * 16 * 1024 * sizeof(float2) / 4;   // partials, upper estimate
    need += (M + 4096) * B * 4 * sizeof(double);       // core-pass rows, upper estimate
    need += far_field_bytes((int)B, (long long)nlines, (int)(M / 64 + 1));
  } else {
    need += std::max(sl.der.cap, nlines) * 32;
  }
  (void)S;
  return need;
}

// Constants of one request on its own grid.
PointConsts request_consts(const gx_server* S, const gx_request& q) {
  gx_consts c{};
  c.temperature = q.temperature;
  c.partition = q.partition;
  c.dfreq = (q.end_nu - q.start_nu) / double(q.npoints);   // BaseProfile.cpp:10
  c.mean_mass = S->o.mean_mass;
  c.pressure = q.pressure;
  c.ref_temp = S->o.ref_temp;
  c.cutoff = q.cutoff;
  c.doppler_widths = q.doppler_widths;
  c.convention = S->o.convention;
  c.precision = q.precision;
  return make_point_consts(q.profile, c, q.pressure);
}

// Widest per-line window of a batch: with per-line Doppler windows (k alpha + dfreq) it
// grows with temperature, so the hottest request decides the line slice and block ranges
// (a request's extra lines have windows that miss its grid and add nothing).
double batch_cutmax(const gx_server* S, const std::vector<Req*>& batch) {
  double c = 0.0;
  for (const Req* q : batch) {
    const PointConsts pc = request_consts(S, q->r);
    c = std::max(c, max_cutoff(pc, S->L->ls.max_nu, (q->r.end_nu - q->r.start_nu) / double(q->r.npoints)));
  }
  return c;
}

// Line slice [lo, hi) of the sorted list that can reach grid [start, start + M*d) with cutoff c.
void line_slice(const gx_lines* L, double start, double end, double cmax, double d, size_t& lo, size_t& hi) {
  const double a = start - cmax - 2 * d, b = end + cmax + 2 * d;
  lo = std::lower_bound(L->h_nu.begin(), L->h_nu.end(), a) - L->h_nu.begin();
  hi = std::upper_bound(L->h_nu.begin(), L->h_nu.end(), b) - L->h_nu.begin();
}

// Grid spacing, own widest window and own line slice of one request.
struct ReqGeo {
  double d, cut;
  size_t lo, hi;
};

ReqGeo request_geo(const gx_server* S, const gx_request& q, const PointConsts& pc) {
  ReqGeo g{};
  g.d = (q.end_nu - q.start_nu) / double(q.npoints);     // BaseProfile.cpp:10
  g.cut = max_cutoff(pc, S->L->ls.max_nu, g.d);
  line_slice(S->L, q.start_nu, q.start_nu + q.npoints * g.d, g.cut, g.d, g.lo, g.hi);
  return g;
}

// Device bytes of a batched FP64 launch: grids and sums, derived data of every slice, the
// item partials (at most two items per block plus one per kFp64Chunk lines of the blocks'
// ranges; a line lies in the ranges of at most (2 cut + 3 dnu) / (256 dnu) + 2 blocks).
constexpr long long kFp64Chunk = 4096;    // smallest chunk of lines per item
size_t f64_batch_need(const gx_server* S, const Slot& sl, const std::vector<Req*>& batch) {
  size_t bins = 0, lines = 0, items = 0, blocks = 0;
  for (const Req* q : batch) {
    const ReqGeo g = request_geo(S, q->r, request_consts(S, q->r));
    const size_t M = (size_t)q->r.npoints, n = g.hi - g.lo, nb = (M + kBinsF64 - 1) / kBinsF64;
    const double per_line = (2.0 * g.cut + 3.0 * g.d) / (kBinsF64 * g.d) + 2.0;
    bins += M;
    lines += n;
    blocks += nb;
    items += 2 * nb + (size_t)std::ceil((double)n * per_line / (double)kFp64Chunk);
  }
  const size_t B = batch.size();
  return std::max(sl.freq_cap, bins) * 8 + std::max(sl.out_cap, bins) * 8 + std::max(sl.der.cap, lines) * 32 +
         std::max(sl.partial_cap, items * kBinsF64) * 8 + items * sizeof(F64Item) + blocks * sizeof(int2) +
         B * (sizeof(F64Task) + kEvalSlots * 8);
}

template <class T>
int grow_copy(T*& d, size_t& cap, const std::vector<T>& h, hipStream_t s, const char* what) {
  int rc = grow(d, cap, h.size(), what);
  if (rc || h.empty()) return rc;
  hipError_t e = hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s);
  return e ? hip_fail(e, what) : GX_OK;
}

// FP64 requests of any grids in one batch: each request's grid, constants and line slice
// (from its own widest window), its bin blocks' line ranges from the host copy of the sorted
// centres, and work items of lines [a, z) cut at multiples of its chunk K (absolute line
// indices; K from its own ranges), so its sums are the same alone and in any batch.
int launch_f64_batch_server(gx_server* S, Slot& sl) {
  const int B = (int)sl.reqs.size();
  const LineStore& ls = S->L->ls;
  const std::vector<double>& hnu = S->L->h_nu;
  hipStream_t s = sl.stream;
  hipError_t e;
  int rc;
  sl.f64 = true;
  sl.h_pcs.resize(B);
  sl.out_off.assign(B, 0);
  std::vector<ReqGeo> geo(B);
  size_t bins = 0, lines = 0;
  for (int k = 0; k < B; ++k) {
    const gx_request& q = sl.reqs[k]->r;
    sl.h_pcs[k] = request_consts(S, q);
    geo[k] = request_geo(S, q, sl.h_pcs[k]);
    sl.out_off[k] = bins;
    bins += (size_t)q.npoints;
    lines += geo[k].hi - geo[k].lo;
  }
  if ((rc = grow(sl.freq, sl.freq_cap, bins, "server: grids")) || (rc = grow(sl.out, sl.out_cap, bins, "server: sums")) ||
      (rc = grow(sl.evals, sl.evals_cap, (size_t)B * kEvalSlots, "server: counters")))
    return rc;
  if ((rc = sl.der.reserve(std::max<size_t>(lines, 1)))) return rc;
  sl.h_t64.assign(B, F64Task{});
  sl.h_items.clear();
  sl.h_bitems.clear();
  size_t dofs = 0;
  long long prep = 0;
  int gblock = 0;
  for (int k = 0; k < B; ++k) {
    const gx_request& q = sl.reqs[k]->r;
    const ReqGeo& g = geo[k];
    const int M = q.npoints;
    F64Task& t = sl.h_t64[k];
    t.pc = sl.h_pcs[k];
    t.gv = GridView{sl.freq + sl.out_off[k], M, 0, M, q.start_nu, g.d};
    t.nu = ls.nu;
    t.line_lo = (int)g.lo;
    t.nlines = (int)(g.hi - g.lo);
    t.d = DerivedF64{sl.der.d.S + dofs, sl.der.d.g + dofs, sl.der.d.y + dofs, sl.der.d.ib + dofs, sl.der.d.ie + dofs};
    t.out = sl.out + sl.out_off[k];
    t.prep_first = prep;
    t.first_block = gblock;
    t.zero_nu = f64_zero_nu(t.pc);
    t.x0 = f64_x0(t.pc, g.d);
    t.evals = sl.evals + (size_t)k * kEvalSlots;
    dofs += (size_t)t.nlines;
    prep += ((long long)t.nlines + 255) / 256 * 256;
    // block ranges (the lines that can reach a block, k_block_ranges' rule) on the host
    const int nb = (M + kBinsF64 - 1) / kBinsF64;
    std::vector<int2> rg(nb);
    long long total = 0;
    for (int b = 0; b < nb; ++b) {
      const int j0 = b * kBinsF64, j1 = std::min(j0 + kBinsF64, M);
      const double lo_nu = q.start_nu + (double)(j0 - 2) * g.d - g.cut;
      const double hi_nu = q.start_nu + (double)(j1 + 1) * g.d + g.cut;
      const int a = (int)(std::lower_bound(hnu.begin() + g.lo, hnu.begin() + g.hi, lo_nu) - hnu.begin());
      const int z = (int)(std::upper_bound(hnu.begin() + a, hnu.begin() + g.hi, hi_nu) - hnu.begin());
      rg[b] = make_int2(a, z);
      total += z - a;
    }
    const long long K = std::max(kFp64Chunk, (total / 2048 + kFp64Chunk - 1) / kFp64Chunk * kFp64Chunk);
    for (int b = 0; b < nb; ++b) {
      const int first = (int)sl.h_items.size();
      const long long a = rg[b].x, z = rg[b].y;
      if (a < z)
        for (long long c = a / K; c <= (z - 1) / K; ++c)
          sl.h_items.push_back(F64Item{k, b, (int)std::max(a, c * K), (int)std::min(z, (c + 1) * K)});
      sl.h_bitems.push_back(make_int2(first, (int)sl.h_items.size() - first));
    }
    gblock += nb;
  }
  const int nitems = (int)sl.h_items.size();
  if ((rc = grow(sl.partial, sl.partial_cap, (size_t)std::max(nitems, 1) * kBinsF64, "server: partials")) ||
      (rc = grow_copy(sl.d_t64, sl.t64_cap, sl.h_t64, s, "server: tasks")) ||
      (rc = grow_copy(sl.d_items, sl.items_cap, sl.h_items, s, "server: items")) ||
      (rc = grow_copy(sl.d_bitems, sl.bitems_cap, sl.h_bitems, s, "server: blocks")))
    return rc;
  if ((e = hipEventRecord(sl.ev_start, s))) return hip_fail(e, "server: event");
  for (int k = 0; k < B; ++k) {
    const gx_request& q = sl.reqs[k]->r;
    if ((e = launch_grid(sl.freq + sl.out_off[k], q.npoints, q.start_nu, geo[k].d, s))) return hip_fail(e, "server: grid");
  }
  if ((e = launch_fill(sl.out, bins, 0.0, s)) ||
      (e = launch_fill(reinterpret_cast<double*>(sl.evals), (size_t)B * kEvalSlots, 0.0, s)))
    return hip_fail(e, "server: fill");
  const LinesView all{ls.nu, ls.elow, ls.aif, ls.gns, ls.gam, ls.nexp, (int)ls.n};
  if ((e = launch_f64_batch(sl.d_t64, B, prep, all, sl.d_items, nitems, gblock, sl.d_bitems, sl.partial, s)))
    return hip_fail(e, "server: FP64 batch");
  if ((e = hipEventRecord(sl.ev_done, s))) return hip_fail(e, "server: event");
  return GX_OK;
}

int launch_batch(gx_server* S, Slot& sl) {
  if (fp64_req(sl.reqs[0]->r)) return launch_f64_batch_server(S, sl);
  const gx_request& r0 = sl.reqs[0]->r;
  const int M = r0.npoints;
  const int B = (int)sl.reqs.size();
  const double d = (r0.end_nu - r0.start_nu) / double(M);     // BaseProfile.cpp:10
  hipStream_t s = sl.stream;
  hipError_t e;
  int rc;
  // per-request constants
  sl.h_pcs.resize(B);
  for (int k = 0; k < B; ++k) sl.h_pcs[k] = request_consts(S, sl.reqs[k]->r);
  const PointConsts& pc0 = sl.h_pcs[0];
  const double cmax = batch_cutmax(S, sl.reqs);
  size_t lo, hi;
  line_slice(S->L, r0.start_nu, r0.start_nu + M * d, cmax, d, lo, hi);
  const int n = (int)(hi - lo);
  const LineStore& ls = S->L->ls;
  LinesView lv{ls.nu + lo, ls.elow + lo, ls.aif + lo, ls.gns + lo, ls.gam ? ls.gam + lo : nullptr,
               ls.nexp ? ls.nexp + lo : nullptr, n};
  if ((rc = grow(sl.freq, sl.freq_cap, (size_t)M, "server: grid"))) return rc;
  if ((rc = grow(sl.out, sl.out_cap, (size_t)M * B, "server: grids"))) return rc;
  if (n > 0 && (size_t)n > sl.win_cap) {
    (void)hipFree(sl.ib);
    (void)hipFree(sl.ie);
    sl.ib = sl.ie = nullptr;
    sl.win_cap = 0;
    if ((e = hipMalloc(&sl.ib, 4 * (size_t)n)) || (e = hipMalloc(&sl.ie, 4 * (size_t)n)))
      return hip_fail(e, "server: windows");
    sl.win_cap = n;
  }
  if ((rc = grow(sl.evals, sl.evals_cap, (size_t)kEvalSlots, "server: counters"))) return rc;
  sl.f64 = false;
  sl.out_off.resize(B);
  for (int k = 0; k < B; ++k) sl.out_off[k] = (size_t)k * M;
  const GridView gv{sl.freq, M, 0, M, r0.start_nu, d};
  if ((rc = sl.der.reserve_mixed((size_t)std::max(n, 1), B, mixed_temperature_rows(sl.h_pcs.data(), B, nullptr))))
    return rc;
  if ((rc = sl.der.set_points(sl.h_pcs.data(), B, s))) return rc;
  // launch shape from the request's own grid only (alone == in a mix)
  if ((rc = sl.run.init())) return rc;   // fixes R (GX_MIXED_R) before the block size is used
  const int bpb = sl.run.bins_per_block();
  const int nblocks = (M + bpb - 1) / bpb;
  if ((rc = sl.rng.reserve(nblocks))) return rc;
  {
    const int slots = sl.run.cus * mixed_occupancy(sl.run.R) * kWavesPerGroup;
    sl.run.forced_nseg = std::min(16, std::max(1, (4 * slots + nblocks - 1) / nblocks));
    sl.run.persistent = false;
    std::vector<int> trow(B);
    mixed_temperature_rows(sl.h_pcs.data(), B, trow.data());
    std::vector<MixedTask> tasks(B);
    for (int k = 0; k < B; ++k)
      tasks[k] = mixed_task(sl.der.m, k, trow[k], sl.h_pcs[k], lv.gamma, sl.out + (size_t)k * M);
    if ((rc = sl.run.prepare(tasks.data(), B, nblocks, gv, n, s))) return rc;
  }
  if ((e = hipEventRecord(sl.ev_start, s))) return hip_fail(e, "server: event");
  if ((e = launch_grid(sl.freq, M, r0.start_nu, d, s))) return hip_fail(e, "server: grid");
  if ((e = launch_fill(sl.out, (size_t)M * B, 0.0, s))) return hip_fail(e, "server: fill");
  if ((e = launch_fill(reinterpret_cast<double*>(sl.evals), kEvalSlots, 0.0, s))) return hip_fail(e, "server: fill");
  if (n > 0) {
    const WindowArgs wa{pc0.cutoff, pc0.convention, 1};
    if ((e = launch_windows(wa, gv, lv.nu, n, sl.ib, sl.ie, sl.evals, s))) return hip_fail(e, "server: windows");
    if ((e = launch_prep_mixed(sl.der.d_pcs, B, lv, gv, sl.ib, sl.ie, sl.der.m, nullptr, s)))
      return hip_fail(e, "server: prep");
    if ((e = launch_block_ranges(gv, lv.nu, n, cmax, bpb, nblocks, sl.rng.r, s))) return hip_fail(e, "server: ranges");
    if ((rc = sl.run.launch(gv, lv.nu, sl.rng.r, sl.ib, sl.ie, s))) return rc;
  }
  if ((e = hipEventRecord(sl.ev_done, s))) return hip_fail(e, "server: event");
  return GX_OK;
}

void retire(gx_server* S, Slot& sl) {
  const int B = (int)sl.reqs.size();
  float ms = 0;
  (void)hipEventElapsedTime(&ms, sl.ev_start, sl.ev_done);
  // evaluation counters: one set per request (batched FP64 path) or one for the batch
  const int nsets = sl.f64 ? B : 1;
  std::vector<unsigned long long> slots((size_t)nsets * kEvalSlots, 0ull);
  hipError_t e = hipMemcpy(slots.data(), sl.evals, sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost);
  std::vector<long long> ev(B, 0);
  for (int k = 0; k < B; ++k)
    for (int q = 0; q < kEvalSlots; ++q) ev[k] += (long long)slots[(size_t)(sl.f64 ? k : 0) * kEvalSlots + q];
  std::vector<std::vector<double>> results(B);
  bool ok = e == hipSuccess;
  for (int k = 0; k < B && ok; ++k) {
    if (sl.reqs[k]->cancel) continue;
    const int M = sl.reqs[k]->r.npoints;
    results[k].resize(M);
    if (hipMemcpy(results[k].data(), sl.out + sl.out_off[k], sizeof(double) * M, hipMemcpyDeviceToHost)) ok = false;
  }
  const auto now = Clock::now();
  std::lock_guard<std::mutex> lk(S->mu);
  for (int k = 0; k < B; ++k) {
    Req* q = sl.reqs[k];
    q->t_done = now;
    q->kernel_ms = ms;
    if (q->cancel) {
      q->state = GX_REQ_CANCELLED;
      ++S->cancelled;
    } else if (!ok) {
      q->state = GX_REQ_FAILED;
      q->error = "device copy failed";
      ++S->failed;
    } else {
      q->host.swap(results[k]);
      q->evals = ev[k];
      q->state = GX_REQ_DONE;
      ++S->completed;
      S->evals += ev[k];
      S->latencies_ms.push_back(std::chrono::duration<double, std::milli>(now - q->t_submit).count());
    }
  }
  S->t_last = now;
  sl.reqs.clear();
  sl.busy = false;
  S->cv_done.notify_all();
}

void scheduler(gx_server* S) {
  (void)hipSetDevice(S->L->device);
  for (;;) {
    bool progressed = false;
    // 1. retire finished batches
    for (auto& up : S->slots) {
      Slot& sl = *up;
      if (sl.busy && hipEventQuery(sl.ev_done) == hipSuccess) {
        retire(S, sl);
        progressed = true;
      }
    }
    // 2. admit
    for (auto& up : S->slots) {
      Slot& sl = *up;
      if (sl.busy) continue;
      std::vector<Req*> batch;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        while (!S->pending.empty()) {
          auto it = S->reqs.find(S->pending.front());
          if (it == S->reqs.end() || it->second->cancel) {
            if (it != S->reqs.end()) {
              it->second->state = GX_REQ_CANCELLED;
              ++S->cancelled;
              S->cv_done.notify_all();
            }
            S->pending.pop_front();
            continue;
          }
          break;
        }
        if (S->pending.empty()) break;
        Req* head = S->reqs[S->pending.front()].get();
        batch.push_back(head);
        for (size_t i = 1; i < S->pending.size() && (int)batch.size() < S->o.max_batch; ++i) {
          auto it = S->reqs.find(S->pending[i]);
          if (it != S->reqs.end() && !it->second->cancel && compatible(it->second->r, head->r))
            batch.push_back(it->second.get());
        }
        // HBM admission: footprint of the other slots plus this slot after growth
        const gx_request& r = head->r;
        const double d = (r.end_nu - r.start_nu) / double(r.npoints);
        size_t lo, hi;
        line_slice(S->L, r.start_nu, r.end_nu, batch_cutmax(S, batch), d, lo, hi);
        size_t others = 0;
        for (auto& o2 : S->slots) if (o2.get() != &sl) others += o2->footprint();
        size_t need = batch_need(S, sl, batch, hi - lo);
        while (batch.size() > 1 && others + need > S->budget) {   // shrink the batch first
          batch.pop_back();
          need = batch_need(S, sl, batch, hi - lo);
        }
        if (others + need > S->budget) {
          // idle slots keep their grown buffers: let them go, largest first, before waiting
          std::vector<Slot*> idle;
          for (auto& o2 : S->slots)
            if (o2.get() != &sl && !o2->busy && o2->footprint() > 0) idle.push_back(o2.get());
          std::sort(idle.begin(), idle.end(), [](Slot* a, Slot* b) { return a->footprint() > b->footprint(); });
          for (Slot* o2 : idle) {
            if (others + need <= S->budget) break;
            others -= o2->footprint();
            o2->release_buffers();
          }
        }
        if (others + need > S->budget) {
          bool any_busy = false;
          for (auto& o2 : S->slots) any_busy |= o2->busy;
          if (any_busy) break;                                     // wait for memory to free up
          if (need > S->budget) {
            head->state = GX_REQ_FAILED;
            head->error = "request does not fit the HBM budget";
            ++S->failed;
            S->pending.pop_front();
            S->cv_done.notify_all();
            progressed = true;
            continue;
          }
        }
        for (Req* q : batch) {
          q->state = GX_REQ_RUNNING;
          S->pending.erase(std::find(S->pending.begin(), S->pending.end(), q->id));
        }
        if (!S->started) {
          S->started = true;
          S->t_first = Clock::now();
        }
        ++S->batches;
        S->batched += (long long)batch.size();
        if (batch.size() > 1) S->shared += (long long)batch.size();
        S->peak_bytes = std::max(S->peak_bytes, others + need);
      }
      sl.reqs = batch;
      sl.busy = true;
      sl.t_launch = Clock::now();
      const int rc = launch_batch(S, sl);
      if (rc) {
        std::string msg = gx_last_error();
        (void)hipStreamSynchronize(sl.stream);
        std::lock_guard<std::mutex> lk(S->mu);
        for (Req* q : sl.reqs) {
          q->state = GX_REQ_FAILED;
          q->error = msg;
          ++S->failed;
        }
        sl.reqs.clear();
        sl.busy = false;
        S->cv_done.notify_all();
      }
      progressed = true;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    bool any_busy = false;
    for (auto& up : S->slots) any_busy |= up->busy;
    if (S->stop && !any_busy && S->pending.empty()) break;
    if (!progressed) {
      if (any_busy)
        S->cv_work.wait_for(lk, std::chrono::microseconds(100));
      else
        S->cv_work.wait_for(lk, std::chrono::milliseconds(20), [&] { return S->stop || !S->pending.empty(); });
    }
  }
}

}  // namespace

extern "C" {

int gx_server_create(gx_lines* L, const gx_server_opts* o, gx_server** out) {
  if (!out) return fail(GX_INVALID_ARG, "server_create: out is NULL");
  *out = nullptr;
  if (!L || !o || !(o->mean_mass > 0) || !(o->ref_temp > 0)) return fail(GX_INVALID_ARG, "server_create: bad arguments");
  hipError_t e;
  if ((e = hipSetDevice(L->device))) return hip_fail(e, "hipSetDevice");
  gx_server* S = new gx_server();
  S->L = L;
  S->o = *o;
  if (S->o.max_inflight <= 0) S->o.max_inflight = 4;
  if (S->o.max_batch <= 0) S->o.max_batch = 16;
  size_t free_b = 0, total_b = 0;
  if ((e = hipMemGetInfo(&free_b, &total_b))) { delete S; return hip_fail(e, "hipMemGetInfo"); }
  const size_t reserve = size_t(4) << 30;
  S->budget = o->hbm_budget ? o->hbm_budget : (free_b > reserve ? free_b - reserve : free_b / 2);
  for (int k = 0; k < S->o.max_inflight; ++k) {
    auto sl = std::make_unique<Slot>();
    if ((e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking)) || (e = hipEventCreate(&sl->ev_start)) ||
        (e = hipEventCreate(&sl->ev_done))) {
      delete S;
      return hip_fail(e, "server_create: streams");
    }
    S->slots.push_back(std::move(sl));
  }
  S->worker = std::thread(scheduler, S);
  *out = S;
  return GX_OK;
}

int gx_server_submit(gx_server* S, const gx_request* r, uint64_t* id) {
  if (!S || !r || !id) return fail(GX_INVALID_ARG, "server_submit: NULL argument");
  if (r->profile != GX_VOIGT && r->profile != GX_DOPPLER) return fail(GX_INVALID_ARG, "server_submit: bad profile");
  if (r->profile == GX_VOIGT && !S->L->voigt) return fail(GX_INVALID_ARG, "server_submit: lines have no broadening");
  if (r->npoints <= 0 || !(r->end_nu > r->start_nu) || !(r->temperature > 0) || !(r->partition > 0) ||
      !std::isfinite(r->pressure))
    return fail(GX_INVALID_ARG, "server_submit: bad request");
  auto q = std::make_unique<Req>();
  q->r = *r;
  if (q->r.profile == GX_DOPPLER) q->r.precision = GX_PREC_FP64;
  if (mixed_req(q->r)) {
    // a grid too fine for the mixed path's W4 core pass runs on the FP64 path (decided from
    // the request alone, so its result does not depend on what it is batched with)
    gx_consts c{};
    c.temperature = r->temperature;
    c.mean_mass = S->o.mean_mass;
    c.convention = S->o.convention;
    const double d = (r->end_nu - r->start_nu) / double(r->npoints);
    const GridView gv{nullptr, r->npoints, 0, r->npoints, r->start_nu, d};
    if (!core_pass_fits(gv, make_point_consts(r->profile, c, r->pressure).dpwcoef)) q->r.precision = GX_PREC_FP64;
  }
  q->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return fail(GX_BUSY, "server_submit: server is stopping");
    q->id = S->next_id++;
    *id = q->id;
    S->pending.push_back(q->id);
    S->reqs[q->id] = std::move(q);
    ++S->submitted;
  }
  S->cv_work.notify_one();
  return GX_OK;
}

int gx_server_poll(gx_server* S, uint64_t id, int* state) {
  if (!S || !state) return fail(GX_INVALID_ARG, "server_poll: NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(GX_NOT_FOUND, "server_poll: unknown request");
  *state = it->second->state;
  return GX_OK;
}

int gx_server_wait(gx_server* S, uint64_t id, double timeout_s, int* state) {
  if (!S) return fail(GX_INVALID_ARG, "server_wait: NULL server");
  std::unique_lock<std::mutex> lk(S->mu);
  auto done = [&] {
    auto it = S->reqs.find(id);
    return it == S->reqs.end() || it->second->state >= GX_REQ_DONE;
  };
  if (timeout_s < 0)
    S->cv_done.wait(lk, done);
  else
    S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(GX_NOT_FOUND, "server_wait: unknown request");
  if (state) *state = it->second->state;
  return GX_OK;
}

int gx_server_result(gx_server* S, uint64_t id, double* out, int n, gx_timing* t) {
  if (!S || !out) return fail(GX_INVALID_ARG, "server_result: NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return fail(GX_NOT_FOUND, "server_result: unknown request");
  Req* q = it->second.get();
  if (q->state == GX_REQ_FAILED) {
    const std::string msg = "server_result: request failed: " + q->error;
    S->reqs.erase(it);   // released: a long-running server keeps no record of it
    return fail(GX_DEVICE_ERROR, msg);
  }
  if (q->state == GX_REQ_CANCELLED) {
    S->reqs.erase(it);
    return fail(GX_NOT_FOUND, "server_result: request was cancelled");
  }
  if (q->state != GX_REQ_DONE) return fail(GX_BUSY, "server_result: request not finished");
  if (n != q->r.npoints) return fail(GX_INVALID_ARG, "server_result: n differs from the request's npoints");
  std::copy(q->host.begin(), q->host.end(), out);
  if (t) {
    t->intensity_ms = 0;
    t->accumulate_ms = q->kernel_ms;
    t->total_ms = (float)std::chrono::duration<double, std::milli>(q->t_done - q->t_submit).count();
    t->evals = q->evals;
    t->lines = (long long)S->L->ls.n;
    t->region1_ms = q->kernel_ms;
    t->core_ms = 0.0f;
    t->core_evals = 0;
    t->far_ms = 0.0f;
    t->far_evals = 0;
    t->interp_pairs = 0;
  }
  S->reqs.erase(it);
  return GX_OK;
}

int gx_server_cancel(gx_server* S, uint64_t id) {
  if (!S) return fail(GX_INVALID_ARG, "server_cancel: NULL server");
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(GX_NOT_FOUND, "server_cancel: unknown request");
    if (it->second->state >= GX_REQ_DONE) return GX_OK;
    it->second->cancel = true;
  }
  S->cv_work.notify_one();
  return GX_OK;
}

int gx_server_stats(gx_server* S, gx_server_summary* out) {
  if (!S || !out) return fail(GX_INVALID_ARG, "server_stats: NULL argument");
  std::lock_guard<std::mutex> lk(S->mu);
  *out = gx_server_summary{};
  out->submitted = S->submitted;
  out->completed = S->completed;
  out->cancelled = S->cancelled;
  out->failed = S->failed;
  out->batches = S->batches;
  out->batched_requests = S->batched;
  out->evals = S->evals;
  out->peak_bytes = S->peak_bytes;
  out->shared_requests = S->shared;
  if (S->started) out->busy_s = std::chrono::duration<double>(S->t_last - S->t_first).count();
  std::vector<double> v = S->latencies_ms;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      const double pos = p * (double)(v.size() - 1);
      const size_t i = (size_t)pos;
      const double f = pos - (double)i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    double s = 0;
    for (double x : v) s += x;
    out->mean_ms = s / (double)v.size();
    out->max_ms = v.back();
  }
  return GX_OK;
}

int gx_server_destroy(gx_server* S) {
  if (!S) return GX_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    for (auto id : S->pending) {
      auto it = S->reqs.find(id);
      if (it != S->reqs.end()) it->second->cancel = true;
    }
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto& sl : S->slots) {
    (void)hipStreamSynchronize(sl->stream);
    sl->release_buffers();
    (void)hipFree(sl->evals);
    (void)hipEventDestroy(sl->ev_start);
    (void)hipEventDestroy(sl->ev_done);
    (void)hipStreamDestroy(sl->stream);
  }
  delete S;
  return GX_OK;
}

}  // extern "C"
