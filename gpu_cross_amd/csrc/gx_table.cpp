// gx_table.cpp -- batched (T,P) opacity tables over one resident line list.
//
// The reference computes one (T,P) point per process: InitializeConstants is
// called once (VoigtProfile.cpp:26) and every cross-section re-reads and
// re-uploads its lines.  Here the sorted line list stays in HBM (gx_lines)
// and a table of points is planned once (gx_table_create):
//
//   fill(out = 0) -> windows per cutoff group -> block ranges per group
//   -> per-point intensities of all points (one 2-D launch)
//   -> ONE persistent accumulation launch whose work queue spans every
//      (point, bin block, line segment) item of the table
//
// and that sequence is captured into a hipGraph on first use and replayed by
// gx_table_run.  Points are admitted by free HBM: when the per-point line
// data of all points does not fit, the table runs in point batches, each its
// own persistent launch inside the same graph.  An autotuner times the
// launch shape (bins per block, work split) on the table itself once per shape.
//
// The FP64 precision of a table is the plain per-point path: one intensity,
// ranges and accumulation launch per point, the denominator of the speedup
// the mixed path is measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/gx.h"
#include "gx_common.h"

using namespace gx;

namespace {

struct Group {                // points that share one cutoff (hence windows and block ranges)
  double cutoff = 0;
  double cutmax = 0;          // widest per-line window of its points (Doppler k*alpha + dfreq can exceed cutoff)
  std::vector<int> points;
  int first_const = 0;        // index of points[0] in the grouped constants table
  int* ib = nullptr;
  int* ie = nullptr;
  Ranges rng;                 // mixed block ranges
  Ranges rng64;               // FP64 block ranges
  int nblocks = 0;
  int nblocks64 = 0;
  unsigned long long* evals = nullptr;   // kEvalSlots
};

struct Launch {               // one persistent accumulation launch (a batch of a group)
  int group = 0;
  int first = 0;              // index into the group's point list
  int count = 0;
  MixedRunner run;
  hipEvent_t ev[3] = {};      // timed runs: region I | core pass
};

// autotune cache: (bins, lines, points per launch) -> (R, items per slot)
std::map<std::tuple<int, size_t, int, int>, std::pair<int, int>> g_tuned;

}  // namespace

struct gx_table {
  gx_lines* L = nullptr;
  gx_grid grid{};
  double dfreq = 0;
  int M = 0;
  std::vector<gx_point> pts;
  gx_table_opts o{};
  std::vector<PointConsts> pcs;
  double* d_freq = nullptr;
  double* d_out = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  std::vector<Group> groups;
  std::vector<Launch> launches;
  Derived der;                // mixed line data for `batch` points, all points' constants
  Derived der64;              // FP64 path
  int batch = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_ran = false;
  float eager_intensity_ms = 0, eager_accumulate_ms = 0;
  float eager_region1_ms = 0, eager_core_ms = 0, eager_far_ms = 0;
  unsigned long long eager_far_evals = 0;
  unsigned long long* core_evals = nullptr;   // kEvalSlots counters of the core pass
  int R = 1, ips = 4;          // launch shape without the autotuner (the one it picks for tables)
};

namespace {

bool mixed_table(const gx_table* t) { return t->o.precision == GX_PREC_MIXED && t->o.profile == GX_VOIGT; }

GridView grid_view(const gx_table* t) {
  return GridView{t->d_freq, t->M, 0, t->M, t->grid.start_nu, t->dfreq};
}

LinesView lines_view(const gx_table* t) {
  const LineStore& ls = t->L->ls;
  return LinesView{ls.nu, ls.elow, ls.aif, ls.gns, ls.gam, ls.nexp, (int)ls.n};
}

// Everything one run enqueues; graph-capturable (kernels only, no memset, no host sync).
// timed: record ev[0..2] around the phases (only outside capture).
int enqueue(gx_table* t, bool timed) {
  hipStream_t s = t->stream;
  hipError_t e;
  const GridView gv = grid_view(t);
  const LinesView lv = lines_view(t);
  if (timed && (e = hipEventRecord(t->ev[0], s))) return hip_fail(e, "event");
  if ((e = launch_fill(t->d_out, (size_t)t->M * t->pts.size(), 0.0, s))) return hip_fail(e, "fill");
  if (t->core_evals && (e = launch_fill(reinterpret_cast<double*>(t->core_evals), kEvalSlots, 0.0, s)))
    return hip_fail(e, "fill");
  const bool mixed = mixed_table(t);
  for (Group& g : t->groups) {
    if ((e = launch_fill(reinterpret_cast<double*>(g.evals), kEvalSlots, 0.0, s))) return hip_fail(e, "fill");
    if (!mixed) continue;   // the FP64 path's per-point prep forms (and counts) each point's own windows
    const WindowArgs wa{g.cutoff, t->o.convention, 1};
    if ((e = launch_windows(wa, gv, lv.nu, lv.n, g.ib, g.ie, g.evals, s))) return hip_fail(e, "windows");
  }
  if (mixed) {
    for (Group& g : t->groups) {
      const double cut = g.cutoff;
      if ((e = launch_block_ranges(gv, lv.nu, lv.n, cut, mixed_bins_per_block(t->R), g.nblocks, g.rng.r, s)))
        return hip_fail(e, "ranges");
    }
    bool first = true;
    for (Launch& l : t->launches) {
      const Group& g = t->groups[l.group];
      // points of one launch are contiguous in the constants table (grouped at create)
      const PointConsts* pc0 = t->der.d_pcs + g.first_const + l.first;
      if ((e = launch_prep_mixed(pc0, l.count, lv, gv, g.ib, g.ie, t->der.m, t->core_evals, s)))
        return hip_fail(e, "prep");
      if (first && timed && (e = hipEventRecord(t->ev[1], s))) return hip_fail(e, "event");
      first = false;
      int rc = l.run.launch(gv, lv.nu, g.rng.r, g.ib, g.ie, s, timed && l.ev[0] ? l.ev : nullptr);
      if (rc) return rc;
    }
  } else {
    if (timed && (e = hipEventRecord(t->ev[1], s))) return hip_fail(e, "event");
    for (Group& g : t->groups) {
      if ((e = launch_block_ranges(gv, lv.nu, lv.n, g.cutmax, kBinsF64, g.nblocks64, g.rng64.r, s)))
        return hip_fail(e, "ranges");
      for (int p : g.points) {
        if ((e = launch_line_prep_f64(t->pcs[p], gv, lv, t->der64.d, g.evals, s))) return hip_fail(e, "prep");
        if ((e = launch_accum_f64(t->pcs[p], gv, lv.nu, t->der64.d, g.rng64.r, g.nblocks64, t->d_out + (size_t)p * t->M,
                                  s)))
          return hip_fail(e, "accumulate");
      }
    }
  }
  if (timed && (e = hipEventRecord(t->ev[2], s))) return hip_fail(e, "event");
  return GX_OK;
}

// GX_MIXED_R (1, 2, 4 or 8): a fixed block size of 64 R bins, else 0.
int env_R() {
  if (const char* env = std::getenv("GX_MIXED_R")) {
    const int r = std::atoi(env);
    if (r == 1 || r == 2 || r == 4 || r == 8) return r;
  }
  return 0;
}

int prepare_launches(gx_table* t) {
  for (Launch& l : t->launches) {
    const Group& g = t->groups[l.group];
    int rc0 = l.run.init();                     // before R is set: init applies GX_MIXED_R to a runner's default
    if (rc0) return rc0;
    l.run.R = t->R;
    l.run.items_per_slot = t->ips;
    l.run.core_counter = t->core_evals;
    // the launch's points are contiguous in the grouped constants k_prep_mixed reads
    std::vector<int> trow(l.count);
    mixed_temperature_rows(t->der.h_pcs.data() + g.first_const + l.first, l.count, trow.data());
    std::vector<MixedTask> tasks(l.count);
    for (int k = 0; k < l.count; ++k) {
      const int p = g.points[l.first + k];
      tasks[k] = mixed_task(t->der.m, k, trow[k], t->pcs[p], lines_view(t).gamma, t->d_out + (size_t)p * t->M);
    }
    const int nb = (t->M + mixed_bins_per_block(t->R) - 1) / mixed_bins_per_block(t->R);
    int rc = l.run.prepare(tasks.data(), l.count, nb, grid_view(t), (int)t->L->ls.n, t->stream);
    if (rc) return rc;
  }
  for (Group& g : t->groups) {
    g.nblocks = (t->M + mixed_bins_per_block(t->R) - 1) / mixed_bins_per_block(t->R);
    int rc = g.rng.reserve(g.nblocks);
    if (rc) return rc;
  }
  return GX_OK;
}

// Time each launch shape on the table itself (every launch, best of two runs: the far field,
// region I and the core pass share the work differently at each block size, and a table's
// points differ widely in how their pairs split between them) and keep the fastest; the choice
// is cached per (bins, lines, points per launch).
int autotune(gx_table* t) {
  const int forced_R = env_R();                 // GX_MIXED_R restricts the choice to that block size
  const auto key = std::make_tuple(t->M, t->L->ls.n, t->launches.empty() ? 0 : t->launches[0].count, forced_R);
  auto it = g_tuned.find(key);
  if (it == g_tuned.end() && !t->launches.empty()) {
    float best = 1e30f;
    int bR = 2, bI = 4;
    int last_R = -1;
    long long last_shape = -1;
    int rc = GX_OK;
    const std::vector<int> cands = forced_R ? std::vector<int>{forced_R} : std::vector<int>{1, 2, 4};
    for (int R : cands) {
      for (int ips : {4, 8, 2}) {
        t->R = R;
        t->ips = ips;
        if ((rc = prepare_launches(t))) break;
        long long shape = 0;                            // the launch shape: segments of every launch
        for (const Launch& l : t->launches) shape = shape * 17 + l.run.plan.nseg;
        if (R == last_R && shape == last_shape) continue;   // same launch shape as a timed one
        last_R = R;
        last_shape = shape;
        float ms = 1e30f;
        for (int rep = 0; rep < 2 && !rc; ++rep) {
          hipError_t e;
          if ((rc = enqueue(t, true))) break;
          if ((e = hipStreamSynchronize(t->stream))) { rc = hip_fail(e, "autotune"); break; }
          float m = 0;
          if ((e = hipEventElapsedTime(&m, t->ev[1], t->ev[2]))) { rc = hip_fail(e, "autotune events"); break; }
          ms = std::min(ms, m);
        }
        if (rc) break;
        if (ms < best) {
          best = ms;
          bR = R;
          bI = ips;
        }
      }
      if (rc) break;
    }
    if (rc) return rc;
    g_tuned[key] = {bR, bI};
    it = g_tuned.find(key);
  }
  if (it != g_tuned.end()) {
    t->R = it->second.first;
    t->ips = it->second.second;
  }
  return prepare_launches(t);
}

// Region-I and core-pass times of the last timed run, summed over the launches.
int split_times(gx_table* t, float* region1, float* core, float* far = nullptr,
                unsigned long long* far_evals = nullptr) {
  *region1 = *core = 0.0f;
  if (far) *far = 0.0f;
  if (far_evals) *far_evals = 0;
  if (!mixed_table(t)) return GX_OK;
  for (Launch& l : t->launches) {
    if (!l.ev[0] || l.run.plan.nitems <= 0) continue;
    float a = 0, b = 0;
    hipError_t e;
    if ((e = hipEventElapsedTime(&a, l.ev[0], l.ev[1])) || (e = hipEventElapsedTime(&b, l.ev[1], l.ev[2])))
      return hip_fail(e, "table: events");
    *region1 += a;
    *core += b;
    if (far) {
      float fm = 0;
      unsigned long long fe = 0;
      const int rc = l.run.far_time(&fm, &fe);
      if (rc) return rc;
      *far += fm;
      if (far_evals) *far_evals += fe;
    }
  }
  return GX_OK;
}

void free_table(gx_table* t) {
  if (!t) return;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  if (t->exec) (void)hipGraphExecDestroy(t->exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  for (Group& g : t->groups) {
    (void)hipFree(g.ib);
    (void)hipFree(g.ie);
    (void)hipFree(g.evals);
    g.rng.release();
    g.rng64.release();
  }
  for (Launch& l : t->launches) {
    l.run.release();
    for (auto& ev : l.ev)
      if (ev) (void)hipEventDestroy(ev);
  }
  (void)hipFree(t->core_evals);
  t->der.release();
  t->der64.release();
  (void)hipFree(t->d_freq);
  (void)hipFree(t->d_out);
  for (auto& ev : t->ev)
    if (ev) (void)hipEventDestroy(ev);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

}  // namespace

extern "C" {

int gx_lines_create(int device, size_t n, const double* energies, const double* nu, const double* aif, const int* gns,
                    const double* gamma_ref, const double* nexp, gx_lines** out) {
  if (!out) return fail(GX_INVALID_ARG, "lines_create: out is NULL");
  *out = nullptr;
  if (n == 0 || !energies || !nu || !aif || !gns) return fail(GX_INVALID_ARG, "lines_create: empty or NULL arrays");
  if ((gamma_ref == nullptr) != (nexp == nullptr)) return fail(GX_INVALID_ARG, "lines_create: gamma_ref and nexp go together");
  if (n > (size_t)INT32_MAX) return fail(GX_INVALID_ARG, "lines_create: more than 2^31-1 lines");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GX_NO_DEVICE, "lines_create: no HIP device");
  if (device < 0 || device >= ndev) return fail(GX_NO_DEVICE, "lines_create: device index out of range");
  hipError_t e;
  if ((e = hipSetDevice(device))) return hip_fail(e, "hipSetDevice");
  gx_lines* L = new gx_lines();
  L->device = device;
  L->voigt = gamma_ref != nullptr;
  if ((e = hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking))) { delete L; return hip_fail(e, "stream"); }
  int rc = L->ls.upload(n, energies, nu, aif, gns, gamma_ref, nexp, L->stream);
  L->ls.release_staging();   // a resident list is uploaded once: give the sort staging back
  if (rc) { gx_lines_destroy(L); return rc; }
  L->h_nu.resize(n);
  if ((e = hipMemcpy(L->h_nu.data(), L->ls.nu, n * sizeof(double), hipMemcpyDeviceToHost))) {
    gx_lines_destroy(L);
    return hip_fail(e, "lines_create: copy back");
  }
  *out = L;
  return GX_OK;
}

int gx_lines_destroy(gx_lines* L) {
  if (!L) return GX_OK;
  (void)hipSetDevice(L->device);
  if (L->stream) (void)hipStreamSynchronize(L->stream);
  L->ls.release();
  if (L->stream) (void)hipStreamDestroy(L->stream);
  delete L;
  return GX_OK;
}

size_t gx_lines_count(const gx_lines* L) { return L ? L->ls.n : 0; }

int gx_table_create(gx_lines* L, const gx_grid* grid, const gx_point* pts, int npts, const gx_table_opts* o,
                    gx_table** out) {
  if (!out) return fail(GX_INVALID_ARG, "table_create: out is NULL");
  *out = nullptr;
  if (!L || !grid || !pts || !o || npts <= 0) return fail(GX_INVALID_ARG, "table_create: bad arguments");
  if (grid->npoints <= 0 || !(grid->end_nu > grid->start_nu)) return fail(GX_INVALID_ARG, "table_create: bad grid");
  if (o->profile != GX_VOIGT && o->profile != GX_DOPPLER) return fail(GX_INVALID_ARG, "table_create: bad profile");
  if (o->profile == GX_VOIGT && !L->voigt) return fail(GX_INVALID_ARG, "table_create: Voigt needs gamma_ref and nexp");
  if (!(o->mean_mass > 0) || !(o->ref_temp > 0)) return fail(GX_INVALID_ARG, "table_create: mean_mass and ref_temp must be > 0");
  for (int p = 0; p < npts; ++p)
    if (!(pts[p].temperature > 0) || !(pts[p].partition > 0) || !std::isfinite(pts[p].pressure))
      return fail(GX_INVALID_ARG, "table_create: bad point");
  hipError_t e;
  if ((e = hipSetDevice(L->device))) return hip_fail(e, "hipSetDevice");
  gx_table* t = new gx_table();
  t->L = L;
  t->grid = *grid;
  t->M = grid->npoints;
  t->dfreq = (grid->end_nu - grid->start_nu) / double(grid->npoints);   // BaseProfile.cpp:10
  t->pts.assign(pts, pts + npts);
  t->o = *o;
  auto bail = [&](int rc) { free_table(t); return rc; };
  if ((e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking))) return bail(hip_fail(e, "stream"));
  for (auto& ev : t->ev)
    if ((e = hipEventCreate(&ev))) return bail(hip_fail(e, "event"));
  // grid nodes, BaseProfile.cpp:17-21
  std::vector<double> f(t->M);
  for (int i = 0; i < t->M; i++) f[i] = grid->start_nu + double(i) * t->dfreq;
  if ((e = hipMalloc(&t->d_freq, sizeof(double) * t->M))) return bail(hip_fail(e, "hipMalloc(grid)"));
  if ((e = hipMemcpy(t->d_freq, f.data(), sizeof(double) * t->M, hipMemcpyHostToDevice))) return bail(hip_fail(e, "grid upload"));
  if ((e = hipMalloc(&t->d_out, sizeof(double) * (size_t)t->M * npts))) return bail(hip_fail(e, "hipMalloc(table grids)"));

  // group points by cutoff; constants in group order so each launch's points are contiguous
  std::vector<int> order;
  std::map<double, int> gidx;
  for (int p = 0; p < npts; ++p) {
    const double c = reference_cutoff(o->profile, pts[p].pressure, o->cutoff);
    auto it = gidx.find(c);
    if (it == gidx.end()) {
      gidx[c] = (int)t->groups.size();
      t->groups.emplace_back();
      t->groups.back().cutoff = c;
    }
    t->groups[gidx[c]].points.push_back(p);
  }
  const size_t n = L->ls.n;
  t->pcs.resize(npts);
  std::vector<PointConsts> grouped;   // constants in group order: a launch's points are contiguous
  for (Group& g : t->groups) {
    g.first_const = (int)grouped.size();
    for (int p : g.points) {
      gx_consts c{};
      c.temperature = pts[p].temperature;
      c.partition = pts[p].partition;
      c.dfreq = t->dfreq;
      c.mean_mass = o->mean_mass;
      c.pressure = pts[p].pressure;
      c.ref_temp = o->ref_temp;
      c.cutoff = g.cutoff;
      c.doppler_widths = o->doppler_widths;
      c.convention = o->convention;
      c.precision = o->precision;
      t->pcs[p] = make_point_consts(o->profile, c, pts[p].pressure);
      grouped.push_back(t->pcs[p]);
      g.cutmax = std::max(g.cutmax, max_cutoff(t->pcs[p], L->ls.max_nu, t->dfreq));
    }
  }
  // a grid too fine for the mixed path's W4 core pass at some point runs as an FP64 table
  if (mixed_table(t))
    for (int p = 0; p < npts; ++p)
      if (!core_pass_fits(grid_view(t), t->pcs[p].dpwcoef)) t->o.precision = GX_PREC_FP64;
  for (Group& g : t->groups) {
    if ((e = hipMalloc(&g.ib, 4 * std::max<size_t>(n, 1))) || (e = hipMalloc(&g.ie, 4 * std::max<size_t>(n, 1))) ||
        (e = hipMalloc(&g.evals, sizeof(unsigned long long) * kEvalSlots)))
      return bail(hip_fail(e, "hipMalloc(windows)"));
    g.nblocks64 = (t->M + kBinsF64 - 1) / kBinsF64;
    int rc = g.rng64.reserve(g.nblocks64);
    if (rc) return bail(rc);
  }
  if (mixed_table(t)) {
    // admission: per-point FP32 line data of as many points as free HBM allows
    size_t free_b = 0, total_b = 0;
    if ((e = hipMemGetInfo(&free_b, &total_b))) return bail(hip_fail(e, "hipMemGetInfo"));
    const size_t headroom = size_t(2) << 30;
    // per line: y and the core range (16 B) per point and amp, 1/alpha (12 B) per temperature
    // (shared by the points at it); the core pass's rows (from the same geometry
    // the launch uses, at the point with the widest cores) and the far-field moments and block
    // plans; segment partials exist only while a launch has fewer items than ~4 per resident
    // wave, so they stay below 2 x 4 x 6144 items x 512 bins x 8 B = 200 MB, inside the headroom
    size_t rows = 0;
    for (int p = 0; p < npts; ++p) rows = std::max(rows, core_rows_bytes(grid_view(t), t->pcs[p].dpwcoef));
    const int temps_all = mixed_temperature_rows(grouped.data(), npts, nullptr);
    const size_t nl = std::max<size_t>(n, 1);
    const size_t per_point = kMixedPointBytes * nl + (kMixedTempBytes * nl * (size_t)temps_all + npts - 1) / npts + rows +
                             (size_t(64) << 10) +
                             far_field_bytes(1, (long long)nl, t->M / 64 + 1);
    const size_t avail = free_b > headroom ? free_b - headroom : 0;
    t->batch = (int)std::min<size_t>((size_t)npts, avail / per_point);
    if (t->batch < 1) return bail(fail(GX_OUT_OF_MEMORY, "table_create: not enough free HBM for one point"));
    int rc = t->der.set_points(grouped.data(), npts, t->stream);
    if (rc) return bail(rc);
    // launches: batches of consecutive points inside each group
    for (int gi = 0; gi < (int)t->groups.size(); ++gi) {
      const int np = (int)t->groups[gi].points.size();
      for (int f0 = 0; f0 < np; f0 += t->batch) {
        t->launches.emplace_back();
        t->launches.back().group = gi;
        t->launches.back().first = f0;
        t->launches.back().count = std::min(t->batch, np - f0);
      }
    }
    // temperature rows: the most any launch needs (each launch numbers its rows from 0)
    int trows = 1;
    for (const Launch& l : t->launches) {
      const Group& g = t->groups[l.group];
      trows = std::max(trows, mixed_temperature_rows(grouped.data() + g.first_const + l.first, l.count, nullptr));
    }
    if ((rc = t->der.reserve_mixed(n, t->batch, trows))) return bail(rc);
    if ((e = hipMalloc(&t->core_evals, sizeof(unsigned long long) * kEvalSlots)))
      return bail(hip_fail(e, "hipMalloc(core counters)"));
    for (Launch& l : t->launches)
      for (auto& ev : l.ev)
        if ((e = hipEventCreate(&ev))) return bail(hip_fail(e, "event"));
    if (const int r = env_R()) t->R = r;
    if ((rc = o->autotune ? autotune(t) : prepare_launches(t))) return bail(rc);
  } else {
    int rc = t->der64.reserve(std::max<size_t>(n, 1));
    if (rc) return bail(rc);
  }
  *out = t;
  return GX_OK;
}

int gx_table_run(gx_table* t) {
  if (!t) return fail(GX_INVALID_ARG, "table_run: NULL table");
  hipError_t e;
  if ((e = hipSetDevice(t->L->device))) return hip_fail(e, "hipSetDevice");
  if (!t->o.use_graph) {
    t->graph_ran = false;
    return enqueue(t, true);
  }
  if (!t->exec) {
    // one eager run first: its events give the kernel split the graph run cannot record
    int rc = enqueue(t, true);
    if (rc) return rc;
    if ((e = hipStreamSynchronize(t->stream))) return hip_fail(e, "table_run: eager");
    if ((e = hipEventElapsedTime(&t->eager_intensity_ms, t->ev[0], t->ev[1])) ||
        (e = hipEventElapsedTime(&t->eager_accumulate_ms, t->ev[1], t->ev[2])))
      return hip_fail(e, "table_run: events");
    if ((rc = split_times(t, &t->eager_region1_ms, &t->eager_core_ms, &t->eager_far_ms, &t->eager_far_evals)))
      return rc;
    if ((e = hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal))) return hip_fail(e, "begin capture");
    rc = enqueue(t, false);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(t->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e2) return hip_fail(e2, "end capture");
    t->graph = g;
    if ((e = hipGraphInstantiate(&t->exec, t->graph, nullptr, nullptr, 0))) return hip_fail(e, "graph instantiate");
  }
  if ((e = hipEventRecord(t->ev[0], t->stream))) return hip_fail(e, "event");
  if ((e = hipGraphLaunch(t->exec, t->stream))) return hip_fail(e, "graph launch");
  if ((e = hipEventRecord(t->ev[2], t->stream))) return hip_fail(e, "event");
  t->graph_ran = true;
  return GX_OK;
}

int gx_table_wait(gx_table* t, gx_timing* timing) {
  if (!t) return fail(GX_INVALID_ARG, "table_wait: NULL table");
  hipError_t e;
  if ((e = hipStreamSynchronize(t->stream))) return hip_fail(e, "table_wait");
  if (!timing) return GX_OK;
  float total = 0, a = 0, b = 0;
  if ((e = hipEventElapsedTime(&total, t->ev[0], t->ev[2]))) return hip_fail(e, "table_wait: events");
  if (t->graph_ran) {
    a = t->eager_intensity_ms;
    b = total - a;
  } else {
    if ((e = hipEventElapsedTime(&a, t->ev[0], t->ev[1])) || (e = hipEventElapsedTime(&b, t->ev[1], t->ev[2])))
      return hip_fail(e, "table_wait: events");
  }
  long long evals = 0;
  for (Group& g : t->groups) {
    unsigned long long slots[kEvalSlots];
    if ((e = hipMemcpy(slots, g.evals, sizeof(slots), hipMemcpyDeviceToHost))) return hip_fail(e, "table_wait: copy");
    unsigned long long s = 0;
    for (auto v : slots) s += v;
    // mixed: one shared window set per group; FP64: every point counted its own
    evals += mixed_table(t) ? (long long)s * (long long)g.points.size() : (long long)s;
  }
  timing->intensity_ms = a;
  timing->accumulate_ms = b;
  timing->total_ms = total;
  timing->evals = evals;
  timing->lines = (long long)t->L->ls.n;
  timing->region1_ms = b;
  timing->core_ms = 0.0f;
  timing->core_evals = 0;
  timing->far_ms = 0.0f;
  timing->far_evals = 0;
  timing->interp_pairs = 0;
  if (mixed_table(t)) {
    for (Launch& l : t->launches) {
      long long ip = 0;
      const int irc = l.run.interp_pairs(&ip);
      if (irc) return irc;
      timing->interp_pairs += ip;
    }
    if (t->graph_ran) {
      timing->region1_ms = t->eager_region1_ms;
      timing->core_ms = t->eager_core_ms;
      timing->far_ms = t->eager_far_ms;
      timing->far_evals = (long long)t->eager_far_evals;
    } else {
      unsigned long long fe = 0;
      int rc = split_times(t, &timing->region1_ms, &timing->core_ms, &timing->far_ms, &fe);
      if (rc) return rc;
      timing->far_evals = (long long)fe;
    }
    if (t->core_evals) {
      unsigned long long slots[kEvalSlots], ce = 0;
      if ((e = hipMemcpy(slots, t->core_evals, sizeof(slots), hipMemcpyDeviceToHost)))
        return hip_fail(e, "table_wait: copy");
      for (auto v : slots) ce += v;
      timing->core_evals = (long long)ce;
    }
  }
  return GX_OK;
}

int gx_table_read(gx_table* t, int point, double* out, int n) {
  if (!t || !out || point < 0 || point >= (int)t->pts.size() || n <= 0 || n > t->M)
    return fail(GX_INVALID_ARG, "table_read: bad arguments");
  hipError_t e;
  if ((e = hipStreamSynchronize(t->stream))) return hip_fail(e, "table_read: sync");
  if ((e = hipMemcpy(out, t->d_out + (size_t)point * t->M, sizeof(double) * n, hipMemcpyDeviceToHost)))
    return hip_fail(e, "table_read: copy");
  return GX_OK;
}

int gx_table_shape(gx_table* t, int* bins_per_lane, int* segments) {
  if (!t) return fail(GX_INVALID_ARG, "table_shape: NULL table");
  if (bins_per_lane) *bins_per_lane = mixed_table(t) ? t->R : 1;
  if (segments) *segments = (mixed_table(t) && !t->launches.empty()) ? t->launches[0].run.plan.nseg : 1;
  return GX_OK;
}

int gx_table_destroy(gx_table* t) {
  free_table(t);
  return GX_OK;
}

}  // extern "C"
