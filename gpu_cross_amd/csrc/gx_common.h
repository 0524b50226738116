// gx_common.h -- host-side building blocks shared by the manager context
// (gx_api.cpp), the batched (T,P) table (gx_table.cpp) and the request
// server (gx_server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gx.h"
#include "gx_internal.h"

namespace gx {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// FP64 constants of one point, derived as OpenMPManager.cpp:57-67,191-192 derive them
// (GPUManager.cpp:24-47 and cross_kernal.cu:938 for the gpu convention).
PointConsts make_point_consts(int profile, const gx_consts& c, double pressure_scale);
double reference_cutoff(int profile, double pressure, double cutoff);
double max_cutoff(const PointConsts& p, double max_nu, double dfreq);

// Device copy of a line list, sorted by nu.
struct LineStore {
  double* nu = nullptr;
  double* elow = nullptr;
  double* aif = nullptr;
  int* gns = nullptr;
  double* gam = nullptr;
  double* nexp = nullptr;
  size_t n = 0;
  size_t cap = 0;
  double max_nu = 0.0;
  bool has_gamma = false;   // the last upload carried per-line gamma and n
  // staging of unsorted uploads (raw copies, sort keys and values, sort scratch); grown on
  // demand and kept, so repeated chunk uploads allocate nothing
  double* s_nu = nullptr;
  double* s_e = nullptr;
  double* s_a = nullptr;
  double* s_gam = nullptr;
  double* s_n = nullptr;
  int* s_g = nullptr;
  uint32_t* s_idx_in = nullptr;
  uint32_t* s_idx_out = nullptr;
  size_t s_cap = 0;            // lines
  bool s_voigt = false;        // s_gam / s_n allocated
  void* s_tmp = nullptr;
  size_t s_tmp_cap = 0;        // bytes
  int reserve(size_t n, bool voigt);
  int reserve_staging(size_t n, bool voigt);
  void release_staging();
  void release();
  int upload(size_t n, const double* e, const double* nu, const double* a, const int* g, const double* gam,
             const double* nexp, hipStream_t s);
};

// Per-point derived line data: FP64 arrays of the plain path (whose ib/ie
// also serve as the shared windows of the mixed path) and FP32 arrays of the
// mixed path for up to `points_cap` points.
struct Derived {
  DerivedF64 d{};
  size_t cap = 0;
  MixedLineData m{};
  size_t mcap = 0;         // lines per row
  int mpoints = 0;         // point rows (core ranges)
  int mtemps = 0;          // temperature rows (amp, 1/alpha)
  PointConsts* d_pcs = nullptr;
  int pcs_cap = 0;
  std::vector<PointConsts> h_pcs;
  int reserve(size_t n);
  int reserve_mixed(size_t n, int npoints, int ntemps);
  // upload per-point constants (waits for the stream only when they change)
  int set_points(const PointConsts* pcs, int npts, hipStream_t s);
  void release();
};

// Per-bin-block line slices.
struct Ranges {
  int2* r = nullptr;
  int cap = 0;
  int reserve(int nblocks);
  void release();
};

// Launch state of the mixed-precision kernel: task table, segment partials,
// work-queue counter.  Host copies of the task table stay alive here until the
// next run so asynchronous uploads never read a dead buffer.
struct MixedRunner {
  MixedTask* d_tasks = nullptr;
  int task_cap = 0;
  float2* partial = nullptr;
  size_t partial_cap = 0;      // float2 elements
  unsigned* counter = nullptr;
  int* starts = nullptr;       // core pass: centre-bin index of the sorted lines
  size_t starts_cap = 0;
  double* core_part = nullptr; // core pass: per-workgroup bin rows
  size_t core_cap = 0;
  int cus = 0;
  // bins per lane (1, 2, 4 or 8): blocks of 64 R bins.  The manager and the server take 1: the
  // far field then takes most lines of a block, the branch-free masked edge loop (R <= 2) and
  // the interpolated lines (R = 1) the rest (one 1e8-line cross-section: 25.3 ms at R = 1, 25.4
  // at 2); tables autotune it
  int R = 1;
  int items_per_slot = 4;      // work items per resident workgroup when splitting blocks
  int forced_nseg = 0;         // > 0: fixed segments per block
  bool persistent = true;      // false: one workgroup per work item, so the hardware interleaves
                               // workgroups of launches on other streams (shared device)
  // far field of region I (gx_far.hip); GX_FAR=0 switches it off
  bool far_on = true;
  double4* far_geom = nullptr;
  size_t far_geom_cap = 0;     // clusters
  FarVal* far_rec = nullptr;
  size_t far_rec_cap = 0;      // values
  int4* far_rng = nullptr;
  int4* edge_rng = nullptr;
  int* far_list = nullptr;
  int* far_cnt = nullptr;
  size_t far_blk_cap = 0;      // (task, block) slots
  int* far_kind = nullptr;
  int far_kind_cap = 0;        // tasks
  int* far_first = nullptr;    // first line above each bin: ie (entries [0, cap)), then ib
  int far_first_cap = 0;       // entries per array
  unsigned long long* far_evals = nullptr;
  // near leaves (real-centred points; GX_NEAR=0 switches them off)
  bool near_on = true;
  int2* near_u = nullptr;
  size_t near_u_cap = 0;       // (task, cluster) entries
  int4* near_rng = nullptr;
  size_t near_rng_cap = 0;     // (task, block) slots
  unsigned long long* core_counter = nullptr;   // optional kEvalSlots counters of core-pass evaluations
  // interpolated lines of the region-I kernel (GX_INTERP=0 switches them off)
  bool interp_on = true;
  float* interp_tab = nullptr;                  // interp_table(), uploaded by init
  unsigned long long* interp_cnt = nullptr;     // kEvalSlots counters, cleared by each launch
  hipEvent_t far_ev[2] = {nullptr, nullptr};
  bool far_timed = false;      // the last launch recorded far_ev
  std::vector<MixedTask> h_tasks;
  MixedPlan plan{};
  int init();
  // Device time of the last timed launch's far-field kernels and the line-bin pairs they
  // took from moments (call after the stream is synchronised).
  int far_time(float* ms, unsigned long long* evals);
  // Line-bin pairs of the last launch that the region-I kernel formed by interpolation
  // (64 per interpolated line and block; call after the stream is synchronised).
  int interp_pairs(long long* pairs);
  int bins_per_block() const { return mixed_bins_per_block(R); }
  // Device bytes the runner holds now (its buffers only grow until release()).
  size_t bytes() const;
  // Upload the task table, size the partials and the core-pass buffers for the
  // chunk gv of the grid and nlines sorted lines, and fix the launch shape.
  // May synchronise the stream; never call it inside a graph capture.
  int prepare(const MixedTask* tasks, int ntasks, int nblocks, const GridView& gv, int nlines, hipStream_t s);
  // Enqueue the kernels of the prepared plan (graph-capturable): the far field (moments,
  // block plans, far sums), region I, the core pass.  ev, when not NULL, gets 3 events:
  // before region I, between region I and the core pass, after the core pass; the far
  // field is then timed by the runner's own events (far_time).
  int launch(const GridView& gv, const double* nu, const int2* ranges, const int* ib, const int* ie, hipStream_t s,
             hipEvent_t* ev = nullptr);
  void release();
};

// Temperature rows of the mixed line data for points pcs[0..n): a new row wherever a point's
// temperature constants differ from the previous point's (as k_prep_mixed numbers them);
// row[p] (may be NULL) gets point p's row.  Returns the number of rows.
int mixed_temperature_rows(const PointConsts* pcs, int n, int* row);
// Task of point `point` (temperature row trow) over the line data m, summing into out.
MixedTask mixed_task(const MixedLineData& m, int point, int trow, const PointConsts& pc, const double* gamma,
                     double* out);

// HBM of the far-field buffers of a mixed launch (ntasks points, nlines lines, nblocks blocks).
size_t far_field_bytes(int ntasks, long long nlines, int nblocks);

// Core pass of the mixed path: bound on |bin - centre bin| inside any W4 core of the
// chunk gv at Doppler width coefficient dpw, or -1 when its bin row does not fit LDS
// (such grids run the FP64 path).
double chunk_max_abs_nu(const GridView& gv);
int core_kmax(const GridView& gv, double dpw);
bool core_pass_fits(const GridView& gv, double dpw);
// Row geometry of the core pass for the chunk gv and a core bound kmax (core_kmax).
struct CoreGeom {
  int cpw;     // centre bins per workgroup
  int span;    // bins of one workgroup's row: cpw + 2 kmax + 1
  int cbase;   // first centre bin that can reach the chunk
  int ncwg;    // workgroups (rows) per task
  int ncent;   // centre bins covered: ncwg * cpw
};
CoreGeom core_geometry(const GridView& gv, int kmax);
// Bytes of one task's core-pass rows (ncwg * span doubles) at Doppler width coefficient dpw.
size_t core_rows_bytes(const GridView& gv, double dpw);

// Launch line prep, block ranges and accumulation of one point (asynchronous).
// ev, when not NULL, holds 7 events: 4 around prep / ranges / accumulation and,
// for the mixed path, 3 around region I and the core pass.  d_core_evals (may be
// NULL) counts the core-pass evaluations of the mixed path.
int run_point(const PointConsts& pc, const GridView& gv, const LineStore& ls, int nlines, int precision,
              Derived& der, Ranges& rng, MixedRunner& mix, double* intens, unsigned long long* d_evals,
              unsigned long long* d_core_evals, hipStream_t s, hipEvent_t* ev);

}  // namespace gx

// Resident line list (gx_lines_* in include/gx.h), shared by tables and the server.
struct gx_lines {
  int device = 0;
  hipStream_t stream = nullptr;
  gx::LineStore ls;
  bool voigt = false;
  std::vector<double> h_nu;   // host copy of the sorted centres, for slicing by frequency
};
