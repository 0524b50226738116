// gx_internal.h -- declarations shared by the kernels (gx_kernels.hip) and the
// host side of the library (gx_api.cpp, gx_table.cpp, gx_server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gx {

// LDS written by the lanes of one wave and read back by the same wave: LDS
// instructions of a wave complete in order; this keeps the compiler from
// moving accesses across the hand-over.
__device__ __forceinline__ void wave_lds_handover() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Constants of one (T, P) point, derived on the host in FP64 exactly as the
// reference's managers derive them (OpenMPManager.cpp:57-67,191-192).
struct PointConsts {
  double beta;            // c2 / T
  double cmcoef;          // 1 / (8 pi c)
  double partition;       // Q(T)
  double dpwcoef;         // Doppler width per unit nu
  double temperature;     // K
  double ref_temp;        // K
  double pressure_scale;  // multiplies gamma (1 when gamma already holds gamma*P)
  double cutoff;          // window half-width, cm^-1
  double doppler_widths;  // > 0: per-line Doppler window k*alpha + dfreq
  int profile;            // kProfGaussian, kProfDoppler, kProfVoigt or kProfVoigtNoGamma
  int convention;         // 0 openmp, 1 gpu
};

// Profiles of PointConsts: the reference's ProfileType numbering (Input.h:9-15) for the first
// three, and the Voigt sum of a line list uploaded without per-line Lorentz widths (the CUDA
// manager's cross_kernal.cu:860-922), which has no number of its own there.  For the Gaussian,
// dpwcoef holds x_ = sqrt(ln2)/HW; for the no-gamma Voigt, the Doppler width per unit nu as
// cross_kernal.cu:872 forms it.
constexpr int kProfGaussian = 0;
constexpr int kProfDoppler = 1;
constexpr int kProfVoigt = 2;
constexpr int kProfVoigtNoGamma = 4;

// Grid and chunk geometry.
struct GridView {
  const double* freq;  // all M nodes
  int npoints_total;   // M
  int start_idx;       // first bin of the chunk
  int points;          // bins in the chunk
  double start_nu;     // freq[0]
  double dfreq;
};

// Centre bin of a line on the grid (nearest node), clamped far inside int range.
__device__ __forceinline__ int centre_bin(double nu_if, const GridView& gv) {
#pragma clang fp contract(off)
  const double c = rint((nu_if - gv.start_nu) / gv.dfreq);
  return (int)fmin(fmax(c, -1073741824.0), 1073741824.0);
}

// Resident line list, sorted by nu (SoA, FP64 as the reference's BaseManager takes it).
struct LinesView {
  const double* nu;
  const double* elow;
  const double* aif;
  const int* gns;
  const double* gamma;  // NULL for Doppler
  const double* nexp;   // NULL for Doppler
  int n;
};

// Per-line derived data of one point (written by the intensity kernel).
struct DerivedF64 {
  double* S;     // intensity (x ISQRTPI for Voigt)
  double* g;     // 1 / (nu * dpwcoef)
  double* y;     // gamma_T * g (Voigt)
  int* ib;       // window [ib, ie) clamped to the chunk
  int* ie;
};

// Line data of the mixed-precision path.  What depends on the temperature only is kept
// once per run of consecutive points at one temperature, [ntemps][stride]: S/alpha, 1/alpha
// and the Lorentz temperature factor (T_ref/T)^n; per point, [npoints][stride], only the W4
// core range, an integer bin range found with the FP64 path's own test
// |nu - f_j|/alpha + y < 15.  y = ((gamma P) (T_ref/T)^n) / alpha is formed where it is used
// (mixed_y), from the resident gamma and these rows, in the FP64 path's order of
// operations, so y and 1/alpha are that path's bits and every W4 region decision is that
// path's.  8 B per line and point plus 20 B per line and temperature (11.3 B per line and
// point over 8 temperatures and 48 points).
constexpr size_t kMixedPointBytes = 8;    // per line and point: the core range
constexpr size_t kMixedTempBytes = 20;    // per line and temperature: S/alpha, 1/alpha, (T_ref/T)^n
struct MixedLineData {
  float* amp;      // [ntemps][stride] S / alpha (FP32)
  double* g;       // [ntemps][stride] 1 / alpha = 1 / (nu * dpwcoef)
  double* ts;      // [ntemps][stride] (T_ref / T)^n
  int2* core;      // [npoints][stride] [k0, k1): bins of the window where |x| + y < 15 (regions II-IV)
  size_t stride;   // lines per row
};

// Window rule shared by all points of a launch.
struct WindowArgs {
  double cutoff;
  int convention;     // 0 round() rule, 1 per-bin |nu - f| <= C
  int zero_to_tiny;   // nu == 0 -> 1e-6 (OpenMPManager.cpp:198)
};

// One (T,P) point of a mixed-precision launch: its line data and its grid.
struct MixedTask {
  const float* amp;      // its temperature's rows of MixedLineData
  const double* g;
  const double* ts;
  const double* gamma;   // the resident line list's gamma (indexed as nu)
  const int2* core;      // its own row
  double pscale;         // the point's pressure_scale
  double dpwcoef;        // Doppler width per unit nu of this point
  double* out;
};

// A load through the global address space: pointers read from memory (a task table) are
// generic to the compiler, and its flat loads also hold the LDS wait counter.
template <class T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ int2 ldg(const int2* p) {
  const unsigned long long v = *(const __attribute__((address_space(1))) unsigned long long*)p;
  return make_int2((int)(unsigned)v, (int)(unsigned)(v >> 32));
}

// y = gamma_T / alpha of line l of a mixed task: ((gamma pressure_scale) (T_ref/T)^n) g, each
// product rounded as the FP64 path rounds it (k_line_prep), so the bits are that path's.
__device__ __forceinline__ double mixed_y(const MixedTask& tk, int l, double g) {
#pragma clang fp contract(off)
  return ((ldg(tk.gamma + l) * tk.pscale) * ldg(tk.ts + l)) * g;
}

// Far field of region I: moments per cluster, leaf size, list length, levels.
constexpr int kFarP = 24;            // moments per cluster
constexpr double kFarEta = 0.5;      // a cluster of radius r serves bins at distance >= r / eta
constexpr int kFarLeaf = 64;         // lines per leaf
constexpr int kFarHead = 6;          // radius, max 1/g, min y, lowest and highest gamma, gamma of the centre
constexpr int kFarRec = kFarHead + 2 * kFarP;   // head, then P real or P complex moments
// Cluster records are stored in FP32: the bound fields (radius, largest Doppler width, lowest y,
// Lorentz-width range) rounded outward, the complex centre's gamma and the moments to nearest;
// every kernel computes in FP64 from them (the tree shifts, the series, the plans).
using FarVal = float;
constexpr double kFarTol = 5e-7;     // bound on the relative truncation error of every far-field sum
constexpr double kFarCplxGamma = 0.05;  // points whose sampled Lorentz widths are all >= this (cm^-1)
                                        // take complex-centred moments
constexpr int kFarList = 128;        // clusters per block at most
constexpr int kFarMaxLevels = 32;
// Near leaves (real-centred points): a leaf whose lines' W4 cores and series-convergence zone
// fit in one bin range U (its near range) has its lines evaluated by the core pass at every bin
// of U (region I outside their cores) and is summed from its moments at the other bins of the
// blocks whose near lines it holds.  U must lie within kmax of each line's centre bin; the core
// pass rows get kNearMargin bins of room for the part of U beyond the cores.
constexpr int kNearMargin = 8;
constexpr double kNearRho = 0.42;    // leaf series used where rho = r / |f - nu_c| <= this

// A mixed-precision launch over ntasks points sharing one grid, one sorted
// line list and one set of block ranges.  Work items are (task, block, segment).
struct MixedPlan {
  GridView gv;
  const double* nu;        // sorted line centres
  const int2* ranges;      // per-block line slices
  const int* ib;           // shared line windows
  const int* ie;
  int nblocks;             // bin blocks of mixed_bins_per_block() bins
  int nseg;                // line segments per block (partials reduced in order when > 1)
  int ntasks;
  const MixedTask* tasks;  // device array [ntasks]
  float2* partial;         // [nitems][bins per block] when nseg > 1
  unsigned* counter;       // work-queue head, reset by the launch
  int nitems;              // ntasks * nblocks * nseg
  int grid;                // workgroups launched (kBlock / 64 waves each)
  int queue;               // 1: waves take items from the counter until none is left; 0: one item per wave
  // W4 core pass (regions II-IV): lines grouped by centre bin c = rint((nu - nu0)/dnu)
  int nlines;              // lines in nu[]
  int* starts;             // [ncent + 1]: first line whose centre bin is >= cbase + c
  int cbase;               // first centre bin that can reach the chunk
  int ncent;               // centre bins [cbase, cbase + ncent)
  int kmax;                // bound on |bin - centre bin| inside any core
  int cpw;                 // centre bins per core workgroup
  int span;                // bins of one core workgroup's row: cpw + 2 kmax + 1
  int ncwg;                // core workgroups per task
  double* core_part;       // [ntasks][ncwg][span] per-workgroup core sums
  // Far field of region I (gx_far.hip): multipole moments of 64-line leaves and of their
  // binary tree; blocks take the lines far from them from the moments instead of line
  // by line.  far == 0: every line of a block is evaluated directly.
  int far;
  int far_bins;            // bins of one block (64 R)
  int far_nlev;            // tree levels (level 0: leaves)
  int far_off[kFarMaxLevels];   // first cluster of each level
  int far_n[kFarMaxLevels];     // clusters of each level
  int far_nclus;           // clusters of all levels
  double4* far_geom;       // [far_nclus]: centre, lowest nu, highest nu (shared by the tasks)
  FarVal* far_rec;         // [ntasks][kFarRec][far_nclus]: head and moments of each cluster, field-major
  int* far_kind;           // [ntasks]: 0 real-centred moments, 1 complex-centred
  int4* far_rng;           // [ntasks][nblocks]: far line ranges [x, y) and [z, w) of each block
  int4* edge_rng;          // [ntasks][nblocks]: window-end lines from moments: [x, y) end inside the
                           // block (bins j < ie), [z, w) start inside it (bins j >= ib)
  int* far_ie1;            // [gv.points + 1]: first line with ie > gv.start_idx - 1 + t (ie, ib are
  int* far_ib1;            // non-decreasing over the sorted lines); far_ib1 likewise for ib
  int far_bin_cap;         // entries of far_ie1 / far_ib1
  int* far_list;           // [ntasks][nblocks][kFarList]: clusters of each block, level << 24 | index
  int* far_cnt;            // [ntasks][nblocks]
  unsigned long long* far_evals;   // line-bin pairs taken from moments
  // near leaves (see kNearMargin): near == 0 leaves every line's exclusion range at its core
  int near;
  int2* near_u;            // [ntasks][far_nclus]: near range U of a leaf (empty when the leaf is not
                           // near-eligible), the hull of its leaves' ranges for a tree cluster
  int4* near_rng;          // [ntasks][nblocks]: runs [x, y) and [z, w) of near leaves (leaf indices)
                           // summed from moments per bin by k_near_eval
  unsigned long long* core_evals;  // kEvalSlots counters of core-pass evaluations (may be NULL):
                                   // the core pass adds the bins of U beyond each line's core
  // interpolated lines of the region-I kernel (64-bin blocks): NULL switches them off
  const float* interp_tab;         // [64][kInterpNodes] barycentric matrix, then the node offsets
  unsigned long long* interp_cnt;  // kEvalSlots counters of interpolated (line, block) pairs (may be NULL)
};

// Chebyshev-Lobatto nodes of a 64-bin block for the interpolated lines of the region-I kernel.
constexpr int kInterpNodes = 16;
constexpr int kInterpTabNodes = 64 * kInterpNodes;   // offset of the node offsets in the table
// The table: A[j][k] = L_k(s_j) with s_j = (j - 31.5) / 31.5 and the Lobatto nodes
// s_k = cos(pi k / 15), then for each node -(its bin position - 32) (the anchored form's
// offset sign); 64 x 16 + 16 floats.
void interp_table(float* tab);

// One request of a batched FP64 launch (launch_f64_batch): its constants, grid, line slice
// [line_lo, line_lo + nlines) of the resident list, derived data (slice-local index) and sums.
struct F64Task {
  PointConsts pc;
  GridView gv;
  const double* nu;                // the whole sorted list
  int line_lo;
  int nlines;
  DerivedF64 d;
  double* out;
  long long prep_first;            // first prep thread (a multiple of 256)
  int first_block;                 // first of its bin blocks among all tasks' blocks
  double zero_nu, x0;              // f64_zero_nu, f64_x0
  unsigned long long* evals;       // kEvalSlots counters of its windows
};

// A chunk of work: lines [a, z) (absolute indices) over bin block `block` of task `task`.
struct F64Item {
  int task, block, a, z;
};

constexpr int kEvalSlots = 64;       // eval counters (spread to avoid one hot atomic)
constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kTile = 256;           // lines staged in LDS per pass (FP64 kernels)
constexpr int kWavesPerGroup = kBlock / 64;

// ---- launchers (gx_kernels.hip) ----
hipError_t launch_iota(uint32_t* idx, int n, hipStream_t s);
hipError_t sort_pairs_by_nu(void* temp, size_t& temp_bytes, const double* keys_in, double* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, int n, hipStream_t s);
hipError_t launch_gather_lines(const uint32_t* perm, const double* elow_in, const double* aif_in, const int* gns_in,
                               const double* gam_in, const double* nexp_in, double* elow, double* aif, int* gns,
                               double* gam, double* nexp, int n, hipStream_t s);
hipError_t launch_line_prep_f64(const PointConsts& pc, const GridView& gv, const LinesView& lv, DerivedF64 out,
                                unsigned long long* evals, hipStream_t s);
hipError_t launch_block_ranges(const GridView& gv, const double* nu, int n, double cut_max, int bins_per_block,
                               int nblocks, int2* ranges, hipStream_t s);
hipError_t launch_accum_f64(const PointConsts& pc, const GridView& gv, const double* nu_sorted, DerivedF64 d,
                            const int2* ranges, int nblocks, double* intens, hipStream_t s);
hipError_t launch_windows(const WindowArgs& wa, const GridView& gv, const double* nu, int n, int* ib, int* ie,
                          unsigned long long* evals, hipStream_t s);
// core_evals (kEvalSlots counters, may be NULL) accumulates the bins of all W4 cores
hipError_t launch_prep_mixed(const PointConsts* d_pcs, int npts, const LinesView& lv, const GridView& gv,
                             const int* ib, const int* ie, MixedLineData out, unsigned long long* core_evals,
                             hipStream_t s);
hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s);
// Batched FP64 path: prep of every task's slice (prep_threads threads), one workgroup per
// item writing partial[item][kBinsF64], then out[j] += the partials of j's block in item
// order (block_items[global block] = (first item, count)).
hipError_t launch_f64_batch(const F64Task* d_tasks, int ntasks, long long prep_threads, const LinesView& all,
                            const F64Item* d_items, int nitems, int nblocks_all, const int2* d_block_items,
                            double* partial, hipStream_t s);
double f64_zero_nu(const PointConsts& pc);   // nu == 0 stand-in of the FP64 sums
double f64_x0(const PointConsts& pc, double dfreq);   // half-bin term of the erf-binned profiles
hipError_t launch_grid(double* f, int n, double start, double dfreq, hipStream_t s);
// R = bins per lane (4 or 8); a work item (one wave) covers mixed_bins_per_block(R) bins.
// launch_accum_mixed: region I (and the segment reduction); launch_core_mixed: the W4 core pass.
hipError_t launch_accum_mixed(const MixedPlan& plan, int R, hipStream_t s);
// Far field (gx_far.hip): cluster geometry, moments and block plans (before region I),
// and the far-field sums added to the grids.
hipError_t launch_far_setup(const MixedPlan& plan, hipStream_t s);
hipError_t launch_far_eval(const MixedPlan& plan, int R, hipStream_t s);
hipError_t launch_core_mixed(const MixedPlan& plan, hipStream_t s);
int mixed_occupancy(int R);   // resident workgroups (kWavesPerGroup waves each) per CU of the mixed kernel
// LDS bytes of a core-pass workgroup with the given row span (0: the row does not fit LDS)
size_t core_lds_bytes(int span);
int mixed_bins_per_block(int R);

// bins covered by one workgroup of the FP64 accumulation kernels
constexpr int kBinsF64 = 256;

}  // namespace gx
