// gx_api.cpp -- host side of libgx.so: the single-GPU manager context behind
// the C ABI of include/gx.h.
//
// One gx_ctx replaces one reference GpuManager (GPUManager.cpp:5-189) with the
// BaseManager lifecycle (BaseManager.h:22-30):
//   create -> initialize_vectors(M) -> transfer_freq -> initialize_constants
//   -> {transfer_vectors(n); execute(points, n, ib)}* -> transfer_results -> cleanup.
// Differences that are deliberate: the uploaded line list is never mutated
// (the reference's stepone overwrites E'' and gamma in place,
// cross_kernal.cu:780,783), per-point derived data lives in separate buffers,
// chunks are sorted by nu on the device when the caller's order is not, and
// errors come back as status codes instead of exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/gx.h"
#include "gx_internal.h"
#include "gx_common.h"
#include "gx_noalloc.h"

namespace gx {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return fail(GX_OUT_OF_MEMORY, m);
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return fail(GX_NO_DEVICE, m);
  return fail(GX_DEVICE_ERROR, m);
}

// ---- constants, exactly as the reference's managers derive them ----
// defines.h:4-20 literals
constexpr double kPlanck = 6.6260693e-27, kVellgt = 2.99792458e+10, kBoltz = 1.380658e-16;
constexpr double kNA = 6.0221412927e23, kLn2 = 0.69314718056, kPi = 3.14159265359;

PointConsts make_point_consts(int profile, const gx_consts& c, double pressure_scale) {
  PointConsts p{};
  if (profile == kProfGaussian || profile == kProfVoigtNoGamma) {
    // sums the reference has only in its CUDA manager: its constants and per-bin windows,
    // whatever the convention (GPUManager.cpp:24-47, cross_kernal.cu:471-625, 731-922)
    p.profile = profile;
    p.convention = GX_CONV_GPU;
    p.temperature = c.temperature;
    p.partition = c.partition;
    p.ref_temp = c.ref_temp;
    p.beta = kPlanck * kVellgt / (kBoltz * c.temperature);
    p.cmcoef = 1.0 / (8.0 * std::acos(-1.0) * 2.99792458e10);
    if (profile == kProfGaussian) {
      // x_ = sqrt(ln2)/HW with ln2 the defines.h:4 literal: the reference's cross_constants.ln2
      // is never set (GPUManager.cpp:24-47), so its value there is undefined
      p.dpwcoef = std::sqrt(kLn2) / c.half_width;
      p.cutoff = c.cutoff > 0.0 ? c.cutoff : c.half_width * 10.0;          // cross_kernal.cu:578
    } else {
      // cross_kernal.cu:872 as written: no Avogadro factor, so with the mean mass in amu the
      // Doppler width is ~1e-16 cm^-1 (every bin then takes the 3-point form)
      p.dpwcoef = std::sqrt(2.0 * kLn2 * kBoltz * c.temperature / (c.mean_mass)) / kVellgt;
      p.cutoff = c.cutoff > 0.0 ? c.cutoff : 500.0 * 0.05;                 // :885-888
    }
    p.pressure_scale = pressure_scale;
    return p;
  }
  p.profile = profile;
  p.convention = c.convention;
  p.temperature = c.temperature;
  p.partition = c.partition;
  p.ref_temp = c.ref_temp;
  p.beta = kPlanck * kVellgt / (kBoltz * c.temperature);               // OpenMPManager.cpp:66
  if (c.convention == GX_CONV_GPU) {
    p.cmcoef = 1.0 / (8.0 * std::acos(-1.0) * 2.99792458e10);           // GPUManager.cpp:31,39
    p.dpwcoef = std::sqrt(2.0 * kBoltz * c.temperature * kNA / (c.mean_mass)) / kVellgt;   // cross_kernal.cu:938
  } else {
    p.cmcoef = 1.0 / (8.0 * kPi * kVellgt);                             // OpenMPManager.cpp:65
    p.dpwcoef = std::sqrt(2.0 * kLn2 * kBoltz * c.temperature * kNA / (c.mean_mass)) / kVellgt;  // :191
  }
  p.cutoff = reference_cutoff(profile, c.pressure, c.cutoff);
  p.doppler_widths = profile == GX_DOPPLER ? c.doppler_widths : 0.0;
  p.pressure_scale = pressure_scale;
  return p;
}

double reference_cutoff(int profile, double pressure, double cutoff) {
  if (cutoff > 0.0) return cutoff;
  if (profile == GX_DOPPLER) return 100.0;     // defines.h:20, DopplerProfile.cpp:4
  return std::min(25.0 * pressure, 100.0);     // OpenMPManager.cpp:192, VoigtProfile.cpp:49
}

double max_cutoff(const PointConsts& p, double max_nu, double dfreq) {
  if (p.profile == GX_DOPPLER && p.doppler_widths > 0.0)
    return p.doppler_widths * (std::max(max_nu, 1e-6) * p.dpwcoef) * (1.0 + 1e-12) + dfreq;
  return p.cutoff;
}

// ---- device line storage shared by the manager, the table and the server ----
int LineStore::reserve(size_t n, bool voigt) {
  if (n <= cap && (!voigt || gam)) return GX_OK;
  release();
  const size_t c = std::max<size_t>(n, 1);
  hipError_t e;
  if ((e = hipMalloc(&nu, c * 8)) != hipSuccess) return hip_fail(e, "hipMalloc(nu)");
  if ((e = hipMalloc(&elow, c * 8)) != hipSuccess) return hip_fail(e, "hipMalloc(elow)");
  if ((e = hipMalloc(&aif, c * 8)) != hipSuccess) return hip_fail(e, "hipMalloc(aif)");
  if ((e = hipMalloc(&gns, c * 4)) != hipSuccess) return hip_fail(e, "hipMalloc(gns)");
  if (voigt) {
    if ((e = hipMalloc(&gam, c * 8)) != hipSuccess) return hip_fail(e, "hipMalloc(gamma)");
    if ((e = hipMalloc(&nexp, c * 8)) != hipSuccess) return hip_fail(e, "hipMalloc(nexp)");
  }
  cap = c;
  return GX_OK;
}

void LineStore::release() {
  hipFree(nu); hipFree(elow); hipFree(aif); hipFree(gns); hipFree(gam); hipFree(nexp);
  nu = elow = aif = gam = nexp = nullptr;
  gns = nullptr;
  cap = 0;
  n = 0;
  release_staging();
}

void LineStore::release_staging() {
  hipFree(s_nu); hipFree(s_e); hipFree(s_a); hipFree(s_g); hipFree(s_gam); hipFree(s_n);
  hipFree(s_idx_in); hipFree(s_idx_out); hipFree(s_tmp);
  s_nu = s_e = s_a = s_gam = s_n = nullptr;
  s_g = nullptr;
  s_idx_in = s_idx_out = nullptr;
  s_tmp = nullptr;
  s_cap = 0;
  s_tmp_cap = 0;
  s_voigt = false;
}

int LineStore::reserve_staging(size_t nl, bool voigt) {
  if (nl <= s_cap && (!voigt || s_voigt)) return GX_OK;
  hipFree(s_nu); hipFree(s_e); hipFree(s_a); hipFree(s_g); hipFree(s_gam); hipFree(s_n);
  hipFree(s_idx_in); hipFree(s_idx_out);
  s_nu = s_e = s_a = s_gam = s_n = nullptr;
  s_g = nullptr;
  s_idx_in = s_idx_out = nullptr;
  s_cap = 0;
  s_voigt = false;
  const size_t c = std::max(nl, s_cap);
  hipError_t e;
  if ((e = hipMalloc(&s_nu, c * 8)) || (e = hipMalloc(&s_e, c * 8)) || (e = hipMalloc(&s_a, c * 8)) ||
      (e = hipMalloc(&s_g, c * 4)) || (e = hipMalloc(&s_idx_in, c * 4)) || (e = hipMalloc(&s_idx_out, c * 4)))
    return hip_fail(e, "transfer_vectors: staging alloc");
  if (voigt && ((e = hipMalloc(&s_gam, c * 8)) || (e = hipMalloc(&s_n, c * 8))))
    return hip_fail(e, "transfer_vectors: staging alloc");
  s_cap = c;
  s_voigt = voigt;
  return GX_OK;
}

// Upload n lines (host SoA) and leave them sorted by nu on the device.
int LineStore::upload(size_t nlines, const double* h_e, const double* h_nu, const double* h_a, const int* h_g,
                      const double* h_gam, const double* h_n, hipStream_t s) {
  const bool voigt = h_gam != nullptr;
  int rc = reserve(nlines, voigt);
  if (rc) return rc;
  has_gamma = voigt;
  bool sorted = true;
  double mx = 0.0;
  for (size_t i = 0; i < nlines; ++i) {
    if (!(h_nu[i] == h_nu[i])) return fail(GX_INVALID_ARG, "transfer_vectors: nu is NaN");
    if (i && h_nu[i] < h_nu[i - 1]) sorted = false;
    mx = std::max(mx, h_nu[i]);
  }
  max_nu = mx;
  n = nlines;
  hipError_t e;
  if (nlines == 0) return GX_OK;
  if (sorted) {
    if ((e = hipMemcpyAsync(nu, h_nu, nlines * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(elow, h_e, nlines * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(aif, h_a, nlines * 8, hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(gns, h_g, nlines * 4, hipMemcpyHostToDevice, s)))
      return hip_fail(e, "transfer_vectors: copy");
    if (voigt && ((e = hipMemcpyAsync(gam, h_gam, nlines * 8, hipMemcpyHostToDevice, s)) ||
                  (e = hipMemcpyAsync(nexp, h_n, nlines * 8, hipMemcpyHostToDevice, s))))
      return hip_fail(e, "transfer_vectors: copy");
    if ((e = hipStreamSynchronize(s))) return hip_fail(e, "transfer_vectors: sync");
    return GX_OK;
  }
  // Unsorted chunk: stage raw arrays, radix-sort (nu, index) pairs, gather the rest.  The
  // staging buffers are kept between uploads (grown only when a chunk is larger).
  const int ni = (int)nlines;
  if ((rc = reserve_staging(nlines, voigt))) return rc;
  if ((e = hipMemcpyAsync(s_nu, h_nu, nlines * 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(s_e, h_e, nlines * 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(s_a, h_a, nlines * 8, hipMemcpyHostToDevice, s)) ||
      (e = hipMemcpyAsync(s_g, h_g, nlines * 4, hipMemcpyHostToDevice, s)))
    return hip_fail(e, "transfer_vectors: copy");
  if (voigt && ((e = hipMemcpyAsync(s_gam, h_gam, nlines * 8, hipMemcpyHostToDevice, s)) ||
                (e = hipMemcpyAsync(s_n, h_n, nlines * 8, hipMemcpyHostToDevice, s))))
    return hip_fail(e, "transfer_vectors: copy");
  if ((e = launch_iota(s_idx_in, ni, s))) return hip_fail(e, "transfer_vectors: iota");
  size_t tmp_bytes = 0;
  if ((e = sort_pairs_by_nu(nullptr, tmp_bytes, s_nu, nu, s_idx_in, s_idx_out, ni, s)))
    return hip_fail(e, "transfer_vectors: sort size");
  if (tmp_bytes > s_tmp_cap) {
    if ((e = hipStreamSynchronize(s))) return hip_fail(e, "transfer_vectors: sync");
    hipFree(s_tmp);
    s_tmp = nullptr;
    s_tmp_cap = 0;
    if ((e = hipMalloc(&s_tmp, std::max<size_t>(tmp_bytes, 16)))) return hip_fail(e, "transfer_vectors: sort alloc");
    s_tmp_cap = std::max<size_t>(tmp_bytes, 16);
  }
  if ((e = sort_pairs_by_nu(s_tmp, tmp_bytes, s_nu, nu, s_idx_in, s_idx_out, ni, s)))
    return hip_fail(e, "transfer_vectors: sort");
  if ((e = launch_gather_lines(s_idx_out, s_e, s_a, s_g, voigt ? s_gam : nullptr, voigt ? s_n : nullptr, elow, aif,
                               gns, gam, nexp, ni, s)))
    return hip_fail(e, "transfer_vectors: gather");
  if ((e = hipStreamSynchronize(s))) return hip_fail(e, "transfer_vectors: sync");
  return GX_OK;
}

int Derived::reserve(size_t n) {
  if (n <= cap) return GX_OK;
  hipFree(d.S); hipFree(d.g); hipFree(d.y); hipFree(d.ib); hipFree(d.ie);
  d = DerivedF64{};
  cap = 0;
  const size_t c = std::max<size_t>(n, 1);
  hipError_t e;
  if ((e = hipMalloc(&d.S, c * 8)) || (e = hipMalloc(&d.g, c * 8)) || (e = hipMalloc(&d.y, c * 8)) ||
      (e = hipMalloc(&d.ib, c * 4)) || (e = hipMalloc(&d.ie, c * 4)))
    return hip_fail(e, "hipMalloc(derived)");
  cap = c;
  return GX_OK;
}

int Derived::reserve_mixed(size_t n, int npoints, int ntemps) {
  if (n <= mcap && npoints <= mpoints && ntemps <= mtemps) return GX_OK;
  hipFree(m.amp); hipFree(m.g); hipFree(m.ts); hipFree(m.core);
  m = MixedLineData{};
  const size_t c = std::max<size_t>({n, mcap, 1});
  const int np = std::max({npoints, mpoints, 1}), nt = std::max({ntemps, mtemps, 1});
  mcap = 0;
  mpoints = 0;
  mtemps = 0;
  const size_t rows_t = c * (size_t)nt, rows_p = c * (size_t)np;
  hipError_t e;
  if ((e = hipMalloc(&m.amp, rows_t * 4)) || (e = hipMalloc(&m.g, rows_t * 8)) || (e = hipMalloc(&m.ts, rows_t * 8)) ||
      (e = hipMalloc(&m.core, rows_p * sizeof(int2))))
    return hip_fail(e, "hipMalloc(mixed line data)");
  m.stride = c;
  mcap = c;
  mpoints = np;
  mtemps = nt;
  return GX_OK;
}

int mixed_temperature_rows(const PointConsts* pcs, int n, int* row) {
  // the same fields k_prep_mixed's same_t compares
  int r = 0;
  for (int p = 0; p < n; ++p) {
    if (p > 0) {
      const PointConsts& a = pcs[p], & b = pcs[p - 1];
      const bool same = a.beta == b.beta && a.cmcoef == b.cmcoef && a.partition == b.partition &&
                        a.dpwcoef == b.dpwcoef && a.ref_temp == b.ref_temp && a.temperature == b.temperature &&
                        a.convention == b.convention;
      if (!same) ++r;
    }
    if (row) row[p] = r;
  }
  return n > 0 ? r + 1 : 0;
}

MixedTask mixed_task(const MixedLineData& m, int point, int trow, const PointConsts& pc, const double* gamma,
                     double* out) {
  const size_t ot = (size_t)trow * m.stride, op = (size_t)point * m.stride;
  return MixedTask{m.amp + ot, m.g + ot, m.ts + ot, gamma, m.core + op, pc.pressure_scale, pc.dpwcoef, out};
}

int Derived::set_points(const PointConsts* pcs, int npts, hipStream_t s) {
  hipError_t e;
  const bool same = h_pcs.size() == (size_t)npts && std::memcmp(h_pcs.data(), pcs, sizeof(PointConsts) * npts) == 0;
  if (same) return GX_OK;
  if ((e = hipStreamSynchronize(s))) return hip_fail(e, "points: sync");
  if (npts > pcs_cap) {
    hipFree(d_pcs);
    d_pcs = nullptr;
    if ((e = hipMalloc(&d_pcs, sizeof(PointConsts) * (size_t)npts))) return hip_fail(e, "hipMalloc(points)");
    pcs_cap = npts;
  }
  h_pcs.assign(pcs, pcs + npts);
  if ((e = hipMemcpy(d_pcs, h_pcs.data(), sizeof(PointConsts) * (size_t)npts, hipMemcpyHostToDevice)))
    return hip_fail(e, "points: upload");
  return GX_OK;
}

void Derived::release() {
  hipFree(d.S); hipFree(d.g); hipFree(d.y); hipFree(d.ib); hipFree(d.ie);
  d = DerivedF64{};
  cap = 0;
  hipFree(m.amp); hipFree(m.g); hipFree(m.ts); hipFree(m.core);
  m = MixedLineData{};
  mcap = 0;
  mpoints = 0;
  mtemps = 0;
  hipFree(d_pcs);
  d_pcs = nullptr;
  pcs_cap = 0;
  h_pcs.clear();
}

int Ranges::reserve(int nb) {
  if (nb <= cap) return GX_OK;
  hipFree(r);
  r = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(&r, sizeof(int2) * (size_t)std::max(nb, 1));
  if (e) return hip_fail(e, "hipMalloc(ranges)");
  cap = std::max(nb, 1);
  return GX_OK;
}

void Ranges::release() {
  hipFree(r);
  r = nullptr;
  cap = 0;
}

int MixedRunner::init() {
  if (counter) return GX_OK;
  // GX_MIXED_R (1, 2, 4 or 8) overrides the default bins per lane; used to compare launch shapes
  if (const char* env = std::getenv("GX_MIXED_R")) {
    const int r = std::atoi(env);
    if (r == 1 || r == 2 || r == 4 || r == 8) R = r;
  }
  if (const char* env = std::getenv("GX_FAR")) far_on = std::atoi(env) != 0;
  if (const char* env = std::getenv("GX_NEAR")) near_on = std::atoi(env) != 0;
  if (const char* env = std::getenv("GX_INTERP")) interp_on = std::atoi(env) != 0;
  hipError_t e = hipMalloc(&counter, sizeof(unsigned));
  if (e) return hip_fail(e, "hipMalloc(counter)");
  if ((e = hipMalloc(&far_evals, sizeof(unsigned long long))) || (e = hipEventCreate(&far_ev[0])) ||
      (e = hipEventCreate(&far_ev[1])))
    return hip_fail(e, "far-field counters");
  {
    std::vector<float> tab(kInterpTabNodes + kInterpNodes);
    interp_table(tab.data());
    if ((e = hipMalloc(&interp_tab, sizeof(float) * tab.size())) ||
        (e = hipMalloc(&interp_cnt, sizeof(unsigned long long) * kEvalSlots)) ||
        (e = hipMemcpy(interp_tab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice)) ||
        (e = hipMemset(interp_cnt, 0, sizeof(unsigned long long) * kEvalSlots)))
      return hip_fail(e, "interpolation table");
  }
  int dev = 0;
  if ((e = hipGetDevice(&dev)) || (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)))
    return hip_fail(e, "device attributes");
  cus = std::max(cus, 1);
  return GX_OK;
}

void MixedRunner::release() {
  hipFree(d_tasks); hipFree(partial); hipFree(counter); hipFree(starts); hipFree(core_part);
  d_tasks = nullptr; partial = nullptr; counter = nullptr; starts = nullptr; core_part = nullptr;
  task_cap = 0; partial_cap = 0; cus = 0; starts_cap = 0; core_cap = 0;
  hipFree(far_geom); hipFree(far_rec); hipFree(far_rng); hipFree(edge_rng); hipFree(far_list); hipFree(far_cnt);
  hipFree(far_evals); hipFree(far_kind); hipFree(far_first);
  far_geom = nullptr; far_rec = nullptr; far_rng = nullptr; edge_rng = nullptr; far_list = nullptr; far_cnt = nullptr;
  far_evals = nullptr; far_kind = nullptr; far_geom_cap = far_rec_cap = far_blk_cap = 0; far_kind_cap = 0;
  far_first = nullptr; far_first_cap = 0;
  hipFree(near_u); hipFree(near_rng);
  near_u = nullptr; near_rng = nullptr; near_u_cap = near_rng_cap = 0;
  hipFree(interp_tab); hipFree(interp_cnt);
  interp_tab = nullptr; interp_cnt = nullptr;
  for (auto& ev : far_ev) {
    if (ev) hipEventDestroy(ev);
    ev = nullptr;
  }
  far_timed = false;
  h_tasks.clear();
}

size_t MixedRunner::bytes() const {
  return (size_t)task_cap * sizeof(MixedTask) + partial_cap * sizeof(float2) + starts_cap * sizeof(int) +
         core_cap * sizeof(double) + far_geom_cap * sizeof(double4) + far_rec_cap * sizeof(FarVal) +
         far_blk_cap * (2 * sizeof(int4) + sizeof(int) * (kFarList + 1)) + (size_t)far_kind_cap * sizeof(int) +
         2 * sizeof(int) * (size_t)far_first_cap + near_u_cap * sizeof(int2) + near_rng_cap * sizeof(int4);
}

int MixedRunner::far_time(float* ms, unsigned long long* evals) {
  *ms = 0.0f;
  *evals = 0;
  if (!far_timed || !plan.far) return GX_OK;
  hipError_t e;
  if ((e = hipEventElapsedTime(ms, far_ev[0], far_ev[1])) ||
      (e = hipMemcpy(evals, far_evals, sizeof(unsigned long long), hipMemcpyDeviceToHost)))
    return hip_fail(e, "far-field timing");
  return GX_OK;
}

int MixedRunner::interp_pairs(long long* pairs) {
  *pairs = 0;
  if (!interp_cnt || plan.nitems <= 0) return GX_OK;
  unsigned long long slots[kEvalSlots], n = 0;
  hipError_t e;
  if ((e = hipMemcpy(slots, interp_cnt, sizeof(slots), hipMemcpyDeviceToHost))) return hip_fail(e, "interpolation counters");
  for (auto v : slots) n += v;
  *pairs = (long long)n * 64;
  return GX_OK;
}

// Barycentric interpolation from the Chebyshev-Lobatto nodes s_k = cos(pi k / (n - 1)) (weights
// (-1)^k, halved at the ends) to the bin positions s_j = (j - 31.5) / 31.5 of a 64-bin block;
// bins 0 and 63 are nodes n - 1 and 0.  FP64, rounded once.
void interp_table(float* tab) {
  constexpr int n = kInterpNodes;
  double sk[n], wk[n];
  for (int k = 0; k < n; ++k) {
    sk[k] = k == 0 ? 1.0 : k == n - 1 ? -1.0 : std::cos(std::acos(-1.0) * k / (n - 1));
    wk[k] = ((k & 1) ? -1.0 : 1.0) * ((k == 0 || k == n - 1) ? 0.5 : 1.0);
  }
  for (int j = 0; j < 64; ++j) {
    const double s = (j - 31.5) / 31.5;
    int hit = -1;
    for (int k = 0; k < n; ++k)
      if (s == sk[k]) hit = k;
    double d = 0.0;
    if (hit < 0)
      for (int k = 0; k < n; ++k) d += wk[k] / (s - sk[k]);
    for (int k = 0; k < n; ++k)
      tab[j * n + k] = hit >= 0 ? (k == hit ? 1.0f : 0.0f) : (float)((wk[k] / (s - sk[k])) / d);
  }
  // the anchored form's offset of each node: x = x(bin 32) + (32 - p_k) e, p_k = 31.5 (1 + s_k)
  for (int k = 0; k < n; ++k) tab[kInterpTabNodes + k] = (float)(0.5 - 31.5 * sk[k]);
}

double chunk_max_abs_nu(const GridView& gv) {
  const double d = gv.dfreq;
  return std::max(std::fabs(gv.start_nu + (double)(gv.start_idx - 1) * d),
                  std::fabs(gv.start_nu + (double)(gv.start_idx + gv.points + 1) * d));
}

// A core bin j of a line satisfies |nu - f_j| < (15 - y) alpha <= 15 dpw nu, so
// nu < F / (1 - 15 dpw) with F the largest |f| of the chunk, and the centre bin c is
// within 15 dpw F / ((1 - 15 dpw) dnu) + 1/2 of j; kmax adds a margin of 3 bins.
int core_kmax(const GridView& gv, double dpw) {
  const double w = 15.0 * dpw;
  if (!(w >= 0.0 && w < 0.5) || !(gv.dfreq > 0.0)) return -1;
  const double kb = w * chunk_max_abs_nu(gv) / ((1.0 - w) * gv.dfreq);
  if (!(kb < 1e6)) return -1;
  const int kmax = (int)std::ceil(kb) + 3;
  return core_lds_bytes(64 + 2 * kmax + 1) ? kmax : -1;
}

bool core_pass_fits(const GridView& gv, double dpw) { return gv.points <= 0 || core_kmax(gv, dpw) >= 0; }

// Centre bins per workgroup from the grid alone (a core half-width estimate at
// dpw = 1e-5), so that a task's rows, hence its sums, do not depend on the other
// tasks of the launch.  Row boundaries sit at multiples of cpw in absolute bins
// for the same reason.
CoreGeom core_geometry(const GridView& gv, int kmax) {
  CoreGeom g{};
  const double kg = std::min(1e6, 15.0 * 1e-5 * chunk_max_abs_nu(gv) / gv.dfreq);
  int cpw = (int)std::min(256.0, std::max(64.0, std::ceil(2.0 * kg / 64.0) * 64.0));
  if (core_lds_bytes(cpw + 2 * kmax + 1) == 0) cpw = 64;
  const long long lo = (long long)gv.start_idx - kmax - 1;
  const long long cbase = (lo >= 0 ? lo / cpw : -((-lo + cpw - 1) / cpw)) * cpw;
  const long long cend = (long long)gv.start_idx + gv.points + kmax + 1;
  g.cpw = cpw;
  g.span = cpw + 2 * kmax + 1;
  g.cbase = (int)cbase;
  g.ncwg = (int)((cend - cbase + cpw - 1) / cpw);
  g.ncent = g.ncwg * cpw;
  return g;
}

size_t core_rows_bytes(const GridView& gv, double dpw) {
  const int kmax = core_kmax(gv, dpw);
  if (kmax < 0 || gv.points <= 0) return 0;
  const CoreGeom g = core_geometry(gv, kmax + kNearMargin);   // room for near ranges
  return (size_t)g.ncwg * (size_t)g.span * sizeof(double);
}

int MixedRunner::prepare(const MixedTask* tasks, int ntasks, int nblocks, const GridView& gv, int nlines,
                         hipStream_t s) {
  int rc = init();
  if (rc) return rc;
  plan = MixedPlan{};
  if (nblocks <= 0 || ntasks <= 0) return GX_OK;
  hipError_t e;
  if (ntasks > task_cap) {
    if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
    h_tasks.clear();
    hipFree(d_tasks);
    d_tasks = nullptr;
    if ((e = hipMalloc(&d_tasks, sizeof(MixedTask) * (size_t)ntasks))) return hip_fail(e, "hipMalloc(tasks)");
    task_cap = ntasks;
  }
  const bool same = h_tasks.size() == (size_t)ntasks &&
                    std::memcmp(h_tasks.data(), tasks, sizeof(MixedTask) * (size_t)ntasks) == 0;
  if (!same) {
    // the previous launch may still read the table: wait before overwriting it
    if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
    h_tasks.assign(tasks, tasks + ntasks);
    if ((e = hipMemcpy(d_tasks, h_tasks.data(), sizeof(MixedTask) * (size_t)ntasks, hipMemcpyHostToDevice)))
      return hip_fail(e, "mixed: task upload");
  }
  const int slots = cus * mixed_occupancy(R) * kWavesPerGroup;   // resident waves
  const long long base = (long long)ntasks * nblocks;
  int nseg = forced_nseg;
  if (nseg <= 0) {
    const long long target = (long long)items_per_slot * slots;
    nseg = (int)std::min<long long>(16, std::max<long long>(1, (target + base - 1) / base));
  }
  const long long nitems = base * nseg;
  if (nitems > INT32_MAX) return fail(GX_INVALID_ARG, "mixed: too many work items");
  if (nseg > 1) {
    const size_t need = (size_t)nitems * (size_t)bins_per_block();
    if (need > partial_cap) {
      if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
      hipFree(partial);
      partial = nullptr;
      if ((e = hipMalloc(&partial, need * sizeof(float2)))) return hip_fail(e, "hipMalloc(partials)");
      partial_cap = need;
    }
  }
  plan.nblocks = nblocks;
  plan.nseg = nseg;
  plan.ntasks = ntasks;
  plan.tasks = d_tasks;
  plan.partial = partial;
  plan.counter = counter;
  plan.nitems = (int)nitems;
  const long long waves = persistent ? std::min<long long>(nitems, slots) : nitems;
  plan.grid = (int)((waves + kWavesPerGroup - 1) / kWavesPerGroup);
  plan.queue = persistent ? 1 : 0;

  plan.nlines = std::max(nlines, 0);
  const bool far_plan = far_on && plan.nlines >= 2 * kFarLeaf && gv.points > 0;
  bool near_plan = far_plan && near_on;
  if (gv.points > 0 && plan.nlines > 0) {
    int kmax = 0;
    for (int k = 0; k < ntasks; ++k) {
      const int kk = core_kmax(gv, tasks[k].dpwcoef);
      if (kk < 0) return fail(GX_INVALID_ARG, "mixed: W4 core too wide for the core pass; use FP64 for this grid");
      kmax = std::max(kmax, kk);
    }
    // near ranges reach a little beyond the cores: kNearMargin more bins of row where LDS holds them
    if (near_plan && core_lds_bytes(64 + 2 * (kmax + kNearMargin) + 1)) kmax += kNearMargin;
    else near_plan = false;
    const CoreGeom cg = core_geometry(gv, kmax);
    plan.kmax = kmax;
    plan.cpw = cg.cpw;
    plan.span = cg.span;
    plan.cbase = cg.cbase;
    plan.ncwg = cg.ncwg;
    plan.ncent = cg.ncent;
    const int span = cg.span;
    const size_t need_starts = (size_t)plan.ncent + 1;
    const size_t need_rows = (size_t)ntasks * plan.ncwg * span;
    if (need_starts > starts_cap || need_rows > core_cap) {
      if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
      if (need_starts > starts_cap) {
        hipFree(starts);
        starts = nullptr;
        starts_cap = 0;
        if ((e = hipMalloc(&starts, need_starts * sizeof(int)))) return hip_fail(e, "hipMalloc(core starts)");
        starts_cap = need_starts;
      }
      if (need_rows > core_cap) {
        hipFree(core_part);
        core_part = nullptr;
        core_cap = 0;
        if ((e = hipMalloc(&core_part, need_rows * sizeof(double)))) return hip_fail(e, "hipMalloc(core rows)");
        core_cap = need_rows;
      }
    }
    plan.starts = starts;
    plan.core_part = core_part;
  }
  // far field: leaves of kFarLeaf lines and their binary tree
  plan.far = 0;
  plan.near = 0;
  plan.core_evals = core_counter;
  plan.interp_tab = interp_on ? interp_tab : nullptr;
  plan.interp_cnt = interp_cnt;
  if (far_plan) {
    int n = (plan.nlines + kFarLeaf - 1) / kFarLeaf, lev = 0, tot = 0;
    while (true) {
      if (lev >= kFarMaxLevels) return fail(GX_INVALID_ARG, "mixed: too many far-field levels");
      plan.far_off[lev] = tot;
      plan.far_n[lev] = n;
      tot += n;
      ++lev;
      if (n == 1) break;
      n = (n + 1) / 2;
    }
    plan.far_nlev = lev;
    plan.far_nclus = tot;
    const size_t need_rec = (size_t)ntasks * tot * kFarRec;
    const size_t need_blk = (size_t)ntasks * nblocks;
    const int need_first = gv.points + 1;
    if ((size_t)tot > far_geom_cap || need_rec > far_rec_cap || need_blk > far_blk_cap || ntasks > far_kind_cap ||
        need_first > far_first_cap) {
      if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
      if ((size_t)tot > far_geom_cap) {
        hipFree(far_geom);
        far_geom = nullptr;
        far_geom_cap = 0;
        if ((e = hipMalloc(&far_geom, sizeof(double4) * (size_t)tot))) return hip_fail(e, "hipMalloc(far geometry)");
        far_geom_cap = tot;
      }
      if (need_rec > far_rec_cap) {
        hipFree(far_rec);
        far_rec = nullptr;
        far_rec_cap = 0;
        if ((e = hipMalloc(&far_rec, sizeof(FarVal) * need_rec))) return hip_fail(e, "hipMalloc(far moments)");
        far_rec_cap = need_rec;
      }
      if (need_blk > far_blk_cap) {
        hipFree(far_rng);
        hipFree(edge_rng);
        hipFree(far_list);
        hipFree(far_cnt);
        far_rng = nullptr;
        edge_rng = nullptr;
        far_list = nullptr;
        far_cnt = nullptr;
        far_blk_cap = 0;
        if ((e = hipMalloc(&far_rng, sizeof(int4) * need_blk)) || (e = hipMalloc(&edge_rng, sizeof(int4) * need_blk)) ||
            (e = hipMalloc(&far_list, sizeof(int) * need_blk * kFarList)) ||
            (e = hipMalloc(&far_cnt, sizeof(int) * need_blk)))
          return hip_fail(e, "hipMalloc(far plans)");
        far_blk_cap = need_blk;
      }
      if (ntasks > far_kind_cap) {
        hipFree(far_kind);
        far_kind = nullptr;
        far_kind_cap = 0;
        if ((e = hipMalloc(&far_kind, sizeof(int) * (size_t)ntasks))) return hip_fail(e, "hipMalloc(far kinds)");
        far_kind_cap = ntasks;
      }
      if (need_first > far_first_cap) {
        hipFree(far_first);
        far_first = nullptr;
        far_first_cap = 0;
        if ((e = hipMalloc(&far_first, 2 * sizeof(int) * (size_t)need_first)))
          return hip_fail(e, "hipMalloc(far bin bounds)");
        far_first_cap = need_first;
      }
    }
    plan.far = 1;
    plan.far_bins = bins_per_block();
    plan.far_geom = far_geom;
    plan.far_rec = far_rec;
    plan.far_kind = far_kind;
    plan.far_rng = far_rng;
    plan.edge_rng = edge_rng;
    plan.far_list = far_list;
    plan.far_cnt = far_cnt;
    plan.far_ie1 = far_first;
    plan.far_ib1 = far_first + far_first_cap;
    plan.far_bin_cap = far_first_cap;
    plan.far_evals = far_evals;
    if (near_plan) {
      const size_t need_u = (size_t)ntasks * tot;
      if (need_u > near_u_cap || need_blk > near_rng_cap) {
        if ((e = hipStreamSynchronize(s))) return hip_fail(e, "mixed: sync");
        if (need_u > near_u_cap) {
          hipFree(near_u);
          near_u = nullptr;
          near_u_cap = 0;
          if ((e = hipMalloc(&near_u, sizeof(int2) * need_u))) return hip_fail(e, "hipMalloc(near ranges)");
          near_u_cap = need_u;
        }
        if (need_blk > near_rng_cap) {
          hipFree(near_rng);
          near_rng = nullptr;
          near_rng_cap = 0;
          if ((e = hipMalloc(&near_rng, sizeof(int4) * need_blk))) return hip_fail(e, "hipMalloc(near runs)");
          near_rng_cap = need_blk;
        }
      }
      plan.near = 1;
      plan.near_u = near_u;
      plan.near_rng = near_rng;
    }
  }
  return GX_OK;
}

// Bytes of the far-field buffers for ntasks points over nlines lines and nblocks blocks.
size_t far_field_bytes(int ntasks, long long nlines, int nblocks) {
  const long long clusters = 2 * ((nlines + kFarLeaf - 1) / kFarLeaf) + 64;
  return (size_t)clusters * (sizeof(double4) + (size_t)ntasks * (kFarRec * sizeof(FarVal) + sizeof(int2))) +
         (size_t)ntasks * nblocks * (3 * sizeof(int4) + sizeof(int) * (kFarList + 1)) + sizeof(int) * (size_t)ntasks +
         2 * sizeof(int) * ((size_t)nblocks * 64 + 1);    // per-bin line bounds (callers count 64-bin blocks)
}

int MixedRunner::launch(const GridView& gv, const double* nu, const int2* ranges, const int* ib, const int* ie,
                        hipStream_t s, hipEvent_t* ev) {
  if (plan.nitems <= 0) return GX_OK;
  plan.gv = gv;
  plan.nu = nu;
  plan.ranges = ranges;
  plan.ib = ib;
  plan.ie = ie;
  hipError_t e;
  far_timed = false;
  if (plan.far) {
    if (ev && (e = hipEventRecord(far_ev[0], s))) return hip_fail(e, "event");
    if ((e = launch_far_setup(plan, s))) return hip_fail(e, "far-field setup launch");
    if ((e = launch_far_eval(plan, R, s))) return hip_fail(e, "far-field launch");
    if (ev && (e = hipEventRecord(far_ev[1], s))) return hip_fail(e, "event");
    far_timed = ev != nullptr;
  }
  if ((e = launch_fill(reinterpret_cast<double*>(interp_cnt), kEvalSlots, 0.0, s)))
    return hip_fail(e, "interpolation counters");
  if (ev && (e = hipEventRecord(ev[0], s))) return hip_fail(e, "event");
  if ((e = launch_accum_mixed(plan, R, s))) return hip_fail(e, "mixed accumulate launch");
  if (ev && (e = hipEventRecord(ev[1], s))) return hip_fail(e, "event");
  if ((e = launch_core_mixed(plan, s))) return hip_fail(e, "mixed core launch");
  if (ev && (e = hipEventRecord(ev[2], s))) return hip_fail(e, "event");
  return GX_OK;
}

// One (T,P) point over [start_idx, start_idx+points) of a grid: prep, ranges, accumulate.
int run_point(const PointConsts& pc, const GridView& gv, const LineStore& ls, int nlines, int precision,
              Derived& der, Ranges& rng, MixedRunner& mix, double* intens, unsigned long long* d_evals,
              unsigned long long* d_core_evals, hipStream_t s, hipEvent_t* ev) {
  LinesView lv{ls.nu, ls.elow, ls.aif, ls.gns, ls.gam, ls.nexp, nlines};
  int rc = der.reserve((size_t)std::max(nlines, 1));
  if (rc) return rc;
  // the mixed form where its W4 core pass fits the grid; the FP64 path (more accurate) otherwise
  const bool mixed = precision == GX_PREC_MIXED && pc.profile == GX_VOIGT && core_pass_fits(gv, pc.dpwcoef);
  if (mixed && (rc = mix.init())) return rc;   // fixes R (GX_MIXED_R) before the block size is used
  const int bpb = mixed ? mix.bins_per_block() : kBinsF64;
  const int nblocks = gv.points > 0 ? (gv.points + bpb - 1) / bpb : 0;
  if ((rc = rng.reserve(nblocks))) return rc;
  hipError_t e;
  if (mixed) {
    if ((rc = der.reserve_mixed((size_t)std::max(nlines, 1), 1, 1))) return rc;
    if ((rc = der.set_points(&pc, 1, s))) return rc;
    MixedTask task = mixed_task(der.m, 0, 0, pc, lv.gamma, intens);
    mix.core_counter = d_core_evals;
    if ((rc = mix.prepare(&task, 1, nblocks, gv, nlines, s))) return rc;
  }
  if (ev && (e = hipEventRecord(ev[0], s))) return hip_fail(e, "event");
  if (mixed) {
    const WindowArgs wa{pc.cutoff, pc.convention, 1};
    if ((e = launch_windows(wa, gv, ls.nu, nlines, der.d.ib, der.d.ie, d_evals, s))) return hip_fail(e, "windows");
    if (d_core_evals && (e = launch_fill(reinterpret_cast<double*>(d_core_evals), kEvalSlots, 0.0, s)))
      return hip_fail(e, "fill");
    if ((e = launch_prep_mixed(der.d_pcs, 1, lv, gv, der.d.ib, der.d.ie, der.m, d_core_evals, s)))
      return hip_fail(e, "mixed prep launch");
  } else {
    if ((e = launch_line_prep_f64(pc, gv, lv, der.d, d_evals, s))) return hip_fail(e, "line prep launch");
  }
  if (ev && (e = hipEventRecord(ev[1], s))) return hip_fail(e, "event");
  const double cut = max_cutoff(pc, ls.max_nu, gv.dfreq);
  if ((e = launch_block_ranges(gv, ls.nu, nlines, cut, bpb, nblocks, rng.r, s))) return hip_fail(e, "ranges launch");
  if (ev && (e = hipEventRecord(ev[2], s))) return hip_fail(e, "event");
  if (mixed) {
    if ((rc = mix.launch(gv, ls.nu, rng.r, der.d.ib, der.d.ie, s, ev ? ev + 4 : nullptr))) return rc;
  } else {
    if ((e = launch_accum_f64(pc, gv, ls.nu, der.d, rng.r, nblocks, intens, s))) return hip_fail(e, "accumulate launch");
  }
  if (ev && (e = hipEventRecord(ev[3], s))) return hip_fail(e, "event");
  return GX_OK;
}

}  // namespace gx

using namespace gx;

struct gx_ctx {
  int device = 0;
  int profile = GX_VOIGT;
  size_t budget = 0;
  size_t ntrans = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[7] = {};       // prep, ranges, accumulate; mixed: region I | core pass split
  bool timed = false;
  bool last_mixed = false;
  unsigned long long* d_core_evals = nullptr;   // kEvalSlots counters
  int M = 0;
  double* d_freq = nullptr;
  double* d_intens = nullptr;
  double start_nu = 0.0;
  gx_consts consts{};
  bool have_consts = false;
  LineStore lines;
  Derived derived;
  Ranges ranges;
  MixedRunner mix;
  unsigned long long* d_evals = nullptr;   // kEvalSlots counters
  int last_lines = 0;
  int noalloc_status = GX_OK;              // set-up of the device's scratch array (gx_execute_noalloc)
  std::string noalloc_error;
};

namespace {

constexpr size_t kBytesPerLine = 44 + 60 + 52;   // resident + derived (FP64 + mixed) + sort staging
// Chunk capacity reported by GetNtrans: the reference's drivers allocate six host arrays of
// that many lines (VoigtProfile.cpp:28-33), so 2^26 lines (3 GB of host buffers) by default;
// GX_MAX_NTRANS raises or lowers it (at most 2^30, which keeps N_ener an int).
size_t max_ntrans() {
  size_t v = size_t(1) << 26;
  if (const char* env = std::getenv("GX_MAX_NTRANS")) {
    const long long x = std::atoll(env);
    if (x > 0) v = std::min<size_t>((size_t)x, size_t(1) << 30);
  }
  return v;
}

int set_device(gx_ctx* c) {
  hipError_t e = hipSetDevice(c->device);
  return e ? hip_fail(e, "hipSetDevice") : GX_OK;
}

// The sum a chunk takes: a Voigt chunk uploaded without gamma and n takes the CUDA manager's
// no-gamma sum (GPUManager.cpp:139-143 chooses it by the gamma buffer being absent).
int sum_profile(const gx_ctx* c) {
  if (c->profile == GX_VOIGT && c->lines.n > 0 && !c->lines.has_gamma) return kProfVoigtNoGamma;
  return c->profile;
}

}  // namespace

extern "C" {

int gx_create(int device, int profile, size_t mem_budget, gx_ctx** out) {
  if (!out) return fail(GX_INVALID_ARG, "gx_create: out is NULL");
  *out = nullptr;
  if (profile != GX_GAUSSIAN && profile != GX_DOPPLER && profile != GX_VOIGT)
    return fail(GX_INVALID_ARG, "gx_create: only the GAUSSIAN (0), DOPPLER (1) and VOIGT (2) profiles are implemented");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e || ndev <= 0) return fail(GX_NO_DEVICE, "gx_create: no HIP device");
  if (device < 0 || device >= ndev) return fail(GX_NO_DEVICE, "gx_create: device index out of range");
  if ((e = hipSetDevice(device))) return hip_fail(e, "hipSetDevice");
  size_t free_b = 0, total_b = 0;
  if ((e = hipMemGetInfo(&free_b, &total_b))) return hip_fail(e, "hipMemGetInfo");
  gx_ctx* c = new gx_ctx();
  c->device = device;
  c->profile = profile;
  // GPUManager.cpp:12 keeps 300 MB back; 1 GiB here leaves room for scratch.
  c->budget = mem_budget ? mem_budget : (total_b > (size_t(1) << 30) ? total_b - (size_t(1) << 30) : total_b / 2);
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking))) { delete c; return hip_fail(e, "stream"); }
  for (auto& ev : c->ev)
    if ((e = hipEventCreate(&ev))) { gx_destroy(c); return hip_fail(e, "event"); }
  if ((e = hipMalloc(&c->d_evals, sizeof(unsigned long long) * kEvalSlots)) ||
      (e = hipMalloc(&c->d_core_evals, sizeof(unsigned long long) * kEvalSlots))) {
    gx_destroy(c);
    return hip_fail(e, "hipMalloc");
  }
  // the scratch array of gx_execute_noalloc: its address on this device, looked up once;
  // a failure is kept for that entry point and does not fail the context
  if ((c->noalloc_status = noalloc_setup(device)) != GX_OK) c->noalloc_error = g_last_error;
  *out = c;
  return GX_OK;
}

int gx_initialize_vectors(gx_ctx* c, int npoints, size_t* ntrans_out) {
  if (!c || npoints <= 0) return fail(GX_INVALID_ARG, "initialize_vectors: bad arguments");
  int rc = set_device(c);
  if (rc) return rc;
  hipFree(c->d_freq);
  hipFree(c->d_intens);
  c->d_freq = c->d_intens = nullptr;
  hipError_t e;
  if ((e = hipMalloc(&c->d_freq, sizeof(double) * (size_t)npoints)) ||
      (e = hipMalloc(&c->d_intens, sizeof(double) * (size_t)npoints)))
    return hip_fail(e, "initialize_vectors: hipMalloc(grid)");
  if ((e = hipMemset(c->d_intens, 0, sizeof(double) * (size_t)npoints))) return hip_fail(e, "memset");
  c->M = npoints;
  const size_t grid_bytes = 16 * (size_t)npoints;
  const size_t avail = c->budget > grid_bytes ? c->budget - grid_bytes : 0;
  c->ntrans = std::min(avail / kBytesPerLine, max_ntrans());   // GPUManager.cpp:55-58
  if (ntrans_out) *ntrans_out = c->ntrans;
  return GX_OK;
}

int gx_initialize_constants(gx_ctx* c, const gx_consts* k) {
  if (!c || !k) return fail(GX_INVALID_ARG, "initialize_constants: NULL argument");
  if (!(k->temperature > 0) || !(k->mean_mass > 0) || !(k->dfreq > 0) || !(k->partition > 0) ||
      !(k->ref_temp > 0) || !std::isfinite(k->pressure))
    return fail(GX_INVALID_ARG, "initialize_constants: temperature, mean_mass, dfreq, partition and ref_temp must be > 0");
  if (k->convention != GX_CONV_OPENMP && k->convention != GX_CONV_GPU)
    return fail(GX_INVALID_ARG, "initialize_constants: unknown convention");
  if (k->precision != GX_PREC_FP64 && k->precision != GX_PREC_MIXED)
    return fail(GX_INVALID_ARG, "initialize_constants: unknown precision");
  if (c->profile == GX_GAUSSIAN && !(k->half_width > 0))
    return fail(GX_INVALID_ARG, "initialize_constants: the Gaussian profile needs half_width > 0");
  c->consts = *k;
  c->have_consts = true;
  return GX_OK;
}

int gx_transfer_freq(gx_ctx* c, const double* freq, const double* intens, int n) {
  if (!c || !freq || n <= 0) return fail(GX_INVALID_ARG, "transfer_freq: bad arguments");
  if (n != c->M) return fail(GX_INVALID_ARG, "transfer_freq: N differs from initialize_vectors");
  int rc = set_device(c);
  if (rc) return rc;
  hipError_t e;
  if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e, "sync");
  if ((e = hipMemcpy(c->d_freq, freq, sizeof(double) * (size_t)n, hipMemcpyHostToDevice)))
    return hip_fail(e, "transfer_freq: copy");
  if (intens)
    e = hipMemcpy(c->d_intens, intens, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
  else
    e = hipMemset(c->d_intens, 0, sizeof(double) * (size_t)n);
  if (e) return hip_fail(e, "transfer_freq: intens");
  c->start_nu = freq[0];   // OpenMPManager.cpp:87
  return GX_OK;
}

int gx_transfer_vectors(gx_ctx* c, size_t nener, const double* energies, const double* nu, const double* aif,
                        const int* gns, const double* gamma, const double* nexp) {
  if (!c) return fail(GX_INVALID_ARG, "transfer_vectors: NULL context");
  if (nener > 0 && (!energies || !nu || !aif || !gns))
    return fail(GX_INVALID_ARG, "transfer_vectors: NULL line array");
  if (c->profile == GX_VOIGT && nener > 0 && (!gamma != !nexp))
    return fail(GX_INVALID_ARG, "transfer_vectors: Voigt takes gamma and n together (or neither)");
  if (nener > c->ntrans && c->ntrans) return fail(GX_INVALID_ARG, "transfer_vectors: more lines than GetNtrans()");
  if (nener > (size_t)INT32_MAX) return fail(GX_INVALID_ARG, "transfer_vectors: chunk too large");
  int rc = set_device(c);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(c->stream);   // GPUManager.cpp:89: wait for the previous chunk
  if (e) return hip_fail(e, "transfer_vectors: sync");
  const bool voigt = c->profile == GX_VOIGT;
  return c->lines.upload(nener, energies, nu, aif, gns, voigt ? gamma : nullptr, voigt ? nexp : nullptr, c->stream);
}

int gx_execute(gx_ctx* c, int points, int n_ener, int start_idx) {
  if (!c) return fail(GX_INVALID_ARG, "execute: NULL context");
  if (!c->have_consts) return fail(GX_INVALID_ARG, "execute: initialize_constants not called");
  if (!c->d_freq) return fail(GX_INVALID_ARG, "execute: initialize_vectors not called");
  if (points < 0 || start_idx < 0 || (long long)start_idx + points > c->M)
    return fail(GX_INVALID_ARG, "execute: bin range outside the grid");
  if (n_ener < 0 || (size_t)n_ener > c->lines.n) return fail(GX_INVALID_ARG, "execute: N_ener exceeds the uploaded chunk");
  int rc = set_device(c);
  if (rc) return rc;
  const gx_consts& k = c->consts;
  PointConsts pc = make_point_consts(sum_profile(c), k, 1.0);
  GridView gv{c->d_freq, c->M, start_idx, points, c->start_nu, k.dfreq};
  hipError_t e = hipMemsetAsync(c->d_evals, 0, sizeof(unsigned long long) * kEvalSlots, c->stream);
  if (e) return hip_fail(e, "execute: memset");
  c->last_lines = n_ener;
  c->timed = true;
  c->last_mixed = k.precision == GX_PREC_MIXED && pc.profile == GX_VOIGT && core_pass_fits(gv, pc.dpwcoef);
  return run_point(pc, gv, c->lines, n_ener, k.precision, c->derived, c->ranges, c->mix, c->d_intens, c->d_evals,
                   c->d_core_evals, c->stream, c->ev);
}

int gx_execute_noalloc(gx_ctx* c, int points, int n_ener, int start_idx) {
  if (!c) return fail(GX_INVALID_ARG, "execute_noalloc: NULL context");
  if (!c->have_consts) return fail(GX_INVALID_ARG, "execute_noalloc: initialize_constants not called");
  if (!c->d_freq) return fail(GX_INVALID_ARG, "execute_noalloc: initialize_vectors not called");
  if (points < 0 || start_idx < 0 || (long long)start_idx + points > c->M)
    return fail(GX_INVALID_ARG, "execute_noalloc: bin range outside the grid");
  if (n_ener < 0 || (size_t)n_ener > c->lines.n)
    return fail(GX_INVALID_ARG, "execute_noalloc: N_ener exceeds the uploaded chunk");
  if (c->consts.precision != GX_PREC_FP64)
    return fail(GX_INVALID_ARG, "execute_noalloc: runs the FP64 path only (precision must be GX_PREC_FP64)");
  if (c->noalloc_status != GX_OK) return fail(c->noalloc_status, c->noalloc_error);
  int rc = set_device(c);
  if (rc) return rc;
  const gx_consts& k = c->consts;
  PointConsts pc = make_point_consts(sum_profile(c), k, 1.0);
  GridView gv{c->d_freq, c->M, start_idx, points, c->start_nu, k.dfreq};
  c->timed = false;   // gx_last_timing covers gx_execute only
  return execute_noalloc(c->device, pc, gv, c->lines, n_ener, c->d_intens, c->stream);
}

int gx_ready(gx_ctx* c) {
  if (!c) return -fail(GX_INVALID_ARG, "ready: NULL context");
  hipError_t e = hipStreamQuery(c->stream);    // GPUManager.cpp:185-188
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return -hip_fail(e, "ready");
}

int gx_transfer_results(gx_ctx* c, double* freq, double* intens, int n) {
  if (!c || !intens || n <= 0) return fail(GX_INVALID_ARG, "transfer_results: bad arguments");
  if (n > c->M) return fail(GX_INVALID_ARG, "transfer_results: N exceeds the grid");
  int rc = set_device(c);
  if (rc) return rc;
  hipError_t e;
  if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e, "transfer_results: kernel");
  if (freq && (e = hipMemcpy(freq, c->d_freq, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost)))
    return hip_fail(e, "transfer_results: copy freq");
  if ((e = hipMemcpy(intens, c->d_intens, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost)))
    return hip_fail(e, "transfer_results: copy intens");
  return GX_OK;
}

int gx_cleanup(gx_ctx* c) {
  if (!c) return fail(GX_INVALID_ARG, "cleanup: NULL context");
  set_device(c);
  hipStreamSynchronize(c->stream);
  hipFree(c->d_freq);
  hipFree(c->d_intens);
  c->d_freq = c->d_intens = nullptr;
  c->M = 0;
  c->lines.release();
  c->derived.release();
  c->ranges.release();
  c->mix.release();
  return GX_OK;
}

int gx_destroy(gx_ctx* c) {
  if (!c) return GX_OK;
  gx_cleanup(c);
  for (auto& ev : c->ev) if (ev) hipEventDestroy(ev);
  if (c->stream) hipStreamDestroy(c->stream);
  hipFree(c->d_evals);
  hipFree(c->d_core_evals);
  delete c;
  return GX_OK;
}

size_t gx_get_ntrans(gx_ctx* c) { return c ? c->ntrans : 0; }

int gx_synchronize(gx_ctx* c) {
  if (!c) return fail(GX_INVALID_ARG, "synchronize: NULL context");
  hipError_t e = hipStreamSynchronize(c->stream);
  return e ? hip_fail(e, "synchronize") : GX_OK;
}

int gx_last_timing(gx_ctx* c, gx_timing* out) {
  if (!c || !out) return fail(GX_INVALID_ARG, "last_timing: NULL argument");
  if (!c->timed) return fail(GX_NOT_FOUND, "last_timing: nothing executed yet");
  hipError_t e;
  if ((e = hipStreamSynchronize(c->stream))) return hip_fail(e, "last_timing: sync");
  float a = 0, b = 0;
  if ((e = hipEventElapsedTime(&a, c->ev[0], c->ev[1])) || (e = hipEventElapsedTime(&b, c->ev[2], c->ev[3])))
    return hip_fail(e, "last_timing: events");
  unsigned long long slots[kEvalSlots], ev = 0;
  if ((e = hipMemcpy(slots, c->d_evals, sizeof(slots), hipMemcpyDeviceToHost))) return hip_fail(e, "last_timing: copy");
  for (unsigned long long v : slots) ev += v;
  float tot = 0;
  if ((e = hipEventElapsedTime(&tot, c->ev[0], c->ev[3]))) return hip_fail(e, "last_timing: events");
  out->intensity_ms = a;
  out->accumulate_ms = b;
  out->total_ms = tot;
  out->evals = (long long)ev;
  out->lines = c->last_lines;
  out->region1_ms = b;
  out->core_ms = 0.0f;
  out->core_evals = 0;
  out->far_ms = 0.0f;
  out->far_evals = 0;
  out->interp_pairs = 0;
  if (c->last_mixed) {
    float r1 = 0, co = 0;
    if ((e = hipEventElapsedTime(&r1, c->ev[4], c->ev[5])) || (e = hipEventElapsedTime(&co, c->ev[5], c->ev[6])))
      return hip_fail(e, "last_timing: events");
    unsigned long long cs[kEvalSlots], ce = 0;
    if ((e = hipMemcpy(cs, c->d_core_evals, sizeof(cs), hipMemcpyDeviceToHost))) return hip_fail(e, "last_timing: copy");
    for (unsigned long long v : cs) ce += v;
    out->region1_ms = r1;
    out->core_ms = co;
    out->core_evals = (long long)ce;
    unsigned long long fe = 0;
    const int frc = c->mix.far_time(&out->far_ms, &fe);
    if (frc) return frc;
    out->far_evals = (long long)fe;
    const int irc = c->mix.interp_pairs(&out->interp_pairs);
    if (irc) return irc;
  }
  return GX_OK;
}

const char* gx_last_error(void) { return g_last_error.c_str(); }

const char* gx_version(void) { return "gx 0.1 (gfx950)"; }

}  // extern "C"
