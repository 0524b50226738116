// gx_kernels.hip -- HIP kernels of the cross-section path for gfx950 (MI355X).
//
//   line prep   one thread per line, FP64: S_ij(T), 1/alpha_D, y = gamma_T/alpha_D and
//               the integer window [ib, ie) of the reference (OpenMPManager.cpp:196-206,
//               cross_kernal.cu:758-790).  HBM-bound.
//   ranges      one thread per bin block: the slice of the nu-sorted list that can touch it.
//   accumulate  bin-parallel: a workgroup owns a contiguous run of bins, streams its line
//               slice through LDS in 256-line tiles (coalesced SoA loads), and every lane
//               sums its own bins in line order, so each grid is deterministic.  No atomics.
//
// Bin blocks are not remapped per XCD: the hardware deals blocks round-robin to the 8 XCDs,
// which interleaves dense and sparse parts of the spectrum evenly, while a block's line
// slice is streamed once per block at ~0.1 B per evaluation, far from any cache limit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "gx_internal.h"
#include "gx_plain.h"
#include "gx_w4.h"

namespace gx {

namespace {

constexpr double kIsqrtPi = 0.564189584;  // defines.h:8 (the truncated literal the reference uses)

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// grid nodes f_i = start + i*dfreq exactly as BaseProfile.cpp:19 forms them (no FMA)
__global__ void k_grid(double* __restrict__ f, int n, double start, double dfreq) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = start + (double)i * dfreq;
}

__global__ void k_iota(uint32_t* idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) idx[i] = (uint32_t)i;
}

__global__ void k_gather(const uint32_t* __restrict__ perm, const double* __restrict__ e_in,
                         const double* __restrict__ a_in, const int* __restrict__ g_in,
                         const double* __restrict__ gam_in, const double* __restrict__ n_in, double* __restrict__ e,
                         double* __restrict__ a, int* __restrict__ g, double* __restrict__ gam, double* __restrict__ nx,
                         int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  e[i] = e_in[p];
  a[i] = a_in[p];
  g[i] = g_in[p];
  if (gam_in) {
    gam[i] = gam_in[p];
    nx[i] = n_in[p];
  }
}

// first j in [lo, hi) with fl(nu - f_j) <= bound (false ... true as j grows)
__device__ int first_le(double nu, double bound, const double* __restrict__ f, int lo, int hi) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (nu - f[mid] <= bound) hi = mid; else lo = mid + 1;
  }
  return lo;
}
__device__ int first_lt(double nu, double bound, const double* __restrict__ f, int lo, int hi) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (nu - f[mid] < bound) hi = mid; else lo = mid + 1;
  }
  return lo;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sum a per-thread count over a 256-thread block, then one atomic per block
// into one of kEvalSlots counters (spread so no single address serialises).
__device__ __forceinline__ void block_count(unsigned long long w, unsigned long long* evals) {
  __shared__ unsigned long long s_w[4];
  w = wave_sum_u64(w);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    if (tot) atomicAdd(&evals[blockIdx.x & (kEvalSlots - 1)], tot);
  }
}

// ---------------------------------------------------------------------------
// Line prep (FP64).  Evaluation order of every expression follows the
// reference so that S, the Doppler width and the window edges round exactly
// as in OpenMPManager.cpp:191-206; contraction into FMA is switched off here.
// ---------------------------------------------------------------------------
// One line i of lv for the point pc: its derived data and window; returns ie - ib.
__device__ __forceinline__ unsigned long long line_prep_one(const PointConsts& pc, const GridView& gv,
                                                            const LinesView& lv, const DerivedF64& out, int i) {
#pragma clang fp contract(off)
  unsigned long long w = 0;
  if (pc.profile == kProfGaussian || pc.profile == kProfVoigtNoGamma) {
    // the CUDA manager's intensity step of both (cross_kernal.cu:471-496, 731-756): its cmcoef,
    // no ISQRTPI, nu == 0 gives S = 0; per-bin windows |nu - f| <= C
    const double nu = lv.nu[i];
    const double a = lv.aif[i];
    const int gi = lv.gns[i];
    const double e = lv.elow[i];
    double S = pc.cmcoef * a * gi * exp(-pc.beta * e) * (1.0 - exp(-pc.beta * nu)) / (nu * nu * pc.partition);
    if (nu == 0) S = 0.0;
    double g, y = 0.0;
    if (pc.profile == kProfGaussian) {
      g = pc.dpwcoef;                              // x_ = sqrt(ln2) / HW (cross_kernal.cu:582)
    } else {
      g = nu * pc.dpwcoef;                         // gamma_G (cross_kernal.cu:885)
      y = kPlainSqrtLn2 * kNoGammaL / g;           // :887
    }
    const int s0 = gv.start_idx, s1 = gv.start_idx + gv.points;
    const int ib = first_le(nu, pc.cutoff, gv.freq, s0, s1);
    int ie = first_lt(nu, -pc.cutoff, gv.freq, s0, s1);
    if (S == 0.0 || ie < ib) ie = ib;
    out.S[i] = S;
    out.g[i] = g;
    out.y[i] = y;
    out.ib[i] = ib;
    out.ie[i] = ie;
    w = (unsigned long long)(ie - ib);
  } else {
    const bool voigt = pc.profile == 2;
    const bool gpu = pc.convention == 1;
    const double nu = lv.nu[i];
    double nu_if = nu;
    if ((voigt || !gpu) && nu_if == 0) nu_if = 1e-6;
    const double a = lv.aif[i];
    const int gi = lv.gns[i];
    const double e = lv.elow[i];
    double S;
    if (voigt) {
      S = pc.cmcoef * a * gi * kIsqrtPi * exp(-pc.beta * e) * (1.0 - exp(-pc.beta * nu_if)) /
          (nu_if * nu_if * pc.partition);
      if (gpu && gi == -1) S = a;  // cross_kernal.cu:777
    } else {
      S = pc.cmcoef * a * gi * exp(-pc.beta * e) * (1.0 - exp(-pc.beta * nu_if)) / (nu_if * nu_if * pc.partition);
      if (gpu && nu == 0) S = 0.0;  // cross_kernal.cu:647
    }
    const double g = 1.0 / (nu_if * pc.dpwcoef);
    double y = 0.0;
    if (voigt) {
      const double gammaL = (lv.gamma[i] * pc.pressure_scale) * pow(pc.ref_temp / pc.temperature, lv.nexp[i]);
      y = gammaL * g;
    }
    double cut = pc.cutoff;
    if (!voigt && pc.doppler_widths > 0.0) cut = pc.doppler_widths * (nu_if * pc.dpwcoef) + gv.dfreq;
    int ib, ie;
    const int s0 = gv.start_idx, s1 = gv.start_idx + gv.points;
    if (gpu) {
      ib = first_le(nu_if, cut, gv.freq, s0, s1);
      ie = first_lt(nu_if, -cut, gv.freq, s0, s1);
    } else {
      ib = (int)fmax(round((nu_if - cut - gv.start_nu) / gv.dfreq), (double)s0);
      ie = (int)fmin(round((nu_if + cut - gv.start_nu) / gv.dfreq), (double)s1);
    }
    if (S == 0.0 || ie < ib) ie = ib;  // OpenMPManager.cpp:163,208: zero lines add nothing
    out.S[i] = S;
    out.g[i] = g;
    if (voigt) out.y[i] = y;
    out.ib[i] = ib;
    out.ie[i] = ie;
    w = (unsigned long long)(ie - ib);
  }
  return w;
}

__global__ __launch_bounds__(256) void k_line_prep(PointConsts pc, GridView gv, LinesView lv, DerivedF64 out,
                                                   unsigned long long* __restrict__ evals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long w = i < lv.n ? line_prep_one(pc, gv, lv, out, i) : 0ull;
  if (evals) block_count(w, evals);
}

// ---------------------------------------------------------------------------
// Integer windows [ib, ie) of every line for one chunk and one cutoff: they
// depend on the grid and the cutoff only, so all (T,P) points with the same
// cutoff share them.  Same expressions as k_line_prep (OpenMPManager.cpp:205-206).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_windows(WindowArgs wa, GridView gv, const double* __restrict__ nu_s, int n,
                                                 int* __restrict__ IB, int* __restrict__ IE,
                                                 unsigned long long* __restrict__ evals) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long w = 0;
  if (i < n) {
    const double nu = nu_s[i];
    const double nu_if = (nu == 0 && wa.zero_to_tiny) ? 1e-6 : nu;
    const int s0 = gv.start_idx, s1 = gv.start_idx + gv.points;
    int ib, ie;
    if (wa.convention == 1) {
      ib = first_le(nu_if, wa.cutoff, gv.freq, s0, s1);
      ie = first_lt(nu_if, -wa.cutoff, gv.freq, s0, s1);
    } else {
      ib = (int)fmax(round((nu_if - wa.cutoff - gv.start_nu) / gv.dfreq), (double)s0);
      ie = (int)fmin(round((nu_if + wa.cutoff - gv.start_nu) / gv.dfreq), (double)s1);
    }
    if (ie < ib) ie = ib;
    IB[i] = ib;
    IE[i] = ie;
    w = (unsigned long long)(ie - ib);
  }
  if (evals) block_count(w, evals);
}

// ---------------------------------------------------------------------------
// Per-point line data of the mixed-precision path: S, 1/alpha and y in FP64
// exactly as k_line_prep, amp = S/alpha stored as FP32, and the line's W4 core
// inside its window: the bins j of [ib, ie) with |nu - f_j| g + y < 15, the FP64
// path's own region-I test (evaluated on the grid array itself).  The set is
// an interval (|nu - f_j| is V-shaped in j); it is found from a bracketing
// estimate one bin wider on each side and trimmed with the exact test.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool in_core(double nu, double g, double y, const double* freq, int j) {
#pragma clang fp contract(off)
  return fabs(nu - freq[j]) * g + y < 15.0;
}

// The temperature-dependent part of one line at one point: S, 1/alpha and (T_ref/T)^n.
struct PrepT {
  double S, g, tscale;
  double alpha;        // nu dpwcoef = 1/g up to rounding: only for the core bracket
};

__device__ __forceinline__ PrepT prep_mixed_t(const PointConsts& pc, double nu_if, double a, int gi, double elow,
                                              double nexp) {
#pragma clang fp contract(off)
  PrepT t;
  t.S = pc.cmcoef * a * gi * kIsqrtPi * exp(-pc.beta * elow) * (1.0 - exp(-pc.beta * nu_if)) /
        (nu_if * nu_if * pc.partition);
  if (pc.convention == 1 && gi == -1) t.S = a;
  t.alpha = nu_if * pc.dpwcoef;
  t.g = 1.0 / t.alpha;
  t.tscale = pow(pc.ref_temp / pc.temperature, nexp);
  return t;
}

// Whether two points share every constant prep_mixed_t reads.
__device__ __forceinline__ bool same_t(const PointConsts& a, const PointConsts& b) {
  return a.beta == b.beta && a.cmcoef == b.cmcoef && a.partition == b.partition && a.dpwcoef == b.dpwcoef &&
         a.ref_temp == b.ref_temp && a.temperature == b.temperature && a.convention == b.convention;
}

// The pressure-dependent part and the point's output: the core range (y as mixed_y forms it).
__device__ __forceinline__ unsigned long long prep_mixed_point(const PointConsts& pc, const PrepT& t, double nu_if,
                                                               double gamma, int ib, int ie, const GridView& gv,
                                                               double inv_dfreq, const MixedLineData& out, int i,
                                                               int p) {
#pragma clang fp contract(off)
  const double g = t.g;
  const double gammaL = (gamma * pc.pressure_scale) * t.tscale;
  const double y = gammaL * g;
  int k0 = 0, k1 = 0;
  if (y < 15.0 && ie > ib) {
    // bracket from products (rounding far below the one-bin margin), trimmed by the exact test
    const double half = (15.0 - y) * t.alpha;
    const double lo = floor((nu_if - half - gv.start_nu) * inv_dfreq) - 1.0;
    const double hi = ceil((nu_if + half - gv.start_nu) * inv_dfreq) + 1.0;
    const int a0 = (int)fmax(lo, (double)ib), z0 = (int)fmin(hi, (double)ie);
    if (a0 < z0) {
      // The core is a contiguous run of the bracket (|nu - f_j| grows away from nu), one or two
      // bins inside each end: the first three bins from each end are tested together (their
      // grid loads in flight at once), the rest one by one only if those miss.
      const int l1 = min(a0 + 1, z0 - 1), l2 = min(a0 + 2, z0 - 1);
      const int r1 = max(z0 - 2, a0), r2 = max(z0 - 3, a0);
      const double* fq = gv.freq;
      const double fl0 = fq[a0], fl1 = fq[l1], fl2 = fq[l2], fr0 = fq[z0 - 1], fr1 = fq[r1], fr2 = fq[r2];
      auto inside = [&](double f) { return fabs(nu_if - f) * g + y < 15.0; };
      const bool tl0 = inside(fl0), tl1 = inside(fl1), tl2 = inside(fl2);
      const bool tr0 = inside(fr0), tr1 = inside(fr1), tr2 = inside(fr2);
      int a, z;
      if (tl0) a = a0;
      else if (a0 + 1 < z0 && tl1) a = a0 + 1;
      else if (a0 + 2 < z0 && tl2) a = a0 + 2;
      else {
        a = min(a0 + 3, z0);
        while (a < z0 && !in_core(nu_if, g, y, fq, a)) ++a;
      }
      if (tr0) z = z0;
      else if (z0 - 2 >= a0 && tr1) z = z0 - 1;
      else if (z0 - 3 >= a0 && tr2) z = z0 - 2;
      else {
        z = max(z0 - 3, a0);
        while (z > a0 && !in_core(nu_if, g, y, fq, z - 1)) --z;
      }
      if (z > a) {
        k0 = a;
        k1 = z;
      }
    }
  }
  out.core[(size_t)p * out.stride + i] = make_int2(k0, k1);
  return (unsigned long long)(k1 - k0);
}

// The temperature row tq of one line.
__device__ __forceinline__ void prep_mixed_store_t(const PrepT& t, const MixedLineData& out, int i, int tq) {
  const size_t o = (size_t)tq * out.stride + i;
  out.amp[o] = (float)(t.S * t.g);
  out.g[o] = t.g;
  out.ts[o] = t.tscale;
}

// One thread per line over all npts points: the line's record is read once, and the
// temperature-dependent part (two exp, one pow) is computed and stored once per run of
// consecutive points at one temperature (row tq counts the runs from 0, as
// mixed_temperature_rows counts them on the host).  Which points start a run is decided once
// per wave, a lane per point, into wave masks (points beyond kPrepMaskPts compare their
// constants in the loop).
constexpr int kPrepMasks = 4;
constexpr int kPrepMaskPts = 64 * kPrepMasks;

__global__ __launch_bounds__(256) void k_prep_mixed(const PointConsts* __restrict__ pcs, int npts, LinesView lv,
                                                    GridView gv, const int* __restrict__ IB,
                                                    const int* __restrict__ IE, MixedLineData out,
                                                    unsigned long long* __restrict__ core_evals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  unsigned long long starts[kPrepMasks];
#pragma unroll
  for (int c = 0; c < kPrepMasks; ++c) {
    const int q = 64 * c + lane;
    starts[c] = __ballot(q > 0 && q < npts && !same_t(pcs[q], pcs[q - 1]));
  }
  unsigned long long w = 0;
  if (i < lv.n) {
    const double nu = lv.nu[i];
    const double nu_if = nu == 0 ? 1e-6 : nu;
    const double a = lv.aif[i], elow = lv.elow[i], gamma = lv.gamma[i], nexp = lv.nexp[i];
    const int gi = lv.gns[i], ib = IB[i], ie = IE[i];
    const double inv_dfreq = 1.0 / gv.dfreq;
    PrepT t = prep_mixed_t(pcs[0], nu_if, a, gi, elow, nexp);
    int tq = 0;
    prep_mixed_store_t(t, out, i, 0);
    for (int p = 0; p < npts; ++p) {
      const bool run = p < kPrepMaskPts ? ((starts[p >> 6] >> (p & 63)) & 1ull) != 0
                                        : !same_t(pcs[p], pcs[p - 1]);
      if (run) {
        t = prep_mixed_t(pcs[p], nu_if, a, gi, elow, nexp);
        prep_mixed_store_t(t, out, i, ++tq);
      }
      w += prep_mixed_point(pcs[p], t, nu_if, gamma, ib, ie, gv, inv_dfreq, out, i, p);
    }
  }
  if (core_evals) block_count(w, core_evals);
}

// ---------------------------------------------------------------------------
// Slice of the sorted list that may touch bin block b: nu within the block's
// frequency span widened by the largest cutoff and two bins.  The exact
// per-line test happens in the accumulation kernels on integer windows.
// ---------------------------------------------------------------------------
__global__ void k_block_ranges(GridView gv, const double* __restrict__ nu, int n, double cut_max, int bpb, int nblocks,
                               int2* __restrict__ ranges) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int j0 = gv.start_idx + b * bpb;
  const int j1 = min(j0 + bpb, gv.start_idx + gv.points);
  const double lo_nu = gv.start_nu + (double)(j0 - 2) * gv.dfreq - cut_max;
  const double hi_nu = gv.start_nu + (double)(j1 + 1) * gv.dfreq + cut_max;
  int lo = 0, hi = n;
  while (lo < hi) {  // first nu >= lo_nu
    const int mid = lo + ((hi - lo) >> 1);
    if (nu[mid] < lo_nu) lo = mid + 1; else hi = mid;
  }
  int a = lo;
  hi = n;
  while (lo < hi) {  // first nu > hi_nu
    const int mid = lo + ((hi - lo) >> 1);
    if (nu[mid] <= hi_nu) lo = mid + 1; else hi = mid;
  }
  ranges[b] = make_int2(a, lo);
}

// ---------------------------------------------------------------------------
// FP64 accumulation: one bin per lane, 256 bins per workgroup.
// ---------------------------------------------------------------------------
// Sum of the lines [a, z) of the sorted list nu_s over bin j (frequency f) of one workgroup
// in line order; the lines' derived data sit at index (line - dofs).  Lines are staged
// through LDS in kTile-line tiles; every lane tests its own bin against each line's window.
struct F64Tiles {
  double nu[kTile], S[kTile], g[kTile], y[kTile];
  int ib[kTile], ie[kTile];
};

template <int PROFILE>
__device__ __forceinline__ double accum_f64_lines(F64Tiles& L, int j, double f, const double* __restrict__ nu_s,
                                                  int a, int z, int dofs, const double* __restrict__ S,
                                                  const double* __restrict__ G, const double* __restrict__ Y,
                                                  const int* __restrict__ IB, const int* __restrict__ IE,
                                                  double zero_nu, double x0) {
  const int t = threadIdx.x;
  double acc = 0.0;
  for (int base = a; base < z; base += kTile) {
    const int i = base + t;
    if (i < z) {
      const double v = nu_s[i];
      L.nu[t] = v == 0.0 ? zero_nu : v;
      L.S[t] = S[i - dofs];
      L.g[t] = G[i - dofs];
      if (PROFILE == kProfVoigt || PROFILE == kProfVoigtNoGamma) L.y[t] = Y[i - dofs];
      L.ib[t] = IB[i - dofs];
      L.ie[t] = IE[i - dofs];
    }
    __syncthreads();
    const int cnt = min(kTile, z - base);
    for (int k = 0; k < cnt; ++k) {
      const int ib = L.ib[k], ie = L.ie[k];
      if (j >= ib && j < ie) {
        const double g = L.g[k];
        if (PROFILE == kProfVoigt) {
          const double x = fabs(L.nu[k] - f) * g;  // OpenMPManager.cpp:211-214
          acc += L.S[k] * w4(x, L.y[k]) * g;
        } else if (PROFILE == kProfGaussian) {
          acc += plain_gauss_term(L.S[k], f - L.nu[k], g, x0);       // cross_kernal.cu:594-604
        } else if (PROFILE == kProfVoigtNoGamma) {
          acc += plain_nogamma_term(L.S[k], L.nu[k] - f, g, L.y[k]);  // cross_kernal.cu:884-904
        } else {
          const double d = fabs(L.nu[k] - f);      // OpenMPManager.cpp:165-171
          const double xm = g * (d - x0);
          if (xm < 6.0) {                           // beyond, erf(xp) == erf(xm) == 1 exactly
            const double xp = g * (d + x0);
            acc += L.S[k] * (erf(xp) - erf(xm));
          }
        }
      }
    }
    __syncthreads();
  }
  return acc;
}

template <int PROFILE>
__global__ __launch_bounds__(kBlock) void k_accum_f64(GridView gv, const double* __restrict__ nu_s,
                                                      const double* __restrict__ S, const double* __restrict__ G,
                                                      const double* __restrict__ Y, const int* __restrict__ IB,
                                                      const int* __restrict__ IE, const int2* __restrict__ ranges,
                                                      int nblocks, double* __restrict__ intens, double zero_nu,
                                                      double x0) {
  __shared__ F64Tiles L;
  const int b = blockIdx.x;
  const int j = gv.start_idx + b * kBinsF64 + threadIdx.x;
  const bool active = j < gv.start_idx + gv.points;
  const double f = active ? gv.freq[j] : 0.0;
  const int2 r = ranges[b];
  const double acc = accum_f64_lines<PROFILE>(L, j, f, nu_s, r.x, r.y, 0, S, G, Y, IB, IE, zero_nu, x0);
  if (active) intens[j] += acc;
}

// ---------------------------------------------------------------------------
// Batched FP64 path: several requests, each with its own grid, constants and line slice,
// in three launches.  Prep: one thread per line of every request's slice (each request's
// threads start on a workgroup boundary, so its windows are counted into its own counters).
// Accumulation: one workgroup per item (request, bin block, chunk of lines), writing its
// partial sums; reduction: per bin, the partials of its block in item order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_line_prep_batch(const F64Task* __restrict__ tasks, int ntasks, LinesView all) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int lo = 0, hi = ntasks - 1;                     // last task with prep_first <= this workgroup's first thread
  const long long t0 = (long long)blockIdx.x * blockDim.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].prep_first <= t0) lo = mid; else hi = mid - 1;
  }
  const F64Task& tk = tasks[lo];
  const long long i = t - tk.prep_first;
  unsigned long long w = 0;
  if (i >= 0 && i < tk.nlines) {
    const LinesView lv{all.nu + tk.line_lo, all.elow + tk.line_lo, all.aif + tk.line_lo, all.gns + tk.line_lo,
                       all.gamma ? all.gamma + tk.line_lo : nullptr, all.nexp ? all.nexp + tk.line_lo : nullptr,
                       tk.nlines};
    w = line_prep_one(tk.pc, tk.gv, lv, tk.d, (int)i);
  }
  block_count(w, tasks[lo].evals);
}

__global__ __launch_bounds__(kBlock) void k_accum_f64_items(const F64Task* __restrict__ tasks,
                                                            const F64Item* __restrict__ items,
                                                            double* __restrict__ partial) {
  __shared__ F64Tiles L;
  const F64Item it = items[blockIdx.x];
  const F64Task& tk = tasks[it.task];
  const GridView& gv = tk.gv;
  const int j = gv.start_idx + it.block * kBinsF64 + threadIdx.x;
  const bool active = j < gv.start_idx + gv.points;
  const double f = active ? gv.freq[j] : 0.0;
  const DerivedF64& d = tk.d;
  double acc;
  switch (tk.pc.profile) {
    case kProfVoigt:
      acc = accum_f64_lines<kProfVoigt>(L, j, f, tk.nu, it.a, it.z, tk.line_lo, d.S, d.g, d.y, d.ib, d.ie,
                                        tk.zero_nu, tk.x0);
      break;
    case kProfDoppler:
      acc = accum_f64_lines<kProfDoppler>(L, j, f, tk.nu, it.a, it.z, tk.line_lo, d.S, d.g, d.y, d.ib, d.ie,
                                          tk.zero_nu, tk.x0);
      break;
    case kProfGaussian:
      acc = accum_f64_lines<kProfGaussian>(L, j, f, tk.nu, it.a, it.z, tk.line_lo, d.S, d.g, d.y, d.ib, d.ie,
                                           tk.zero_nu, tk.x0);
      break;
    default:
      acc = accum_f64_lines<kProfVoigtNoGamma>(L, j, f, tk.nu, it.a, it.z, tk.line_lo, d.S, d.g, d.y, d.ib, d.ie,
                                               tk.zero_nu, tk.x0);
      break;
  }
  partial[(size_t)blockIdx.x * kBinsF64 + threadIdx.x] = acc;
}

__global__ __launch_bounds__(kBinsF64) void k_reduce_f64_items(const F64Task* __restrict__ tasks, int ntasks,
                                                               const int2* __restrict__ block_items,
                                                               const double* __restrict__ partial) {
  const int gb = blockIdx.x;                       // block over all tasks
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].first_block <= gb) lo = mid; else hi = mid - 1;
  }
  const F64Task& tk = tasks[lo];
  const int b = gb - tk.first_block;
  const int j = tk.gv.start_idx + b * kBinsF64 + threadIdx.x;
  if (j >= tk.gv.start_idx + tk.gv.points) return;
  const int2 bi = block_items[gb];
  double sum = 0.0;
  for (int c = 0; c < bi.y; ++c) sum += partial[(size_t)(bi.x + c) * kBinsF64 + threadIdx.x];
  tk.out[j] += sum;
}

// ---------------------------------------------------------------------------
// Mixed-precision Voigt accumulation (persistent, work-queue driven).
//
// Work item = (point, bin block, line segment); a block is the 64 R contiguous
// bins of ONE wave (lane l sums bins w0 + l + 64q, q < R), and the four waves of
// a workgroup take items independently: no workgroup barrier, so a wave whose
// bins see fewer lines never waits for another.  A wave streams its line
// segment in chunks of 64, one line per lane, staged FP64 -> FP32 into its own
// LDS area:
//   c      nearest bin to the centre (kept as an exact float integer)
//   r      (nu - f_{c}) / alpha      (|r| <= e/2)
//   e      dfreq / alpha
// so x_j = r - (j - c) e with an exact integer factor and the FP32 offset
// error stays relative to x itself; lines with small e use anchored offsets
// (one FMA per bin, below).  Region I of W4 runs in FP32 (region1()), and bins
// with |x| + y < 15 (the core, fixed per line by k_prep_mixed as an integer
// range, so the region choice equals the FP64 path's) are left to the core
// pass (k_core_mixed), which evaluates regions II-IV line by line.
// Window and core tests are scalar per (wave, line): a line whose window
// covers all of the wave's bins and whose core misses them runs the
// branch-free loop.  Per-bin sums: plain FP32 over 32 lines, then a compensated
// step into (sum, part) (kahan_step).  Segments of one block land in partial buffers and
// are added in segment order by k_reduce_segments, so the result does not
// depend on scheduling.
// ---------------------------------------------------------------------------
template <int R>
struct MixedShape {
  static constexpr int kBins = 64 * R;     // bins of one work item (one wave)
};

// Staged line record, 2 + ceil(R/4) float4 per line:
//   [0]      ib, ie, k0, k1 (ints: window and core ranges, absolute bins); anchored fast
//            lines keep x at the middles of slices 0-3 (those of the R slices) here instead
//            (they need no ranges); edge lines with R <= 2 keep the 64-bit lane mask of each
//            slice (bins inside the window and outside the core) as two ints per slice
//   [1]      e, h, k, a      e = dnu/alpha; h = y^2 - 1/2, k = 2 y^2, a = amp c y (scaled)
//   [2]      b, r, cf, s     b = a (scaled); r = x at the centre bin; cf: centre bin
//                            relative to the wave's first bin; s: slice codes of an edge
//                            line (int bits); anchored edge lines (R <= 2) keep x at the
//                            middles of slices 0 and 1 in r and cf
//   [3 ...]  anchored fast lines: x at the middles of slices 4-7 (R = 8)
// so the fast loop reads everything of a line from one LDS address.
template <int R>
struct StagedRec {
  static constexpr int kGroups = (R + 3) / 4;   // float4 groups of slice anchors
  static constexpr int kVec = 2 + kGroups;
};

// Anchored offsets: x at bin jb + o is A - o e, A = x(jb) rounded once from FP64 and o an
// exact small integer, one FMA per bin.  With jb the middle of each 64-bin slice, |o| <= 32,
// the error of x is eps (|x| + |A|) <= eps (2|x| + 32 e), and since region I has
// |x| + y >= 15 the relative error of the profile stays below eps (1 + 3e); the form is
// used for lines with e = dnu/alpha <= 8 (25 eps), the exact-integer form elsewhere.
constexpr double kAnchorMaxE = 8.0;

// Slice codes of an edge line (2 bits per 64-bin slice of the wave's bins).
constexpr int kSliceMiss = 0;   // the window misses the slice
constexpr int kSliceFull = 1;   // the window covers the slice and the core misses it
constexpr int kSlicePart = 2;   // per-lane window and core test

// Bits [a, b) of a 64-bit lane mask, a and b clamped to [0, 64].
__device__ __forceinline__ unsigned long long lane_range(int a, int b) {
  a = min(max(a, 0), 64);
  b = min(max(b, 0), 64);
  if (b <= a) return 0ull;
  const unsigned long long below_b = b >= 64 ? ~0ull : (1ull << b) - 1ull;
  return below_b & ~((1ull << a) - 1ull);     // a < 64 here
}

// Kahan summation with the compensation kept in the running partial sum: part holds
// the terms added since the last step plus the rounding error of that step (Fast2Sum
// error, exact while |sum| >= |part|, as in Kahan's form; all terms here are >= 0), so
// the next partial sum starts from the correction and the value is sum + part.
__device__ __forceinline__ void kahan_step(float& sum, float& part) {
  const float tk = sum + part;
  part = part - (tk - sum);
  sum = tk;
}

// W4 region I, c y (1/2 + x^2 + y^2) / ((1/2 + y^2 - x^2)^2 + 4 x^2 y^2), rewritten with
// u = x^2 + y^2 - 1/2 as c y (u + 1) / (u^2 + 2 y^2): the same rational function (the
// quartic's constant term reduces to 2y^2), 4 FMA + 1 reciprocal.  In region I
// |x| >= 15 - y, so u has no cancellation.
__device__ __forceinline__ float region1(float x, float h, float k, float a, float b) {
  const float u = __builtin_fmaf(x, x, h);
  const float den = __builtin_fmaf(u, u, k);
  const float num = __builtin_fmaf(a, u, b);
  return num * __builtin_amdgcn_rcpf(den);
}

// Interpolated lines (64-bin blocks): a line whose window covers the block, whose W4 core
// misses it and whose Lorentz width is at least kInterpG half-blocks is evaluated at the 16
// Chebyshev-Lobatto nodes of the block's frequency span instead of at its 64 bins, four lines
// side by side in a wave; after the item's last line the node sums are interpolated to the
// bins (barycentric Lobatto weights, tabulated by the host in P.interp_tab: the 64 x 16 matrix,
// then each node's offset).  The line's region-I function is rational in f with its poles
// gamma from the real axis, so the sum over such lines is analytic in the Bernstein ellipse of
// the block with rho = G + sqrt(G^2 + 1) = 3.3 at G = 1.5, and the interpolation error stays
// near rho^-16 ~ 5e-9 of the sum (2e-10 measured for one line at G = 2).
constexpr double kInterpG = 1.5;

__global__ void k_reset_counter(unsigned* c) { *c = 0u; }

template <int R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(R <= 4 ? 6 : 3))) void k_accum_voigt_mixed(MixedPlan P) {
  constexpr int kBins = MixedShape<R>::kBins;
  constexpr int kWaves = kBlock / 64;
  constexpr int kVec = StagedRec<R>::kVec;
  __shared__ float4 s_l[kWaves][64][kVec];

  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const float no = (float)(32 - lane);             // -(offset of this lane from its slice middle)
  const GridView gv = P.gv;
  const double inv_dfreq = 1.0 / gv.dfreq;
  const int jend = gv.start_idx + gv.points;
  const int per_task = P.nblocks * P.nseg;
  float4 (*my_l)[kVec] = s_l[wave];
  // bin offsets of this lane from the wave's first bin, exact in FP32
  float lanef_q[R];
#pragma unroll
  for (int q = 0; q < R; ++q) lanef_q[q] = (float)(lane + 64 * q);
  // interpolated lines: lane l evaluates node l & 15 of line l >> 4 of each group of four
  const bool interp = R == 1 && P.interp_tab != nullptr;
  const float nk = interp ? P.interp_tab[kInterpTabNodes + (lane & 15)] : 0.0f;

  int pass = 0;
  for (;;) {
    int item;
    if (P.queue) {                                 // persistent: take items from the queue
      unsigned v = 0;
      if (lane == 0) v = atomicAdd(P.counter, 1u);
      item = __builtin_amdgcn_readfirstlane((int)v);
    } else {                                       // one item per wave
      if (pass++) break;
      item = blockIdx.x * kWaves + wave;
    }
    if (item >= P.nitems) break;
    const int task = item / per_task;
    const int rem = item - task * per_task;
    const int b = rem / P.nseg;
    const int seg = rem - b * P.nseg;
    const MixedTask tk = P.tasks[task];
    const int w0 = gv.start_idx + b * kBins;
    const int w1 = w0 + kBins;
    const int2 rg = P.ranges[b];
    const int len = rg.y - rg.x;
    const int lo = rg.x + (int)(((long long)len * seg) / P.nseg);
    const int hi = rg.x + (int)(((long long)len * (seg + 1)) / P.nseg);
    // lines whose sums this block takes from far-field moments (k_far_eval, k_far_edges)
    int4 fr = P.far ? P.far_rng[(size_t)task * P.nblocks + b] : make_int4(0, 0, 0, 0);
    int4 er = P.far ? P.edge_rng[(size_t)task * P.nblocks + b] : make_int4(0, 0, 0, 0);
    fr.x = __builtin_amdgcn_readfirstlane(fr.x);
    fr.y = __builtin_amdgcn_readfirstlane(fr.y);
    fr.z = __builtin_amdgcn_readfirstlane(fr.z);
    fr.w = __builtin_amdgcn_readfirstlane(fr.w);
    er.x = __builtin_amdgcn_readfirstlane(er.x);
    er.y = __builtin_amdgcn_readfirstlane(er.y);
    er.z = __builtin_amdgcn_readfirstlane(er.z);
    er.w = __builtin_amdgcn_readfirstlane(er.w);

    float sum[R], part[R];
#pragma unroll
    for (int q = 0; q < R; ++q) sum[q] = part[q] = 0.0f;
    int since_flush = 0;
    float isum = 0.0f, ipart = 0.0f;               // node sums of the interpolated lines
    int isteps = 0, ilines = 0;

    // Kahan step after every 32 lines: the lines of a class are added in blocks that
    // end where the 32-line count does, so the flush test runs once per block
    auto run_block = [&](int n, auto body) {
      for (int i = 0; i < n;) {
        const int m = min(n - i, 32 - since_flush);
        body(i, i + m);
        i += m;
        since_flush += m;
        if (since_flush == 32) {
          since_flush = 0;
#pragma unroll
          for (int q = 0; q < R; ++q) kahan_step(sum[q], part[q]);
        }
      }
    };

    for (int base = lo; base < hi; base += 64) {
      if (base >= er.x && base < er.y) base = er.y;
      if (base >= fr.x && base < fr.y) base = fr.y;
      if (base >= fr.z && base < fr.w) base = fr.w;
      if (base >= er.z && base < er.w) base = er.w;
      if (base >= hi) break;
      const int i = base + lane;
      bool fa = false, fx = false, ea = false, ed = false, fi = false;
      float4 rec[kVec];
      if (i < hi && !(i >= fr.x && i < fr.y) && !(i >= fr.z && i < fr.w) && !(i >= er.x && i < er.y) &&
          !(i >= er.z && i < er.w)) {
#pragma clang fp contract(off)
        const double v0 = P.nu[i];
        const double nu = v0 == 0.0 ? 1e-6 : v0;
        const double g = ldg(tk.g + i), y = mixed_y(tk, i, g);      // the FP64 path's bits
        const double cj = rint((nu - gv.start_nu) * inv_dfreq);     // nearest bin (r below is consistent)
        const double fc = gv.start_nu + cj * gv.dfreq;
        const int ib = P.ib[i], ie = P.ie[i];
        int2 core = ldg(tk.core + i);           // bins left to the core pass: the W4 core, or the
        if (P.near) {                           // near range of the line's leaf when it has one
          const int2 u = P.near_u[(size_t)task * P.far_nclus + i / kFarLeaf];
          if (u.y > u.x) core = u;
        }
        const float amp = ldg(tk.amp + i);
        // Power-of-two rescaling x' = sigma x keeps x'^4 far from FP32 overflow for lines
        // whose window spans huge x (tiny nu => huge 1/alpha); region I is invariant under
        // (A1, A2, R, P2) -> (A1 s^4, A2 s^2, R s^2, P2 s^4).  sigma = 1 for ordinary lines.
        const double xmax = (fabs((double)ib - cj) + fabs((double)ie - cj) + 2.0) * gv.dfreq * g;
        const int kexp = xmax > 65536.0 ? (int)ceil(log2(xmax / 65536.0)) : 0;
        const double sg = ldexp(1.0, -kexp), sg2 = sg * sg, sg4 = sg2 * sg2;
        // scaled: u' = x'^2 + (y^2 - 1/2) s^2, den' = u'^2 + 2 y^2 s^4, num' = a s^2 (u' + s^2)
        const double y2 = y * y;
        const double ac = (double)amp * 0.5641896 * y;
        rec[1] = make_float4((float)(gv.dfreq * g * sg), (float)((y2 - 0.5) * sg2),
                             (float)(2.0 * y2 * sg4), (float)(ac * sg2));
        rec[2] = make_float4((float)(ac * sg4), (float)((nu - fc) * g * sg),
                             (float)(cj - (double)w0), 0.0f);
        const bool an = gv.dfreq * g <= kAnchorMaxE;
        const bool touch = ie > w0 && ib < w1;
        const bool full = ib <= w0 && ie >= w1;
        const bool cor = core.y > core.x && core.y > w0 && core.x < w1;
        fa = full && !cor && an;
        fx = full && !cor && !an;
        fi = interp && fa && y >= kInterpG * 31.5 * gv.dfreq * g;   // gamma >= G half-blocks
        ed = touch && !(full && !cor);
        ea = R <= 2 && ed && an;                   // edge lines with slice-middle anchors
        ed = ed && !ea;
        if (fa) {
          // slice q of the wave has its middle at bin w0 + 64 q + 32
          const double ed_ = gv.dfreq * g * sg;
          const double x32 = (nu - fc) * g * sg + (cj - (double)(w0 + 32)) * ed_;
#pragma unroll
          for (int qg = 0; qg < StagedRec<R>::kGroups; ++qg) {
            const double b0 = x32 - (double)(256 * qg) * ed_;
            float4 h = make_float4((float)b0, 0.0f, 0.0f, 0.0f);     // the slices there are
            if (4 * qg + 1 < R) h.y = (float)(b0 - 64.0 * ed_);
            if (4 * qg + 2 < R) h.z = (float)(b0 - 128.0 * ed_);
            if (4 * qg + 3 < R) h.w = (float)(b0 - 192.0 * ed_);
            rec[qg == 0 ? 0 : 2 + qg] = h;
          }
        } else {
          // slice codes of an edge line, 2 bits per 64-bin slice
          int codes = 0;
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const int s0 = w0 + 64 * q, s1 = s0 + 64;
            const int slo = max(ib, s0), shi = min(ie, s1);
            const int c = slo >= shi ? kSliceMiss
                          : (slo == s0 && shi == s1 && (core.y <= core.x || core.y <= s0 || core.x >= s1))
                              ? kSliceFull
                              : kSlicePart;
            codes |= c << (2 * q);
          }
          if constexpr (R <= 2) {
            // per slice, the lanes inside the window and outside the core, as a lane mask
            unsigned long long mk[2] = {0ull, 0ull};
#pragma unroll
            for (int q = 0; q < R; ++q) {
              const int s0 = w0 + 64 * q;
              const unsigned long long cm = core.y > core.x ? lane_range(core.x - s0, core.y - s0) : 0ull;
              mk[q] = lane_range(ib - s0, ie - s0) & ~cm;
            }
            rec[0] = make_float4(__uint_as_float((unsigned)mk[0]), __uint_as_float((unsigned)(mk[0] >> 32)),
                                 __uint_as_float((unsigned)mk[1]), __uint_as_float((unsigned)(mk[1] >> 32)));
            if (ea) {
              // x at the middles of slices 0 and 1, as the anchored fast lines keep them
              const double ed_ = gv.dfreq * g * sg;
              const double x32 = (nu - fc) * g * sg + (cj - (double)(w0 + 32)) * ed_;
              rec[2].y = (float)x32;
              rec[2].z = (float)(x32 - 64.0 * ed_);
            }
          } else {
            rec[0] = make_float4(__int_as_float(ib), __int_as_float(ie), __int_as_float(core.x),
                                 __int_as_float(core.y));
          }
          rec[2].w = __int_as_float(codes);
        }
      }
      // Compact the chunk's records by class, in line order: anchored fast lines first, then
      // exact-offset fast lines, anchored edge lines (R <= 2), then the other edge lines;
      // lines that miss the wave are dropped.
      fa = fa && !fi;                              // interpolated lines last
      const unsigned long long ma = __ballot(fa), mx = __ballot(fx), mea = __ballot(ea), me = __ballot(ed);
      const unsigned long long mi = __ballot(fi);
      const int na = __popcll(ma), nx = __popcll(mx), nea = __popcll(mea), ne = __popcll(me), ni = __popcll(mi);
      // a chunk of one class (the common case: all its lines fast or all interpolated) keeps the
      // lines' order among themselves; others are sorted by class
      const unsigned long long mall = ma | mx | mea | me | mi;
      int pos;
      if (mall == ma || mall == mi || mall == mx || mall == mea || mall == me) {
        pos = __builtin_amdgcn_mbcnt_hi((unsigned)(mall >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mall, 0u));
      } else {
        const unsigned long long mine = fa ? ma : (fx ? mx : (ea ? mea : (ed ? me : mi)));
        pos = (fx ? na : 0) + (ea ? na + nx : 0) + (ed ? na + nx + nea : 0) + (fi ? na + nx + nea + ne : 0) +
              __builtin_amdgcn_mbcnt_hi((unsigned)(mine >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mine, 0u));
      }
      wave_lds_handover();                         // the previous chunk's reads are done
      if (fa || fx || ea || ed || fi) {
        my_l[pos][0] = rec[0];
        my_l[pos][1] = rec[1];
        my_l[pos][2] = rec[2];
        if (fa) {
#pragma unroll
          for (int v = 3; v < kVec; ++v) my_l[pos][v] = rec[v];
        }
      }
      wave_lds_handover();                         // staged records visible to every lane

      // Fast lines: branch-free FP32 region I for all R bins of every lane.
      run_block(na, [&](int a, int z) {
#pragma unroll 2
        for (int k = a; k < z; ++k) {
          const float4 E = my_l[k][1];
          const float bb = my_l[k][2].x;
#pragma unroll
          for (int qg = 0; qg < StagedRec<R>::kGroups; ++qg) {
            const float4 H = my_l[k][qg == 0 ? 0 : 2 + qg];
            const float hq[4] = {H.x, H.y, H.z, H.w};
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
              if (4 * qg + qq < R) part[4 * qg + qq] += region1(__builtin_fmaf(no, E.x, hq[qq]), E.y, E.z, E.w, bb);
          }
        }
      });
      run_block(nx, [&](int a, int z) {
        for (int k = na + a; k < na + z; ++k) {
          const float4 E = my_l[k][1];
          const float4 B = my_l[k][2];
#pragma unroll
          for (int q = 0; q < R; ++q)
            part[q] += region1(__builtin_fmaf(B.z - lanef_q[q], E.x, B.y), E.y, E.z, E.w, B.x);
        }
      });
      // edge lines: per 64-bin slice the staged slice code says whether the line misses it
      // (skipped), covers it clear of its core (no per-lane test) or covers part of it (the
      // per-lane window and core test); the core bins [k0, k1) belong to the core pass.  With
      // R <= 2 the test is the staged lane mask (one select), and the lines with
      // e <= kAnchorMaxE, a class of their own, form x from the slice-middle anchors (one FMA,
      // as the anchored fast lines do).
      auto edge_slice = [&](int q, int c, float x, const float4& E, const float4& B, const float4& Iv) {
        const float u = __builtin_fmaf(x, x, E.y);
        const float rden = __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, E.z));
        float num = __builtin_fmaf(E.w, u, B.x);
        if (c == kSlicePart) {
          if constexpr (R <= 2) {
            const unsigned lo = __builtin_amdgcn_readfirstlane(__float_as_uint(q == 0 ? Iv.x : Iv.z));
            const unsigned hi = __builtin_amdgcn_readfirstlane(__float_as_uint(q == 0 ? Iv.y : Iv.w));
            num = __builtin_amdgcn_inverse_ballot_w64(((unsigned long long)hi << 32) | lo) ? num : 0.0f;
          } else {
            const int jj = w0 + lane + 64 * q;
            const bool ok = jj >= __float_as_int(Iv.x) && jj < __float_as_int(Iv.y) &&
                            (jj < __float_as_int(Iv.z) || jj >= __float_as_int(Iv.w));
            num = ok ? num : 0.0f;
          }
        }
        part[q] = __builtin_fmaf(num, rden, part[q]);
      };
      if constexpr (R <= 2) {
        // the lane mask is zero on a slice the line misses and all ones on one it covers clear
        // of its core, so the loops take every slice without a branch and two lines at a time
        auto masked = [&](int q, float x, const float4& E, const float4& B, const float4& Iv) {
          const float u = __builtin_fmaf(x, x, E.y);
          const float rden = __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, E.z));
          const unsigned lo = __builtin_amdgcn_readfirstlane(__float_as_uint(q == 0 ? Iv.x : Iv.z));
          const unsigned hi = __builtin_amdgcn_readfirstlane(__float_as_uint(q == 0 ? Iv.y : Iv.w));
          const float num = __builtin_amdgcn_inverse_ballot_w64(((unsigned long long)hi << 32) | lo)
                                ? __builtin_fmaf(E.w, u, B.x)
                                : 0.0f;
          part[q] = __builtin_fmaf(num, rden, part[q]);
        };
        run_block(nea, [&](int a, int z) {
#pragma unroll 2
          for (int k = na + nx + a; k < na + nx + z; ++k) {
            const float4 Iv = my_l[k][0];
            const float4 E = my_l[k][1];
            const float4 B = my_l[k][2];
#pragma unroll
            for (int q = 0; q < R; ++q) masked(q, __builtin_fmaf(no, E.x, q == 0 ? B.y : B.z), E, B, Iv);
          }
        });
        run_block(ne, [&](int a, int z) {
#pragma unroll 2
          for (int k = na + nx + nea + a; k < na + nx + nea + z; ++k) {
            const float4 Iv = my_l[k][0];
            const float4 E = my_l[k][1];
            const float4 B = my_l[k][2];
#pragma unroll
            for (int q = 0; q < R; ++q) masked(q, __builtin_fmaf(B.z - lanef_q[q], E.x, B.y), E, B, Iv);
          }
        });
      } else {
        run_block(ne, [&](int a, int z) {
          for (int k = na + nx + nea + a; k < na + nx + nea + z; ++k) {
            const float4 Iv = my_l[k][0];
            const float4 E = my_l[k][1];
            const float4 B = my_l[k][2];
            const int codes = __builtin_amdgcn_readfirstlane(__float_as_int(B.w));
#pragma unroll
            for (int q = 0; q < R; ++q) {
              const int c = (codes >> (2 * q)) & 3;
              if (c != kSliceMiss) edge_slice(q, c, __builtin_fmaf(B.z - lanef_q[q], E.x, B.y), E, B, Iv);
            }
          }
        });
      }
      if constexpr (R == 1) {
        // interpolated lines, four side by side: lane l at node l & 15 of line k + (l >> 4)
        const int i0 = na + nx + nea + ne;
        for (int k = i0; k < i0 + ni; k += 4) {
          const int ln = k + (lane >> 4);
          if (ln < i0 + ni) {
            const float4 Ek = my_l[ln][1];
            const float bk = my_l[ln][2].x, hk = my_l[ln][0].x;
            ipart += region1(__builtin_fmaf(nk, Ek.x, hk), Ek.y, Ek.z, Ek.w, bk);
          }
          if (++isteps == 32) {
            isteps = 0;
            kahan_step(isum, ipart);
          }
        }
        ilines += ni;
      }
    }
    if (R == 1 && ilines > 0) {
      // node sums S_k (the four line groups added), then bin lane = sum_k A[lane][k] S_k
      kahan_step(isum, ipart);
      float v = isum + ipart;
      v += __shfl_xor(v, 16);
      v += __shfl_xor(v, 32);
      const float* arow = P.interp_tab + lane * kInterpNodes;
      float p = 0.0f;
#pragma unroll
      for (int k = 0; k < kInterpNodes; ++k) p = __builtin_fmaf(arow[k], __shfl(v, k), p);
      part[0] += p;
      if (lane == 0 && P.interp_cnt) atomicAdd(&P.interp_cnt[item & (kEvalSlots - 1)], (unsigned long long)ilines);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) kahan_step(sum[q], part[q]);
    if (P.nseg == 1) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int jj = w0 + lane + 64 * q;
        if (jj < jend) tk.out[jj] += (double)sum[q] + (double)part[q];
      }
    } else {
      float2* dst = P.partial + (size_t)item * kBins;
#pragma unroll
      for (int q = 0; q < R; ++q) dst[lane + 64 * q] = make_float2(sum[q], -part[q]);
    }
  }
}

// out[j] += sum over segments, in segment order, of (sum - comp); the region-I kernel
// writes comp = -part
__global__ void k_reduce_segments(MixedPlan P, int kBinsMixed) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per_task = (long long)P.nblocks * kBinsMixed;
  if (idx >= per_task * P.ntasks) return;
  const int task = (int)(idx / per_task);
  const int off = (int)(idx - task * per_task);
  const int b = off / kBinsMixed, m = off - b * kBinsMixed;
  const int j = P.gv.start_idx + off;
  if (j >= P.gv.start_idx + P.gv.points) return;
  const size_t item0 = ((size_t)task * P.nblocks + b) * P.nseg;
  double acc = 0.0;
  for (int sg = 0; sg < P.nseg; ++sg) {
    const float2 v = P.partial[(item0 + sg) * kBinsMixed + m];
    acc += (double)v.x - (double)v.y;
  }
  P.tasks[task].out[j] += acc;
}

// ---------------------------------------------------------------------------
// W4 core pass: regions II-IV of every (line, bin) pair with bin in the line's
// core [k0, k1), added to the grids after the region-I kernel.
//
// Lines are taken in nu order, 64 RL per wave, RL consecutive lines per lane
// (RL = 4 where lines are dense, else 1); the wave walks the offsets
// k = bin - base bin together, so all lanes sit at about the same x and take
// the same W4 region (the branches stay coherent).  A lane's value at offset k
// is the sum over its lines at bin c + k, c its base bin (its first line's
// centre bin).  Lanes with equal c form contiguous segments (nu is sorted); a
// segment of one lane adds its value straight into the workgroup's row of bin
// sums, a longer one is first summed across lanes by a fixed DPP tree.  A
// workgroup (one wave) owns the
// lines of P.cpw consecutive centre bins and a row of FP64 sums over
// the bins they can reach; k_core_gather adds the rows covering a bin in row
// order.  Every sum is in a fixed order: the grids are deterministic.
// ---------------------------------------------------------------------------

// first line whose centre bin is >= cbase + c, for c in [0, ncent]
__global__ void k_centre_starts(GridView gv, const double* __restrict__ nu, int n, int cbase, int ncent,
                                int* __restrict__ starts) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncent) return;
  const int target = cbase + c;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    const double v = nu[mid];
    if (centre_bin(v == 0.0 ? 1e-6 : v, gv) < target) lo = mid + 1; else hi = mid;
  }
  starts[c] = lo;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Segmented inclusive sum over the wave in a fixed order: v += m * v(lane - 1, 2, 4, 8) inside
// rows of 16 (row_shr, 0 from outside the row), then += m15 * v(lane 15 of the row before) on
// rows 1 and 3 and += m31 * v(lane 31) on rows 2 and 3 (row_bcast; the other rows are left
// as they are).  The m are 0 or 1 per lane (is the source lane in my segment), so each
// segment's last lane ends with its total.  Each step is one DPP-modified FMA; the s_nops
// give the two wait states a DPP read of a just-written VGPR needs (five before the first,
// which also covers a VALU write of EXEC).  Every lane of the wave must be active.
__device__ __forceinline__ float seg_scan(float v, float m1, float m2, float m4, float m8, float m15, float m31) {
  asm volatile(
      "s_nop 4\n\t"
      "v_fmac_f32_dpp %0, %0, %1 row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %2 row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %3 row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %4 row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %5 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %6 row_bcast:31 row_mask:0xc bank_mask:0xf"
      : "+v"(v)
      : "v"(m1), "v"(m2), "v"(m4), "v"(m8), "v"(m15), "v"(m31));
  return v;
}

// One line of a core chunk, held in registers for the offset loop.
struct CoreLine {
  double nu, g, y;
  float amp;
  int lo, hi;          // core offsets from the lane's base bin, inside the row; empty when lo >= hi
  W4Region2 r2;
};

// Lines per lane where a workgroup's lines are dense (at least kCoreDenseLines per centre
// bin on average): a lane's kCoreLanesLines consecutive lines then share their centre bin,
// or nearly, and the scan and the LDS add run once for all of them.
constexpr int kCoreLaneLines = 4;
constexpr int kCoreDenseLines = 2 * kCoreLaneLines;

// The chunks [l0, l1) of a core workgroup with RL lines per lane (lane L takes lines
// chunk + RL L + r).  The lane's base bin cb is its first line's centre bin; all its lines
// are evaluated at bin cb + k (each with its own x, so a line whose centre bin differs from cb
// only shifts its core range).  A line of a near leaf (one with a near range U, which holds
// the W4 cores of all its lines) is evaluated at every bin of U: region I where |x| + y >= 15
// (the FP64 decision, so exactly its bins outside its core), regions II-IV elsewhere; the far
// field and the region-I kernel take its other bins.  kFreqLds: grid values of the row staged
// in LDS, interleaved with the sums (entry i at s_core[2i], its frequency at s_core[2i + 1]);
// else the sums alone and the frequencies read from HBM.
template <bool kFreqLds, int RL>
__device__ __forceinline__ void core_chunks(const MixedPlan& P, const MixedTask& tk, double* s_core, int base,
                                            int l0, int l1) {
  constexpr int kStride = kFreqLds ? 2 : 1;
  const int lane = threadIdx.x;
  const GridView gv = P.gv;
  const int span = P.span;
  for (int chunk = l0; chunk < l1; chunk += 64 * RL) {
    const int first = chunk + RL * lane;
    const bool act = first < l1;
    CoreLine L[RL];
    int cb = 0x7fffffff - 64 + lane;            // lanes without lines: singleton segments out of reach
    int lo = 0x3fffffff, hi = -0x3fffffff;
#pragma unroll
    for (int r = 0; r < RL; ++r) {
      const int l = first + r;
      L[r].nu = 1e-6;
      L[r].g = 0.0;
      L[r].y = 15.0;
      L[r].amp = 0.0f;
      L[r].lo = 0x3fffffff;
      L[r].hi = -0x3fffffff;
      if (l < l1) {
        const double v0 = P.nu[l];
        L[r].nu = v0 == 0.0 ? 1e-6 : v0;
        if (r == 0) cb = centre_bin(L[r].nu, gv);
        L[r].g = ldg(tk.g + l);
        L[r].y = mixed_y(tk, l, L[r].g);
        L[r].amp = ldg(tk.amp + l);
        int2 core = ldg(tk.core + l);
        if (P.near) {                            // the near range of the line's leaf (it holds the
          const int2 u = P.near_u[(size_t)blockIdx.y * P.far_nclus + l / kFarLeaf];   // line's core)
          if (u.y > u.x) core = u;
        }
        if (core.y > core.x) {
          L[r].lo = max(core.x, base) - cb;
          L[r].hi = min(core.y, base + span) - cb;
        }
        lo = min(lo, L[r].lo);
        hi = max(hi, L[r].hi);
      }
    }
    if (!__ballot(hi > lo)) continue;
#pragma unroll
    for (int r = 0; r < RL; ++r) L[r].r2 = W4Region2(L[r].y);
    const int kmin = __builtin_amdgcn_readfirstlane(wave_min(lo));
    const int kend = __builtin_amdgcn_readfirstlane(wave_max(hi));
    // segments: lanes with equal base bins (contiguous, nu is sorted); each lane's distance
    // from its segment's first lane fixes which steps of the scan it takes
    const int cprev = __shfl_up(cb, 1), cnext = __shfl_down(cb, 1);
    const bool head = lane == 0 || cprev != cb;
    const bool tail = lane == 63 || cnext != cb;
    const unsigned long long upto = __ballot(head) & ((2ull << lane) - 1ull);
    const int hd = lane - (63 - __builtin_clzll(upto));
    const float m1 = hd >= 1 ? 1.0f : 0.0f, m2 = hd >= 2 ? 1.0f : 0.0f, m4 = hd >= 4 ? 1.0f : 0.0f;
    const float m8 = hd >= 8 ? 1.0f : 0.0f;
    const float m15 = hd >= (lane & 15) + 1 ? 1.0f : 0.0f;   // lane 16 row - 1 is in my segment
    const float m31 = hd >= lane - 31 ? 1.0f : 0.0f;          // lane 31 is in my segment
    // the lane's row entries for offset 0 (a segment's lanes share them; lanes without
    // lines never use theirs)
    const int rel = act ? cb - base : 0;
    double* const acc0 = s_core + kStride * rel;
    const double* const fq0 = kFreqLds ? acc0 + 1 : gv.freq + (base + rel);
    for (int k = kmin; k < kend; ++k) {
      float v = 0.0f;
      if (k >= lo && k < hi) {
        const double f = fq0[kStride * k];
#pragma unroll
        for (int r = 0; r < RL; ++r)
          if (k >= L[r].lo && k < L[r].hi) {
            const double x = fabs(L[r].nu - f) * L[r].g;
            v = __builtin_fmaf(L[r].amp, w4_near_mixed(x, L[r].y, L[r].r2), v);
          }
      }
      v = seg_scan(v, m1, m2, m4, m8, m15, m31);
      // a nonzero total has a line of the segment on at k, so entry cb + k - base is in the row
      if (tail && v != 0.0f) atomicAdd(acc0 + kStride * k, (double)v);   // one lane per address
    }
  }
}

// At most 128 VGPRs (four waves per SIMD; with 4 lines per lane the compiler would take 136
// and run three): the offset loop waits on the scan's DPP steps and the LDS, so the fourth
// wave pays (176 ms against 191 ms per 48-point table); tighter limits spill.
template <bool kFreqLds>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_core_mixed(MixedPlan P) {
  extern __shared__ double s_core[];
  constexpr int kStride = kFreqLds ? 2 : 1;
  const int lane = threadIdx.x;
  const int w = blockIdx.x;
  const MixedTask tk = P.tasks[blockIdx.y];
  const GridView gv = P.gv;
  const int span = P.span;
  const int c0 = P.cbase + w * P.cpw;
  const int base = c0 - P.kmax;                 // row entry i is bin base + i
  const int s0 = gv.start_idx, s1 = gv.start_idx + gv.points;
  for (int i = lane; i < span; i += 64) {
    s_core[kStride * i] = 0.0;
    if (kFreqLds) {
      const int jj = base + i;
      s_core[2 * i + 1] = (jj >= s0 && jj < s1) ? gv.freq[jj] : 0.0;
    }
  }
  __syncthreads();
  const int l0 = P.starts[w * P.cpw];
  const int l1 = P.starts[min((w + 1) * P.cpw, P.ncent)];
  if (P.near && P.core_evals) {
    // the bins of near ranges beyond the lines' cores, counted apart from the offset loops
    // (the lines' data is then in cache for them)
    const int2* const near_u = P.near_u + (size_t)blockIdx.y * P.far_nclus;
    unsigned long long extra = 0;
    for (int l = l0 + lane; l < l1; l += 64) {
      const int2 u = near_u[l / kFarLeaf];
      if (u.y > u.x) {
        const int2 core = tk.core[l];
        extra += (unsigned long long)(u.y - u.x) - (unsigned long long)max(core.y - core.x, 0);
      }
    }
    extra = wave_sum_u64(extra);
    if (lane == 0 && extra) atomicAdd(&P.core_evals[blockIdx.x & (kEvalSlots - 1)], extra);
  }
  if (l1 - l0 >= kCoreDenseLines * P.cpw)
    core_chunks<kFreqLds, kCoreLaneLines>(P, tk, s_core, base, l0, l1);
  else
    core_chunks<kFreqLds, 1>(P, tk, s_core, base, l0, l1);
  __syncthreads();
  double* row = P.core_part + ((size_t)blockIdx.y * P.ncwg + w) * span;
  for (int i = lane; i < span; i += 64) row[i] = s_core[kStride * i];
}

// out[j] += the core rows covering bin j, in row order
__global__ void k_core_gather(MixedPlan P) {
  const int j = P.gv.start_idx + blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= P.gv.start_idx + P.gv.points) return;
  const int task = blockIdx.y;
  const int r = j - P.cbase + P.kmax;                       // >= 0
  const int whi = min(r / P.cpw, P.ncwg - 1);
  const int t = r - P.span;                                 // rows with w * cpw > t
  const int wlo = t < 0 ? 0 : t / P.cpw + 1;
  const double* rows = P.core_part + (size_t)task * P.ncwg * P.span;
  double acc = 0.0;
  for (int w = wlo; w <= whi; ++w) acc += rows[(size_t)w * P.span + (r - w * P.cpw)];
  P.tasks[task].out[j] += acc;
}

}  // namespace


hipError_t launch_iota(uint32_t* idx, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_iota<<<(n + 255) / 256, 256, 0, s>>>(idx, n);
  return hipGetLastError();
}

hipError_t sort_pairs_by_nu(void* temp, size_t& temp_bytes, const double* keys_in, double* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, int n, hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0,
                                   8 * sizeof(double), s);
}

hipError_t launch_gather_lines(const uint32_t* perm, const double* elow_in, const double* aif_in, const int* gns_in,
                               const double* gam_in, const double* nexp_in, double* elow, double* aif, int* gns,
                               double* gam, double* nexp, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_gather<<<(n + 255) / 256, 256, 0, s>>>(perm, elow_in, aif_in, gns_in, gam_in, nexp_in, elow, aif, gns, gam, nexp,
                                           n);
  return hipGetLastError();
}

hipError_t launch_line_prep_f64(const PointConsts& pc, const GridView& gv, const LinesView& lv, DerivedF64 out,
                                unsigned long long* evals, hipStream_t s) {
  if (lv.n <= 0) return hipSuccess;
  k_line_prep<<<(lv.n + 255) / 256, 256, 0, s>>>(pc, gv, lv, out, evals);
  return hipGetLastError();
}

hipError_t launch_block_ranges(const GridView& gv, const double* nu, int n, double cut_max, int bins_per_block,
                               int nblocks, int2* ranges, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  k_block_ranges<<<(nblocks + 255) / 256, 256, 0, s>>>(gv, nu, n, cut_max, bins_per_block, nblocks, ranges);
  return hipGetLastError();
}

hipError_t launch_accum_f64(const PointConsts& pc, const GridView& gv, const double* nu_sorted, DerivedF64 d,
                            const int2* ranges, int nblocks, double* intens, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  const bool zero_to_tiny = pc.profile == kProfVoigt || (pc.profile == kProfDoppler && pc.convention == 0);
  const double zero_nu = zero_to_tiny ? 1e-6 : 0.0;
  if (pc.profile == kProfGaussian)
    k_accum_f64<kProfGaussian><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks,
                                                          intens, zero_nu, pc.dpwcoef * gv.dfreq * 0.5);
  else if (pc.profile == kProfVoigtNoGamma)
    k_accum_f64<kProfVoigtNoGamma><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges,
                                                              nblocks, intens, zero_nu, 0.0);
  else if (pc.profile == 2)
    k_accum_f64<2><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks, intens,
                                              zero_nu, 0.0);
  else
    k_accum_f64<1><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks, intens,
                                              zero_nu, gv.dfreq * 0.5);
  return hipGetLastError();
}

hipError_t launch_f64_batch(const F64Task* d_tasks, int ntasks, long long prep_threads, const LinesView& all,
                            const F64Item* d_items, int nitems, int nblocks_all, const int2* d_block_items,
                            double* partial, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipError_t e;
  if (prep_threads > 0) {
    k_line_prep_batch<<<(unsigned)((prep_threads + 255) / 256), 256, 0, s>>>(d_tasks, ntasks, all);
    if ((e = hipGetLastError())) return e;
  }
  if (nitems > 0) {
    k_accum_f64_items<<<nitems, kBlock, 0, s>>>(d_tasks, d_items, partial);
    if ((e = hipGetLastError())) return e;
  }
  if (nblocks_all > 0) {
    k_reduce_f64_items<<<nblocks_all, kBinsF64, 0, s>>>(d_tasks, ntasks, d_block_items, partial);
    if ((e = hipGetLastError())) return e;
  }
  return hipSuccess;
}

double f64_zero_nu(const PointConsts& pc) {
  const bool zero_to_tiny = pc.profile == kProfVoigt || (pc.profile == kProfDoppler && pc.convention == 0);
  return zero_to_tiny ? 1e-6 : 0.0;
}

double f64_x0(const PointConsts& pc, double dfreq) {
  if (pc.profile == kProfGaussian) return pc.dpwcoef * dfreq * 0.5;
  if (pc.profile == kProfDoppler) return dfreq * 0.5;
  return 0.0;
}

hipError_t launch_windows(const WindowArgs& wa, const GridView& gv, const double* nu, int n, int* ib, int* ie,
                          unsigned long long* evals, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_windows<<<(n + 255) / 256, 256, 0, s>>>(wa, gv, nu, n, ib, ie, evals);
  return hipGetLastError();
}

hipError_t launch_prep_mixed(const PointConsts* d_pcs, int npts, const LinesView& lv, const GridView& gv,
                             const int* ib, const int* ie, MixedLineData out, unsigned long long* core_evals,
                             hipStream_t s) {
  if (lv.n <= 0 || npts <= 0) return hipSuccess;
  k_prep_mixed<<<(lv.n + 255) / 256, 256, 0, s>>>(d_pcs, npts, lv, gv, ib, ie, out, core_evals);
  return hipGetLastError();
}

size_t core_lds_bytes(int span) {
  if (span <= 0) return 0;
  if ((size_t)span * 16 <= 65536) return (size_t)span * 16;   // row sums + grid values
  if ((size_t)span * 8 <= 65536) return (size_t)span * 8;     // row sums; grid values from HBM
  return 0;
}

hipError_t launch_grid(double* f, int n, double start, double dfreq, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_grid<<<(n + 255) / 256, 256, 0, s>>>(f, n, start, dfreq);
  return hipGetLastError();
}

hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 1 << 20);
  k_fill<<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

int mixed_bins_per_block(int R) { return 64 * R; }

hipError_t launch_accum_mixed(const MixedPlan& p, int R, hipStream_t s) {
  if (p.nitems <= 0) return hipSuccess;
  k_reset_counter<<<1, 1, 0, s>>>(p.counter);
  switch (R) {
    case 1: k_accum_voigt_mixed<1><<<p.grid, kBlock, 0, s>>>(p); break;
    case 2: k_accum_voigt_mixed<2><<<p.grid, kBlock, 0, s>>>(p); break;
    case 4: k_accum_voigt_mixed<4><<<p.grid, kBlock, 0, s>>>(p); break;
    default: k_accum_voigt_mixed<8><<<p.grid, kBlock, 0, s>>>(p); break;
  }
  hipError_t e = hipGetLastError();
  if (e) return e;
  if (p.nseg > 1) {
    const int bpb = mixed_bins_per_block(R);
    const long long n = (long long)p.nblocks * bpb * p.ntasks;
    k_reduce_segments<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, bpb);
    if ((e = hipGetLastError())) return e;
  }
  return hipSuccess;
}

hipError_t launch_core_mixed(const MixedPlan& p, hipStream_t s) {
  if (p.nitems <= 0 || p.nlines <= 0 || p.ncwg <= 0 || p.gv.points <= 0) return hipSuccess;
  hipError_t e;
  k_centre_starts<<<(p.ncent + 1 + 255) / 256, 256, 0, s>>>(p.gv, p.nu, p.nlines, p.cbase, p.ncent, p.starts);
  if ((e = hipGetLastError())) return e;
  const size_t lds = core_lds_bytes(p.span);
  if (lds == 0) return hipErrorInvalidValue;
  if (lds == (size_t)p.span * 16)
    k_core_mixed<true><<<dim3(p.ncwg, p.ntasks), 64, lds, s>>>(p);
  else
    k_core_mixed<false><<<dim3(p.ncwg, p.ntasks), 64, lds, s>>>(p);
  if ((e = hipGetLastError())) return e;
  k_core_gather<<<dim3((p.gv.points + 255) / 256, p.ntasks), 256, 0, s>>>(p);
  return hipGetLastError();
}

int mixed_occupancy(int R) {
  int n = 0;
  hipError_t e;
  switch (R) {
    case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_accum_voigt_mixed<1>, kBlock, 0); break;
    case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_accum_voigt_mixed<2>, kBlock, 0); break;
    case 4: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_accum_voigt_mixed<4>, kBlock, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_accum_voigt_mixed<8>, kBlock, 0); break;
  }
  if (e != hipSuccess) n = 1;
  return n > 0 ? n : 1;
}

}  // namespace gx
