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
// Workgroups are mapped XCD-aware: hardware dispatches block b to XCD b % 8, and the remap
// gives each XCD one contiguous range of bin blocks so neighbouring blocks, which stream
// almost the same lines, share that XCD's L2.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "gx_internal.h"
#include "gx_w4.h"

namespace gx {

namespace {

constexpr double kIsqrtPi = 0.564189584;  // defines.h:8 (the truncated literal the reference uses)
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int xcd = bid % kXcds, idx = bid / kXcds;
  const int per = nblocks / kXcds, rem = nblocks % kXcds;
  return xcd * per + min(xcd, rem) + idx;
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

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

// ---------------------------------------------------------------------------
// Line prep (FP64).  Evaluation order of every expression follows the
// reference so that S, the Doppler width and the window edges round exactly
// as in OpenMPManager.cpp:191-206; contraction into FMA is switched off here.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_line_prep(PointConsts pc, GridView gv, LinesView lv, DerivedF64 out,
                                                   unsigned long long* __restrict__ evals) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long w = 0;
  if (i < lv.n) {
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
  if (evals) {
    w = wave_sum_u64(w);
    if ((threadIdx.x & 63) == 0 && w) atomicAdd(evals, w);
  }
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
template <int PROFILE>
__global__ __launch_bounds__(kBlock) void k_accum_f64(GridView gv, const double* __restrict__ nu_s,
                                                      const double* __restrict__ S, const double* __restrict__ G,
                                                      const double* __restrict__ Y, const int* __restrict__ IB,
                                                      const int* __restrict__ IE, const int2* __restrict__ ranges,
                                                      int nblocks, double* __restrict__ intens, double zero_nu,
                                                      double x0) {
  __shared__ double s_nu[kTile], s_S[kTile], s_g[kTile], s_y[kTile];
  __shared__ int s_ib[kTile], s_ie[kTile];
  const int t = threadIdx.x;
  const int b = xcd_remap(blockIdx.x, nblocks);
  const int j = gv.start_idx + b * kBinsF64 + t;
  const bool active = j < gv.start_idx + gv.points;
  const double f = active ? gv.freq[j] : 0.0;
  const int2 r = ranges[b];
  double acc = 0.0;
  for (int base = r.x; base < r.y; base += kTile) {
    const int i = base + t;
    if (i < r.y) {
      const double v = nu_s[i];
      s_nu[t] = v == 0.0 ? zero_nu : v;
      s_S[t] = S[i];
      s_g[t] = G[i];
      if (PROFILE == 2) s_y[t] = Y[i];
      s_ib[t] = IB[i];
      s_ie[t] = IE[i];
    }
    __syncthreads();
    const int cnt = min(kTile, r.y - base);
    for (int k = 0; k < cnt; ++k) {
      const int ib = s_ib[k], ie = s_ie[k];
      if (j >= ib && j < ie) {
        const double g = s_g[k];
        if (PROFILE == 2) {
          const double x = fabs(s_nu[k] - f) * g;  // OpenMPManager.cpp:211-214
          acc += s_S[k] * w4(x, s_y[k]) * g;
        } else {
          const double d = fabs(s_nu[k] - f);      // OpenMPManager.cpp:165-171
          const double xm = g * (d - x0);
          if (xm < 6.0) {                           // beyond, erf(xp) == erf(xm) == 1 exactly
            const double xp = g * (d + x0);
            acc += s_S[k] * (erf(xp) - erf(xm));
          }
        }
      }
    }
    __syncthreads();
  }
  if (active) intens[j] += acc;
}

// ---------------------------------------------------------------------------
// Mixed-precision Voigt accumulation.
//
// A workgroup owns kBinsMixed = 4 waves x 64 lanes x R bins; lane l of wave w
// sums bins w0 + l + 64q (q < R), w0 = j0 + 256 w.  Per staged line (FP64 -> LDS):
//   c      nearest bin to the centre, relative to the block origin j0
//   r      (nu - f_{j0+c}) / alpha   (|r| <= e/2, FP32)
//   e      dfreq / alpha             (FP32)
// so x_j = r - (j - j0 - c) e with an exact integer factor: the FP32 offset
// error stays relative to x itself (no large-number cancellation).
// Region I of W4 is evaluated in FP32 in the closed form
//   (A1 + A2 x^2) / (x^2 (x^2 + R) + P2),
//   A1 = amp c y (0.5+y^2), A2 = amp c y, R = 2y^2 - 1, P2 = (0.5+y^2)^2.
// Bins with |x| + y < 15 (the core, found per line in FP64 as an integer
// range so the region choice equals the FP64 path's) are evaluated by the FP64
// W4.  Window and core tests are wave-uniform per line: a wave whose 256 bins
// lie inside the window and outside the core runs the branch-free FP32 loop.
// Per-bin sums: plain FP32 over 32 lines, then a Kahan step into (sum, comp).
// ---------------------------------------------------------------------------
constexpr int kR = 4;
constexpr int kBinsMixed = kBlock * kR;

__global__ __launch_bounds__(kBlock) void k_accum_voigt_mixed(GridView gv, const double* __restrict__ nu_s,
                                                              const double* __restrict__ S,
                                                              const double* __restrict__ G,
                                                              const double* __restrict__ Y,
                                                              const int* __restrict__ IB, const int* __restrict__ IE,
                                                              const int2* __restrict__ ranges, int nblocks,
                                                              double* __restrict__ intens, double zero_nu) {
  __shared__ float s_r[kTile], s_e[kTile], s_a1[kTile], s_a2[kTile], s_rr[kTile], s_p2[kTile], s_amp[kTile];
  __shared__ int s_c[kTile], s_ib[kTile], s_ie[kTile], s_k0[kTile], s_k1[kTile];
  __shared__ double s_nu[kTile], s_g[kTile], s_y[kTile];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int b = xcd_remap(blockIdx.x, nblocks);
  const int j0 = gv.start_idx + b * kBinsMixed;
  const int jend = gv.start_idx + gv.points;
  const int w0 = j0 + wave * 256;
  const int2 r = ranges[b];

  float sum[kR], comp[kR], part[kR];
#pragma unroll
  for (int q = 0; q < kR; ++q) sum[q] = comp[q] = part[q] = 0.0f;

  for (int base = r.x; base < r.y; base += kTile) {
    const int i = base + t;
    if (i < r.y) {
#pragma clang fp contract(off)
      // FP64 staging of one line
      const double v0 = nu_s[i];
      const double nu = v0 == 0.0 ? zero_nu : v0;
      const double g = G[i], y = Y[i];
      const double amp = S[i] * g;
      const double pos = (nu - gv.start_nu) / gv.dfreq;  // bin coordinate of the centre
      const double cj = rint(pos);
      const double fc = gv.start_nu + cj * gv.dfreq;
      s_c[t] = (int)cj - j0;
      s_r[t] = (float)((nu - fc) * g);
      s_e[t] = (float)(gv.dfreq * g);
      const double y2 = y * y;
      s_a2[t] = (float)(amp * 0.5641896 * y);
      s_a1[t] = (float)(amp * 0.5641896 * y * (0.5 + y2));
      s_rr[t] = (float)(2.0 * y2 - 1.0);
      s_p2[t] = (float)((0.5 + y2) * (0.5 + y2));
      s_amp[t] = (float)amp;
      s_ib[t] = IB[i];
      s_ie[t] = IE[i];
      s_nu[t] = nu;
      s_g[t] = g;
      s_y[t] = y;
      // core bins: |nu - f_j| g + y < 15, as an exact integer range [k0, k1)
      int k0 = 0, k1 = 0;  // empty
      if (y < 15.0) {
        const double half = (15.0 - y) / g;
        int a = (int)floor((nu - half - gv.start_nu) / gv.dfreq) - 1;
        int bnd = (int)ceil((nu + half - gv.start_nu) / gv.dfreq) + 1;
        auto core = [&](int jj) {
          const double fj = gv.start_nu + (double)jj * gv.dfreq;
          return fabs(nu - fj) * g + y < 15.0;
        };
        while (a < bnd && !core(a)) ++a;
        while (bnd > a && !core(bnd - 1)) --bnd;
        k0 = a;
        k1 = bnd;
      }
      s_k0[t] = k0;
      s_k1[t] = k1;
    }
    __syncthreads();
    const int cnt = min(kTile, r.y - base);
    for (int k = 0; k < cnt; ++k) {
      const int ib = uniform_int(s_ib[k]), ie = uniform_int(s_ie[k]);
      if (ie > w0 && ib < w0 + 256) {
        const int c = uniform_int(s_c[k]);
        const int k0 = uniform_int(s_k0[k]), k1 = uniform_int(s_k1[k]);
        const float e = s_e[k], rr0 = s_r[k];
        const float e64 = 64.0f * e;
        // x for bin q of this lane: r - (w0 - j0 + lane + 64 q - c) e
        const float x0 = __builtin_fmaf(-(float)(w0 - j0 + lane - c), e, rr0);
        const float a1 = s_a1[k], a2 = s_a2[k], rr = s_rr[k], p2 = s_p2[k];
        const bool fast = ib <= w0 && ie >= w0 + 256 && (k0 >= k1 || k1 <= w0 || k0 >= w0 + 256);
        if (fast) {
#pragma unroll
          for (int q = 0; q < kR; ++q) {
            const float x = __builtin_fmaf(-(float)q, e64, x0);
            const float x2 = x * x;
            const float num = __builtin_fmaf(a2, x2, a1);
            const float den = __builtin_fmaf(x2, x2 + rr, p2);
            part[q] += num * __builtin_amdgcn_rcpf(den);
          }
        } else {
#pragma unroll
          for (int q = 0; q < kR; ++q) {
            const int jj = w0 + lane + 64 * q;
            if (jj >= ib && jj < ie) {
              if (jj >= k0 && jj < k1) {
                const double f = gv.freq[jj];
                const double x = fabs(s_nu[k] - f) * s_g[k];
                part[q] += s_amp[k] * (float)w4_inner(x, s_y[k]);
              } else {
                const float x = __builtin_fmaf(-(float)q, e64, x0);
                const float x2 = x * x;
                const float num = __builtin_fmaf(a2, x2, a1);
                const float den = __builtin_fmaf(x2, x2 + rr, p2);
                part[q] += num * __builtin_amdgcn_rcpf(den);
              }
            }
          }
        }
      }
      if ((k & 31) == 31) {  // compensated step every 32 lines
#pragma unroll
        for (int q = 0; q < kR; ++q) {
          const float yk = part[q] - comp[q];
          const float tk = sum[q] + yk;
          comp[q] = (tk - sum[q]) - yk;
          sum[q] = tk;
          part[q] = 0.0f;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kR; ++q) {
      const float yk = part[q] - comp[q];
      const float tk = sum[q] + yk;
      comp[q] = (tk - sum[q]) - yk;
      sum[q] = tk;
      part[q] = 0.0f;
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < kR; ++q) {
    const int jj = w0 + lane + 64 * q;
    if (jj < jend) intens[jj] += (double)sum[q] - (double)comp[q];
  }
}

}  // namespace

int mixed_bins_per_block() { return kBinsMixed; }

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
  const bool zero_to_tiny = pc.profile == 2 || pc.convention == 0;
  const double zero_nu = zero_to_tiny ? 1e-6 : 0.0;
  if (pc.profile == 2)
    k_accum_f64<2><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks, intens,
                                              zero_nu, 0.0);
  else
    k_accum_f64<1><<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks, intens,
                                              zero_nu, gv.dfreq * 0.5);
  return hipGetLastError();
}

hipError_t launch_accum_mixed(const PointConsts& pc, const GridView& gv, const double* nu_sorted, DerivedF64 d,
                              const int2* ranges, int nblocks, double* intens, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (pc.profile != 2) return hipErrorInvalidValue;
  k_accum_voigt_mixed<<<nblocks, kBlock, 0, s>>>(gv, nu_sorted, d.S, d.g, d.y, d.ib, d.ie, ranges, nblocks, intens,
                                                 1e-6);
  return hipGetLastError();
}

}  // namespace gx
