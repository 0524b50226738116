// gx_internal.h -- declarations shared by the kernels (gx_kernels.hip) and the
// host side of the library (gx_api.cpp, gx_table.cpp, gx_server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gx {

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
  int profile;            // 1 Doppler, 2 Voigt
  int convention;         // 0 openmp, 1 gpu
};

// Grid and chunk geometry.
struct GridView {
  const double* freq;  // all M nodes
  int npoints_total;   // M
  int start_idx;       // first bin of the chunk
  int points;          // bins in the chunk
  double start_nu;     // freq[0]
  double dfreq;
};

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

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kTile = 256;           // lines staged in LDS per pass

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
hipError_t launch_accum_mixed(const PointConsts& pc, const GridView& gv, const double* nu_sorted, DerivedF64 d,
                              const int2* ranges, int nblocks, double* intens, hipStream_t s);

// bins covered by one workgroup of the accumulation kernels
constexpr int kBinsF64 = 256;
int mixed_bins_per_block();

}  // namespace gx
