// ref_harness.cpp -- drives the reference's own CPU manager (OpenMPManager,
// compiled unmodified from the reference sources by oracle/Makefile) through
// the BaseManager lifecycle, exactly as BaseProfile/VoigtProfile/DopplerProfile
// do, and writes the resulting grid as raw FP64.
//
// TEST INFRASTRUCTURE ONLY: produces the golden vectors under tests/golden/.
// Never built into or loaded by the product library.
//
// Lifecycle followed (reference files):
//   VoigtProfile.cpp:20-26 / DopplerProfile.cpp:19-25  InitializeVectors, TransferFreq, InitializeConstants
//   VoigtProfile.cpp:60-118 / DopplerProfile.cpp:47-99 range filter and chunk flush
//   BaseProfile.cpp:49-62                             TransferVectors, chunk ib/ie, ExecuteCrossSection
//   BaseProfile.cpp:65-67, HybridManager.cpp:58-80    TransferResults into a zeroed host grid
//
// Input file (little-endian, written by tests/golden/make_golden.py):
//   int32  profile, nthreads, npoints, max_points, capacity
//   double start_nu, end_nu, temperature, partition, mean_mass, pressure, ref_temp
//   int64  nlines
//   double energies[n], nu[n], aif[n]; int32 gns[n]; double gamma[n], nexp[n]  (last two Voigt only)
// Output file: double intens[npoints].  stderr: "manager_seconds <s>", the wall time from the
// first TransferVectors to TransferResults (the CPU baseline timing; file I/O excluded).
#include <malloc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "OpenMPManager.h"

namespace {

template <class T>
bool read_n(FILE* f, T* dst, size_t n) { return fread(dst, sizeof(T), n, f) == n; }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  // Every grid the reference allocates with new[] must come from a fresh,
  // zeroed mmap: OpenMPManager never initialises g_intens (OpenMPManager.cpp:23).
  mallopt(M_MMAP_THRESHOLD, 64 * 1024);

  FILE* in = fopen(argv[1], "rb");
  if (!in) { perror("input"); return 2; }
  int32_t hdr[5];
  double par[7];
  int64_t n = 0;
  if (!read_n(in, hdr, 5) || !read_n(in, par, 7) || !read_n(in, &n, 1)) { fprintf(stderr, "bad header\n"); return 2; }
  const int profile = hdr[0], nthreads = hdr[1], npoints = hdr[2], max_points = hdr[3], capacity = hdr[4];
  const double start_nu = par[0], end_nu = par[1], temperature = par[2], partition = par[3];
  const double mean_mass = par[4], pressure = par[5], ref_temp = par[6];
  if (npoints * sizeof(double) < 64 * 1024) { fprintf(stderr, "npoints too small for zeroed grids\n"); return 2; }
  std::vector<double> E(n), nu(n), A(n), gam(n, 0.0), nexp(n, 0.0);
  std::vector<int> g(n);
  bool ok = read_n(in, E.data(), n) && read_n(in, nu.data(), n) && read_n(in, A.data(), n) && read_n(in, g.data(), n);
  if (profile == VOIGT) ok = ok && read_n(in, gam.data(), n) && read_n(in, nexp.data(), n);
  fclose(in);
  if (!ok) { fprintf(stderr, "short input\n"); return 2; }

  // BaseProfile.cpp:7-21: the grid
  const double dfreq = (end_nu - start_nu) / double(npoints);
  std::vector<double> freq(npoints), intens(npoints, 0.0);
  for (int i = 0; i < npoints; i++) freq[i] = start_nu + double(i) * dfreq;

  // capacity lines of 44 (Voigt) / 28 (Doppler) bytes; OpenMPManager keeps half its budget
  const size_t per_line = profile == VOIGT ? 44 : 28;
  const size_t budget = 2 * (size_t(capacity) * per_line + size_t(npoints) * 8 * (2 + nthreads)) + 2 * per_line;
  OpenMPManager* mgr = new OpenMPManager(ProfileType(profile), nthreads, budget);
  mgr->InitializeVectors(npoints);
  mgr->TransferFreq(freq.data(), intens.data(), npoints);
  const int fit = int(std::min<size_t>(mgr->GetNtrans(), size_t(capacity)));
  mgr->InitializeConstants(0.0, temperature, partition, dfreq, mean_mass, pressure, ref_temp);

  const double hw = profile == VOIGT ? std::min(25.0 * pressure, 100.0) : 100.0;  // VoigtProfile.cpp:49, DopplerProfile.cpp:4
  std::vector<double> hE(fit), hnu(fit), hA(fit), hgam(fit), hn(fit);
  std::vector<int> hg(fit);
  int ntrans = 0;
  double min_nu = 99999.0, max_nu = 0.0;
  const auto t_start = std::chrono::steady_clock::now();
  auto perform = [&]() {  // BaseProfile.cpp:49-62
    mgr->TransferVectors(ntrans, hE.data(), hnu.data(), hA.data(), hg.data(),
                         profile == VOIGT ? hgam.data() : nullptr, profile == VOIGT ? hn.data() : nullptr);
    int ib = std::max(round((min_nu - hw - start_nu) / dfreq), 0.0);
    int ie = std::min(round((max_nu + hw - start_nu) / dfreq), double(npoints));
    mgr->ExecuteCrossSection(ie - ib, ntrans, ib);
  };
  for (int64_t i = 0; i < n; i++) {
    if (nu[i] < start_nu - hw) continue;
    if (nu[i] > end_nu + hw) break;
    hE[ntrans] = E[i]; hnu[ntrans] = nu[i]; hA[ntrans] = A[i]; hg[ntrans] = g[i];
    hgam[ntrans] = gam[i]; hn[ntrans] = nexp[i];
    min_nu = std::min(min_nu, hnu[ntrans]);
    max_nu = std::max(max_nu, hnu[ntrans]);
    int tib = std::max(round((min_nu - hw - start_nu) / dfreq), 0.0);
    int tie = std::min(round((max_nu + hw - start_nu) / dfreq), double(npoints));
    ntrans++;
    bool span = profile == VOIGT && (tie - tib) > max_points;  // VoigtProfile.cpp:90 (Doppler flushes on capacity only)
    if (ntrans >= fit || span) {
      perform();
      min_nu = 99999.0; max_nu = 0.0; ntrans = 0;
    }
  }
  if (ntrans > 0) perform();

  // HybridManager.cpp:58-73: worker result added into the zeroed host grid
  std::vector<double> tmp_freq(npoints), tmp(npoints, 0.0);
  mgr->TransferResults(tmp_freq.data(), tmp.data(), npoints);
  for (int i = 0; i < npoints; i++) intens[i] += tmp[i];
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  fprintf(stderr, "manager_seconds %.6f\n", secs);

  FILE* out = fopen(argv[2], "wb");
  if (!out) { perror("output"); return 2; }
  fwrite(intens.data(), sizeof(double), npoints, out);
  fclose(out);
  return 0;
}
