// drive_manager.cpp -- runs one cross-section through the C++ Mi355xManager
// with the reference's driver flow, from the same input file format that
// oracle/ref_harness.cpp feeds the reference's own CPU manager, so the two
// outputs compare bin for bin.
//
// Flow (reference files): grid BaseProfile.cpp:7-21; manager set-up
// VoigtProfile.cpp:20-26 / DopplerProfile.cpp:19-25; range filter and chunk
// flush VoigtProfile.cpp:60-118 / DopplerProfile.cpp:47-99; per chunk
// BaseProfile.cpp:49-62; results HybridManager.cpp:58-80.
//
// usage: drive_manager input.bin output.bin [precision(0|1)] [convention(0|1)]
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mi355x_manager.h"

template <class T>
static bool read_n(FILE* f, T* dst, size_t n) { return fread(dst, sizeof(T), n, f) == n; }

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s input.bin output.bin [precision] [convention]\n", argv[0]);
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) { perror("input"); return 2; }
  int32_t hdr[5];
  double par[7];
  int64_t n = 0;
  if (!read_n(in, hdr, 5) || !read_n(in, par, 7) || !read_n(in, &n, 1)) { fprintf(stderr, "bad header\n"); return 2; }
  const int profile = hdr[0], npoints = hdr[2], max_points = hdr[3], capacity = hdr[4];
  const double start_nu = par[0], end_nu = par[1], temperature = par[2], partition = par[3];
  const double mean_mass = par[4], pressure = par[5], ref_temp = par[6];
  std::vector<double> E(n), nu(n), A(n), gam(n, 0.0), nexp(n, 0.0);
  std::vector<int> g(n);
  bool ok = read_n(in, E.data(), n) && read_n(in, nu.data(), n) && read_n(in, A.data(), n) && read_n(in, g.data(), n);
  if (profile == VOIGT) ok = ok && read_n(in, gam.data(), n) && read_n(in, nexp.data(), n);
  fclose(in);
  if (!ok) { fprintf(stderr, "short input\n"); return 2; }

  const double dfreq = (end_nu - start_nu) / double(npoints);
  std::vector<double> freq(npoints), intens(npoints, 0.0);
  for (int i = 0; i < npoints; i++) freq[i] = start_nu + double(i) * dfreq;

  Mi355xManager::Options opt;
  if (argc > 3) opt.precision = std::atoi(argv[3]);
  if (argc > 4) opt.convention = std::atoi(argv[4]);
  BaseManager* mgr = new Mi355xManager(ProfileType(profile), opt);
  mgr->InitializeVectors(npoints);
  mgr->TransferFreq(freq.data(), intens.data(), npoints);
  const int fit = int(std::min<size_t>(mgr->GetNtrans(), size_t(capacity)));
  mgr->InitializeConstants(0.0, temperature, partition, dfreq, mean_mass, pressure, ref_temp);

  const double hw = profile == VOIGT ? std::min(25.0 * pressure, 100.0) : 100.0;
  std::vector<double> hE(fit), hnu(fit), hA(fit), hgam(fit), hn(fit);
  std::vector<int> hg(fit);
  int ntrans = 0;
  double min_nu = 99999.0, max_nu = 0.0;
  auto perform = [&]() {
    mgr->TransferVectors(ntrans, hE.data(), hnu.data(), hA.data(), hg.data(),
                         profile == VOIGT ? hgam.data() : nullptr, profile == VOIGT ? hn.data() : nullptr);
    const int ib = int(std::max(std::round((min_nu - hw - start_nu) / dfreq), 0.0));
    const int ie = int(std::min(std::round((max_nu + hw - start_nu) / dfreq), double(npoints)));
    mgr->ExecuteCrossSection(ie - ib, ntrans, ib);
  };
  for (int64_t i = 0; i < n; i++) {
    if (nu[i] < start_nu - hw) continue;
    if (nu[i] > end_nu + hw) break;
    hE[ntrans] = E[i]; hnu[ntrans] = nu[i]; hA[ntrans] = A[i]; hg[ntrans] = g[i];
    hgam[ntrans] = gam[i]; hn[ntrans] = nexp[i];
    min_nu = std::min(min_nu, hnu[ntrans]);
    max_nu = std::max(max_nu, hnu[ntrans]);
    const int tib = int(std::max(std::round((min_nu - hw - start_nu) / dfreq), 0.0));
    const int tie = int(std::min(std::round((max_nu + hw - start_nu) / dfreq), double(npoints)));
    ntrans++;
    const bool span = profile == VOIGT && (tie - tib) > max_points;
    if (ntrans >= fit || span) {
      perform();
      min_nu = 99999.0; max_nu = 0.0; ntrans = 0;
    }
  }
  if (ntrans > 0) perform();
  while (!mgr->ReadyForWork()) {
  }
  std::vector<double> tmp_freq(npoints), tmp(npoints, 0.0);
  mgr->TransferResults(tmp_freq.data(), tmp.data(), npoints);
  for (int i = 0; i < npoints; i++) intens[i] += tmp[i];
  mgr->Cleanup();
  delete static_cast<Mi355xManager*>(mgr);

  FILE* out = fopen(argv[2], "wb");
  if (!out) { perror("output"); return 2; }
  fwrite(intens.data(), sizeof(double), npoints, out);
  fclose(out);
  return 0;
}
