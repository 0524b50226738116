// mi355x_manager.h -- single-MI355X compute backend with the reference's
// BaseManager interface, over the C ABI of libgx.so (include/gx.h).
//
// Drop-in: construct it where the reference constructs its HybridManager
// (VoigtProfile.cpp:20, DopplerProfile.cpp:19):
//     manager = new Mi355xManager(profile);
// The lifecycle, argument meaning and blocking behaviour follow GpuManager
// (GPUManager.cpp:5-189): TransferVectors and TransferResults wait for prior
// work, ExecuteCrossSection is asynchronous and adds into a persistent grid,
// ReadyForWork polls the stream.  Device errors print a message and exit(-1)
// as the reference's CheckCudaError does (cuda_utils.cu:8-17), unless an
// error handler is installed with set_error_handler().
#pragma once

#include <cstdio>
#include <cstdlib>

#include "../../include/gx.h"

#ifndef GX_BASE_MANAGER_HEADER
#define GX_BASE_MANAGER_HEADER "gx_base_manager.h"
#endif
#include GX_BASE_MANAGER_HEADER

class Mi355xManager final : public BaseManager {
 public:
  struct Options {
    int device = 0;
    size_t memory = 0;               // device budget in bytes; 0: device total less 1 GiB
    int convention = GX_CONV_OPENMP; // widths / windows / constants of the CPU manager
    int precision = GX_PREC_FP64;
    double cutoff = 0.0;             // <= 0: reference rule (Voigt min(25P,100), Doppler 100)
    double doppler_widths = 0.0;     // Doppler: > 0 per-line window k*alpha + dfreq
  };
  using ErrorHandler = void (*)(int status, const char* where, const char* message);

  explicit Mi355xManager(ProfileType pprofile) : Mi355xManager(pprofile, Options()) {}
  Mi355xManager(ProfileType pprofile, const Options& opt) : BaseManager(pprofile), opt_(opt) {
    check(gx_create(opt.device, (int)pprofile, opt.memory, &ctx_), "create");
  }
  ~Mi355xManager() { gx_destroy(ctx_); }

  static void set_error_handler(ErrorHandler h) { handler() = h; }

  void InitializeVectors(int Npoints) override {
    size_t n = 0;
    check(gx_initialize_vectors(ctx_, Npoints, &n), "InitializeVectors");
    N_trans = n;
  }
  void InitializeConstants(double half_width, double temperature, double partition, double dfreq,
                           double meanmass, double pressure = 10.0, double ref_temp = 296.0) override {
    gx_consts c{};
    c.half_width = half_width;
    c.temperature = temperature;
    c.partition = partition;
    c.dfreq = dfreq;
    c.mean_mass = meanmass;
    c.pressure = pressure;
    c.ref_temp = ref_temp;
    c.cutoff = opt_.cutoff;
    c.doppler_widths = opt_.doppler_widths;
    c.convention = opt_.convention;
    c.precision = opt_.precision;
    check(gx_initialize_constants(ctx_, &c), "InitializeConstants");
  }
  void TransferVectors(size_t Nener, double* h_energies, double* h_nu, double* h_aif, int* h_gns,
                       double* h_gamma = NULL, double* h_n = NULL) override {
    check(gx_transfer_vectors(ctx_, Nener, h_energies, h_nu, h_aif, h_gns, h_gamma, h_n), "TransferVectors");
  }
  void TransferFreq(double* h_freq, double* h_intens, int N) override {
    check(gx_transfer_freq(ctx_, h_freq, h_intens, N), "TransferFreq");
  }
  void ExecuteCrossSection(int N, int N_ener, int start_idx) override {
    check(gx_execute(ctx_, N, N_ener, start_idx), "ExecuteCrossSection");
  }
  void TransferResults(double* h_freq, double* h_intens, int N) override {
    check(gx_transfer_results(ctx_, h_freq, h_intens, N), "TransferResults");
  }
  void Cleanup() override { check(gx_cleanup(ctx_), "Cleanup"); }
  bool ReadyForWork() override {
    const int r = gx_ready(ctx_);
    if (r < 0) check(-r, "ReadyForWork");
    return r == 1;
  }

  gx_ctx* context() { return ctx_; }

 private:
  gx_ctx* ctx_ = nullptr;
  Options opt_;

  static ErrorHandler& handler() {
    static ErrorHandler h = nullptr;
    return h;
  }
  static void check(int status, const char* where) {
    if (status == GX_OK) return;
    if (handler()) {
      handler()(status, where, gx_last_error());
      return;
    }
    std::printf("[%s] MI355X error: %s\n", where, gx_last_error());
    std::exit(-1);
  }
};
