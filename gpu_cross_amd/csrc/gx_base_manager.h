// gx_base_manager.h -- the compute-backend interface the Mi355xManager plugs
// into, restated from its signatures (BaseManager.h:6-33 of the reference;
// ProfileType from Input.h:9-15).  A reference build that already has its own
// BaseManager.h compiles mi355x_manager.h against that one instead
// (define GX_BASE_MANAGER_HEADER to "BaseManager.h"); the virtual signatures
// are identical.
#pragma once

#include <cstddef>

enum ProfileType { GAUSSIAN, DOPPLER, VOIGT, PSUEDO_VOIGT };

class BaseManager {
 private:
  size_t available_memory = 0;
  size_t total_memory = 0;

 protected:
  size_t N_trans = 0;
  ProfileType profile;
  void InitializeMemory(size_t bytes) { available_memory = total_memory = bytes; }
  void TrackMemory(size_t bytes) { available_memory -= bytes; }
  void FreeMemory(size_t bytes) { available_memory += bytes; }
  size_t GetAvailableMemory() { return available_memory; }
  explicit BaseManager(ProfileType pprofile) : profile(pprofile) {}
  ~BaseManager() {}

 public:
  virtual void InitializeVectors(int Npoints) = 0;
  virtual void InitializeConstants(double half_width, double temperature, double partition, double dfreq,
                                   double meanmass, double pressure = 10.0, double ref_temp = 296.0) = 0;
  virtual void TransferVectors(size_t Nener, double* h_energies, double* h_nu, double* h_aif, int* h_gns,
                               double* h_gamma = NULL, double* h_n = NULL) = 0;
  virtual void TransferFreq(double* h_freq, double* h_intens, int N) = 0;
  virtual void ExecuteCrossSection(int N, int N_ener, int start_idx) = 0;
  virtual void TransferResults(double* h_freq, double* h_intens, int N) = 0;
  virtual void Cleanup() = 0;
  virtual bool ReadyForWork() { return true; }
  size_t GetNtrans() { return N_trans; }
};
