// gx_common.h -- host-side building blocks shared by the manager context
// (gx_api.cpp), the batched (T,P) table (gx_table.cpp) and the request
// server (gx_server.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

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
  int reserve(size_t n, bool voigt);
  void release();
  int upload(size_t n, const double* e, const double* nu, const double* a, const int* g, const double* gam,
             const double* nexp, hipStream_t s);
};

// Per-point derived line data (FP64).
struct Derived {
  DerivedF64 d{};
  size_t cap = 0;
  int reserve(size_t n);
  void release();
};

// Per-bin-block line slices.
struct Ranges {
  int2* r = nullptr;
  int cap = 0;
  int reserve(int nblocks);
  void release();
};

// Launch line prep, block ranges and accumulation of one point (asynchronous).
// ev, when not NULL, holds 4 events recorded around the kernels.
int run_point(const PointConsts& pc, const GridView& gv, const LineStore& ls, int nlines, int precision,
              Derived& der, Ranges& rng, double* intens, unsigned long long* d_evals, hipStream_t s, hipEvent_t* ev);

}  // namespace gx
