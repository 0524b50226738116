// gx_noalloc.hip -- gx_execute_noalloc: the FP64 path of gx_execute for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The per-line intensities, Doppler widths, Lorentz ratios and
// windows, the per-block line slices and the evaluation counters live in one
// file-scope device array, `scratch`, instead of buffers the context allocates.
// Its device address is looked up once per GPU (noalloc_setup, called from
// gx_create with that GPU current) and only that address is handed to kernels.
//
// Every call on one GPU shares that GPU's array and finds the previous call's
// contents in it: the counters are cleared by a fill kernel at the start of each
// call, everything else is fully written before it is read.  Calls that overlap
// on one GPU (two contexts on different streams, or host threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gx.h"
#include "gx_common.h"
#include "gx_internal.h"
#include "gx_noalloc.h"

namespace gx {

namespace {

// Layout of the array (byte offsets, each 256-aligned).
constexpr size_t kLinesCap = GX_NOALLOC_MAX_LINES;
constexpr size_t kBlocksCap = (size_t)GX_NOALLOC_MAX_BINS / kBinsF64;
constexpr size_t kOffS = 0;
constexpr size_t kOffG = kOffS + 8 * kLinesCap;
constexpr size_t kOffY = kOffG + 8 * kLinesCap;
constexpr size_t kOffIb = kOffY + 8 * kLinesCap;
constexpr size_t kOffIe = kOffIb + 4 * kLinesCap;
constexpr size_t kOffRanges = kOffIe + 4 * kLinesCap;
constexpr size_t kOffEvals = kOffRanges + sizeof(int2) * kBlocksCap;
constexpr size_t kScratchBytes = kOffEvals + sizeof(unsigned long long) * kEvalSlots;
static_assert(kScratchBytes == GX_NOALLOC_SCRATCH_BYTES, "scratch size differs from include/gx.h");
static_assert(GX_NOALLOC_MAX_BINS % kBinsF64 == 0, "bins limit must be whole blocks");

constexpr int kMaxDevices = 64;

std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};   // device address of `scratch` per GPU

}  // namespace

static __device__ __align__(256) unsigned long long scratch[kScratchBytes / 8];

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(GX_NO_DEVICE, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return GX_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "noalloc: hipGetSymbolAddress(scratch)");
  if (!p) return fail(GX_DEVICE_ERROR, "noalloc: scratch has no device address");
  g_ptr[device] = static_cast<unsigned char*>(p);
  return GX_OK;
}

// Enqueue prep, block ranges and accumulation of one FP64 point on stream s with
// every intermediate buffer inside the device's scratch array.  Allocates nothing.
int execute_noalloc(int device, const PointConsts& pc, const GridView& gv, const LineStore& ls, int nlines,
                    double* intens, hipStream_t s) {
  if (nlines < 0 || (size_t)nlines > kLinesCap)
    return fail(GX_INVALID_ARG, "execute_noalloc: more than GX_NOALLOC_MAX_LINES lines");
  if (gv.points < 0 || (size_t)gv.points > (size_t)GX_NOALLOC_MAX_BINS)
    return fail(GX_INVALID_ARG, "execute_noalloc: more than GX_NOALLOC_MAX_BINS bins");
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (device >= 0 && device < kMaxDevices) base = g_ptr[device];
  }
  if (!base) return fail(GX_DEVICE_ERROR, "execute_noalloc: scratch not set up on this device");
  DerivedF64 d{reinterpret_cast<double*>(base + kOffS), reinterpret_cast<double*>(base + kOffG),
               reinterpret_cast<double*>(base + kOffY), reinterpret_cast<int*>(base + kOffIb),
               reinterpret_cast<int*>(base + kOffIe)};
  int2* ranges = reinterpret_cast<int2*>(base + kOffRanges);
  unsigned long long* evals = reinterpret_cast<unsigned long long*>(base + kOffEvals);
  const int nblocks = gv.points > 0 ? (gv.points + kBinsF64 - 1) / kBinsF64 : 0;
  LinesView lv{ls.nu, ls.elow, ls.aif, ls.gns, ls.gam, ls.nexp, nlines};
  hipError_t e;
  // counters start at zero: a fill kernel (all-zero bits), not a memset node
  if ((e = launch_fill(reinterpret_cast<double*>(evals), kEvalSlots, 0.0, s))) return hip_fail(e, "noalloc: fill");
  if ((e = launch_line_prep_f64(pc, gv, lv, d, evals, s))) return hip_fail(e, "noalloc: line prep launch");
  const double cut = max_cutoff(pc, ls.max_nu, gv.dfreq);
  if ((e = launch_block_ranges(gv, ls.nu, nlines, cut, kBinsF64, nblocks, ranges, s)))
    return hip_fail(e, "noalloc: ranges launch");
  if ((e = launch_accum_f64(pc, gv, ls.nu, d, ranges, nblocks, intens, s)))
    return hip_fail(e, "noalloc: accumulate launch");
  return GX_OK;
}

}  // namespace gx
