// gx_ingest_kernels.hip -- kernels of the ExoMol .trans reader (driven by gx_ingest.cpp).
//
// A piece of .trans text in HBM is read in four passes, all HBM-bound byte and integer
// work:
//   k_tok_count / k_tok_emit  token starts (a non-space byte after a space, C-locale
//                             isspace), 16 bytes per thread, counted per 4 KiB tile and
//                             scattered to their rank after a one-block scan;
//   k_records                 record r = tokens 3r, 3r+1, 3r+2 read as int, int, double
//                             (gx_decimal.h), then the state lookups of ReadNextState
//                             (ExomolStateReader.cpp:156-162): nu = E_f - E_i, E'' = E_i,
//                             g = g_tot(f), the broadening key of (J_i, J_f);
//   k_keep_count / k_keep_emit  the read loop's filter (VoigtProfile.cpp:65-68) as an
//                             ordered compaction, with gamma/n carried forward from the last
//                             record whose key matched (GetGammaN leaves them unchanged
//                             otherwise, ExomolStateReader.cpp:7-37).
#include <hip/hip_runtime.h>

#include "gx_decimal.h"
#include "gx_ingest_internal.h"

namespace gx {
namespace {

__device__ const uint64_t g_pow5_words[] = {GX_POW5_WORDS};
__device__ const int g_pow5_shifts[] = {GX_POW5_SHIFTS};

// Exclusive prefix sum over a 256-thread block; *total = block sum.
__device__ __forceinline__ unsigned block_excl_sum(unsigned v, unsigned* total) {
  __shared__ unsigned s_w[kTokBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  unsigned pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kTokBlock / 64; ++w) {
    if (w < wid) pre += s_w[w];
    tot += s_w[w];
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

// Inclusive prefix max over a 256-thread block; *total = block max.
__device__ __forceinline__ int block_incl_max(int v, int* total) {
  __shared__ int s_w[kTokBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x = max(x, y);
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  int pre = -1, tot = -1;
#pragma unroll
  for (int w = 0; w < kTokBlock / 64; ++w) {
    if (w < wid) pre = max(pre, s_w[w]);
    tot = max(tot, s_w[w]);
  }
  __syncthreads();
  *total = tot;
  return max(pre, x);
}

// Token-start bits of the 16 bytes at base (bytes at or beyond len count as spaces).
__device__ __forceinline__ unsigned start_bits(const unsigned char* __restrict__ text, size_t len, size_t base) {
  if (base >= len) return 0u;
  const uint4 q = *reinterpret_cast<const uint4*>(text + base);
  const unsigned w[4] = {q.x, q.y, q.z, q.w};
  bool prev_space = base == 0 ? true : dec_is_space(text[base - 1]);
  unsigned m = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const unsigned char c = (unsigned char)(w[k >> 2] >> (8 * (k & 3)));
    const bool sp = base + k >= len || dec_is_space(c);
    if (!sp && prev_space) m |= 1u << k;
    prev_space = sp;
  }
  return m;
}

__global__ __launch_bounds__(kTokBlock) void k_tok_count(const unsigned char* __restrict__ text, size_t len,
                                                         unsigned* __restrict__ counts) {
  const size_t base = (size_t)blockIdx.x * kTokTile + (size_t)threadIdx.x * 16;
  unsigned tot;
  block_excl_sum(__popc(start_bits(text, len, base)), &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kTokBlock) void k_tok_emit(const unsigned char* __restrict__ text, size_t len,
                                                        const unsigned* __restrict__ offsets,
                                                        unsigned* __restrict__ tokpos) {
  const size_t base = (size_t)blockIdx.x * kTokTile + (size_t)threadIdx.x * 16;
  unsigned m = start_bits(text, len, base);
  unsigned tot;
  unsigned idx = offsets[blockIdx.x] + block_excl_sum(__popc(m), &tot);
  while (m) {
    const int k = __ffs(m) - 1;
    m &= m - 1;
    tokpos[idx++] = (unsigned)(base + k);
  }
}

// One block of 1024 threads: v[0..n) -> exclusive prefix sums in place, *total = sum.
__global__ __launch_bounds__(1024) void k_scan_sum(unsigned* __restrict__ v, int n,
                                                   unsigned long long* __restrict__ total) {
  __shared__ unsigned long long s_w[16];
  const int per = (n + 1023) / 1024;
  const int a = min(n, (int)threadIdx.x * per), z = min(n, a + per);
  unsigned long long mine = 0;
  for (int i = a; i < z; ++i) mine += v[i];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long x = mine;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  unsigned long long run = x - mine;
  for (int w = 0; w < wid; ++w) run += s_w[w];
  for (int i = a; i < z; ++i) {
    const unsigned c = v[i];
    v[i] = (unsigned)run;
    run += c;
  }
  if (threadIdx.x == 1023) *total = run;
}

// One block of 1024 threads: v[0..n) -> exclusive prefix max in place (-1 before the first).
__global__ __launch_bounds__(1024) void k_scan_max(int* __restrict__ v, int n) {
  __shared__ int s_w[16];
  const int per = (n + 1023) / 1024;
  const int a = min(n, (int)threadIdx.x * per), z = min(n, a + per);
  int mine = -1;
  for (int i = a; i < z; ++i) mine = max(mine, v[i]);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = mine;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x = max(x, y);
  }
  if (lane == 63) s_w[wid] = x;
  __syncthreads();
  int run = __shfl_up(x, 1, 64);
  if (lane == 0) run = -1;
  for (int w = 0; w < wid; ++w) run = max(run, s_w[w]);
  for (int i = a; i < z; ++i) {
    const int c = v[i];
    v[i] = run;
    run = max(run, c);
  }
}

__global__ void k_reset_ctl(FeedCtl* ctl) {
  ctl->min_over = 0x7fffffff;
  ctl->min_badid = 0x7fffffff;
  ctl->nhost = 0;
  ctl->pad = 0;
}

__global__ __launch_bounds__(256) void k_records(const unsigned char* __restrict__ text, size_t len,
                                                 const unsigned* __restrict__ tokpos, int nrec, IngestView v,
                                                 RecordArrays o, FeedCtl* __restrict__ ctl,
                                                 int* __restrict__ host_list) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const char* t = reinterpret_cast<const char*>(text);
  const char* end = t + len;
  int idf = 0, idi = 0;
  double a = 0.0;
  int rc = dec_parse_int(t + tokpos[3 * r], end, &idf);
  rc |= dec_parse_int(t + tokpos[3 * r + 1], end, &idi);
  rc |= dec_parse_double(t + tokpos[3 * r + 2], end, g_pow5_words, g_pow5_shifts, &a);
  if (rc != GX_TOK_OK) {
    o.flag[r] = 1;
    host_list[atomicAdd(&ctl->nhost, 1)] = r;
    return;
  }
  const int f = idf - 1, i = idi - 1;                 // ExomolStateReader.cpp:156-157
  if (f < 0 || f >= v.nstates || i < 0 || i >= v.nstates) {
    o.flag[r] = 2;
    atomicMin(&ctl->min_badid, r);
    return;
  }
  o.flag[r] = 0;
  const double ei = v.E[i];
  const double nu = v.E[f] - ei;                      // ExomolStateReader.cpp:158-160
  o.nu[r] = nu;
  o.elow[r] = ei;
  o.aif[r] = a;
  o.gns[r] = v.g[f];
  int b = -1;
  if (v.use_broad) {
    const int ji = v.J[i] - v.jlo, jf = v.J[f] - v.jlo;   // GetGammaN(state_i, state_f)
    if ((unsigned)ji < (unsigned)v.jn && (unsigned)jf < (unsigned)v.jn) b = v.btab[(size_t)ji * v.jn + jf];
  }
  o.bidx[r] = b;
  if (nu > v.hi) atomicMin(&ctl->min_over, r);
}

__global__ void k_gather_pos(const unsigned* __restrict__ tokpos, const int* __restrict__ recs, int n,
                             unsigned* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = tokpos[3 * recs[k]];
}

__device__ __forceinline__ bool kept(const RecordArrays& r, int k, int nread, int brk, double lo) {
  return k < nread && k != brk && r.nu[k] >= lo;
}

__global__ __launch_bounds__(kRecBlock) void k_keep_count(RecordArrays r, int nread, int brk, double lo,
                                                          unsigned* __restrict__ counts, int* __restrict__ maxf) {
  const int k = blockIdx.x * kRecBlock + threadIdx.x;
  unsigned tot;
  block_excl_sum(kept(r, k, nread, brk, lo) ? 1u : 0u, &tot);
  int mx;
  block_incl_max(k < nread && r.bidx[k] >= 0 ? k : -1, &mx);
  if (threadIdx.x == 0) {
    counts[blockIdx.x] = tot;
    maxf[blockIdx.x] = mx;
  }
}

__global__ __launch_bounds__(kRecBlock) void k_keep_emit(RecordArrays r, int nread, int brk, double lo,
                                                         const double2* __restrict__ bval, double2* __restrict__ carry,
                                                         int in_slot, LinesOut out, size_t base,
                                                         const unsigned* __restrict__ offsets,
                                                         const int* __restrict__ maxf) {
  const int k = blockIdx.x * kRecBlock + threadIdx.x;
  const bool keep = kept(r, k, nread, brk, lo);
  unsigned tot;
  const unsigned pos = offsets[blockIdx.x] + block_excl_sum(keep ? 1u : 0u, &tot);
  int mx;
  const int last = max(maxf[blockIdx.x], block_incl_max(k < nread && r.bidx[k] >= 0 ? k : -1, &mx));
  const double2 gn = last >= 0 ? bval[r.bidx[last]] : carry[in_slot];
  if (keep) {
    const size_t o = base + pos;
    out.elow[o] = r.elow[k];
    out.nu[o] = r.nu[k];
    out.aif[o] = r.aif[k];
    out.gns[o] = r.gns[k];
    out.gamma[o] = gn.x;
    out.nexp[o] = gn.y;
  }
  if (k == nread - 1) carry[1 - in_slot] = gn;
}

}  // namespace

hipError_t launch_tok_count(const unsigned char* text, size_t len, unsigned* counts, hipStream_t s) {
  const size_t nblk = (len + kTokTile - 1) / kTokTile;
  if (nblk == 0) return hipSuccess;
  k_tok_count<<<(unsigned)nblk, kTokBlock, 0, s>>>(text, len, counts);
  return hipGetLastError();
}

hipError_t launch_scan_sum(unsigned* v, int n, unsigned long long* total, hipStream_t s) {
  k_scan_sum<<<1, 1024, 0, s>>>(v, n, total);
  return hipGetLastError();
}

hipError_t launch_tok_emit(const unsigned char* text, size_t len, const unsigned* offsets, unsigned* tokpos,
                           hipStream_t s) {
  const size_t nblk = (len + kTokTile - 1) / kTokTile;
  if (nblk == 0) return hipSuccess;
  k_tok_emit<<<(unsigned)nblk, kTokBlock, 0, s>>>(text, len, offsets, tokpos);
  return hipGetLastError();
}

hipError_t launch_reset_ctl(FeedCtl* ctl, hipStream_t s) {
  k_reset_ctl<<<1, 1, 0, s>>>(ctl);
  return hipGetLastError();
}

hipError_t launch_records(const unsigned char* text, size_t len, const unsigned* tokpos, int nrec,
                          const IngestView& v, const RecordArrays& r, FeedCtl* ctl, int* host_list,
                          hipStream_t s) {
  if (nrec <= 0) return hipSuccess;
  k_records<<<(nrec + 255) / 256, 256, 0, s>>>(text, len, tokpos, nrec, v, r, ctl, host_list);
  return hipGetLastError();
}

hipError_t launch_gather_pos(const unsigned* tokpos, const int* recs, int n, unsigned* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_gather_pos<<<(n + 255) / 256, 256, 0, s>>>(tokpos, recs, n, out);
  return hipGetLastError();
}

hipError_t launch_keep(const RecordArrays& r, int nread, int brk, double lo, const double2* bval, double2* carry,
                       int in_slot, LinesOut out, size_t base, unsigned* counts, int* maxf,
                       unsigned long long* nkept, hipStream_t s) {
  if (nread <= 0) return hipSuccess;
  const int nblk = (nread + kRecBlock - 1) / kRecBlock;
  k_keep_count<<<nblk, kRecBlock, 0, s>>>(r, nread, brk, lo, counts, maxf);
  k_scan_sum<<<1, 1024, 0, s>>>(counts, nblk, nkept);
  k_scan_max<<<1, 1024, 0, s>>>(maxf, nblk);
  k_keep_emit<<<nblk, kRecBlock, 0, s>>>(r, nread, brk, lo, bval, carry, in_slot, out, base, counts, maxf);
  return hipGetLastError();
}

}  // namespace gx
