// gx_far.hip -- far field of the mixed path's region I.
//
// W4's region-I term (|x| + y >= 15) is T c / (1/2 + T^2) with T = y - i x and c the
// 0.5641896 of humlic (profiles.cpp:16-18).  Since T / (1/2 + T^2) = (1/(T - ia) +
// 1/(T + ia)) / 2 with a = 1/sqrt(2), and x = (nu - f) g, y = gamma g, one line adds
//     S g Re w_I = Re[ i kappa (1/(zeta+ - f) + 1/(zeta- - f)) ],
//     kappa = c S / 2,  zeta+- = nu +- a/g + i gamma          (1/g = alpha, y/g = gamma)
// to bin f: two complex poles per line.  For a cluster of lines with centre nu_c and
// every pole within r of it, and a bin at d = f - nu_c with |d| >= r / eta,
//     sum over the cluster = (1/d) sum_{k>=1} Q_k (r/d)^k,
//     Q_k = r^-k sum_lines kappa (Im (zeta+ - nu_c)^k + Im (zeta- - nu_c)^k),
// a real series in the real variable r/d.  Each line's Im (zeta - nu_c)^k is gamma times
// a polynomial of degree k - 1, so the terms beyond k = 24 add at most
// 25 rho^24 (1 + rho)^2 / (1 - rho)^2 of the cluster's sum (rho = r/|d|); the planner
// holds that below 5e-7 (rho <= 0.43).
//
// The radius of a real-centred cluster is at least its largest gamma, so with broad lines
// no cluster serves bins within about twice that distance.  Points whose lines are all
// broad (Lorentz widths >= 0.05 cm^-1, where the gamma_min of the bound below stays
// moderate) use a complex centre c = nu_c + i gamma_c instead:
//     sum over the cluster = Re[ -i (1/D) sum_{k=0..23} M_k (r/D)^k ],  D = f - c,
//     M_k = sum_lines kappa ((zeta+ - c)^k + (zeta- - c)^k) / r^k,
// where r bounds |zeta - c|.  A line's share is kappa gamma / |zeta - f|^2, and the
// tail of its series is at most kappa rho^24 / ((1 - rho) |D|), rho = r / |D|, so the
// relative truncation error is at most rho^24 (1 + rho)^2 |D| / ((1 - rho) gamma),
// which the planner holds below 5e-7 with the cluster's lowest gamma.
//
// Clusters: leaves of 64 consecutive lines of the sorted list and their binary tree.
// Leaf moments come from the lines (k_far_leaf), a parent's from its children's by
// the shift of the centre (k_far_up).  For each (point, bin block), k_far_plan picks
// clusters among the lines whose windows cover the whole block (a contiguous slice of
// the sorted list), far enough that the series converges and such that no line's W4
// core reaches the block (so every pair is a region-I pair of the FP64 path too); the
// region-I kernel skips those lines and k_far_eval adds their sums in FP64.
//
// Cluster record (kFarRec FP32 values): radius, max 1/g, min y, lowest and highest gamma,
// gamma_c (0 for real centres), then Q_1..Q_24 (real centres) or (Re, Im) of M_0..M_23.
// Each task's records are stored field-major (FarRec): field f of cluster k at
// f * nclus + k, so threads on consecutive clusters read and write consecutive values.
#include <hip/hip_runtime.h>

#include "gx_internal.h"

// Moments, series and direct sums here are approximations held to a truncation bound, so
// they let the compiler fuse multiply-adds; the tests that decide which clusters serve which
// bins (series_ok, serves, the plans) keep every rounding explicit.
namespace gx {
namespace {

constexpr double kPoleShift = 0.70710678118654752440;   // a = 1/sqrt(2)
constexpr double kKappa = 0.5 * 0.5641896;             // c / 2, c of humlic's region I

__device__ __forceinline__ double nu_if(double v) { return v == 0.0 ? 1e-6 : v; }

// 1/d for normal d != 0: the hardware reciprocal (4.6e-8 relative at most, measured over
// 4e6 values spread over 2^+-12 on MI355X) refined by one Newton step, which squares the
// error (~2e-15; the series are approximations held to a bound far above this rounding).
__device__ __forceinline__ double rcp_nr(double d) {
  const double r = __builtin_amdgcn_rcp(d);
  return __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
}


// One cluster's record inside a task's field-major record array: rc[f] is field f.
template <class T>
struct FarRec {
  T* p;            // field 0 of the cluster
  size_t stride;   // clusters of the task (the distance between fields)
  __device__ __forceinline__ T& operator[](int f) const { return p[(size_t)f * stride]; }
};
using FarRecC = FarRec<const FarVal>;

// Cluster geometry, for every level: centre = midpoint of its lines' nu range.
__global__ void k_far_geom(MixedPlan P, double4* __restrict__ geom) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.far_nclus) return;
  int lev = 0;
  while (lev + 1 < P.far_nlev && k >= P.far_off[lev + 1]) ++lev;
  const int idx = k - P.far_off[lev];
  const long long l0 = (long long)idx * kFarLeaf << lev;
  const long long l1 = min((long long)P.nlines, l0 + ((long long)kFarLeaf << lev));
  const double lo = nu_if(P.nu[l0]), hi = nu_if(P.nu[l1 - 1]);
  geom[k] = make_double4(0.5 * (lo + hi), lo, hi, 0.0);
}

// Moment kind of each point: complex centres when the Lorentz widths of 4096 lines
// spread over the list are all >= kFarCplxGamma.  The kind changes which clusters
// serve which bins, never the bounds every served pair meets.
__global__ __launch_bounds__(256) void k_far_kind(MixedPlan P) {
  const MixedTask tk = P.tasks[blockIdx.x];
  double gmin = 1e300;
  for (int k = threadIdx.x; k < 4096; k += 256) {
    const long long l = (long long)k * P.nlines / 4096;
    const double g = tk.g[l];
    gmin = fmin(gmin, mixed_y(tk, (int)l, g) * (1.0 / g));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) gmin = fmin(gmin, __shfl_xor(gmin, o));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = gmin;
  __syncthreads();
  if (threadIdx.x == 0)
    P.far_kind[blockIdx.x] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3])) >= kFarCplxGamma ? 1 : 0;
}

// FP32 bound of a non-negative FP64 value, rounded up (UP) or down: a maximum or minimum of
// such bounds over the wave bounds the FP64 one (one ulp step on the bits, v >= 0).
template <bool UP>
__device__ __forceinline__ float f32_bound(double v) {
  float f = (float)v;
  if (UP ? (double)f < v : (double)f > v) f = __int_as_float(__float_as_int(f) + (UP ? 1 : -1));
  return f;
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {   // every source lane valid for these controls
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}

// gfx950 lane swaps: permlane32_swap exchanges the upper 32 lanes of a with the lower 32 lanes
// of b, permlane16_swap the odd 16-lane rows of a with the even rows of b.  After the swap, in
// every lane one of (a, b) holds the lane's own value and the other the value of its partner at
// lane distance D (a's in the lower half or the even rows, b's in the others).
template <int D>
__device__ __forceinline__ void lane_swap(float& a, float& b) {
  static_assert(D == 32 || D == 16, "lane distance of a swap");
  if constexpr (D == 32) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
  } else {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
  }
}

// Maximum (MAX) or minimum over the wave, in every lane: quad swaps and rotations by 4 and
// 8 inside each row of 16 lanes, then across rows by the lane swaps.
template <bool MAX>
__device__ __forceinline__ float wave_extreme_f32(float v) {
  auto pick = [](float a, float b) { return MAX ? fmaxf(a, b) : fminf(a, b); };
  v = pick(v, dpp_f32<0xB1>(v));          // quad_perm [1, 0, 3, 2]
  v = pick(v, dpp_f32<0x4E>(v));          // quad_perm [2, 3, 0, 1]
  v = pick(v, dpp_f32<0x124>(v));         // row_ror 4
  v = pick(v, dpp_f32<0x128>(v));         // row_ror 8
  float a = v, b = v;
  lane_swap<16>(a, b);
  v = pick(a, b);
  a = v;
  b = v;
  lane_swap<32>(a, b);
  return pick(a, b);
}

// Sums over the 64 lanes of N values per lane (N a multiple of 8), by halving: at lane
// distance 32, 16 and 8 each lane keeps half of its values and adds the partner's copy of
// that half (one lane swap of the value pair at 32 and 16, a row rotation by 8 inside each
// row of 16), then the N/8 values left are summed over the 8 lanes (row rotation by 4, quad
// swaps by 2 and 1) into lane 8m + 4 of each group of eight.  There lane 8m + 4 holds the
// totals of values first(lane) + [0, N/8), first(lane) = N/2 bit5 + N/4 bit4 + N/8 bit3 of its
// lane index.  Fixed order (a + b = b + a in IEEE): the totals do not depend on scheduling.
// Every lane of the wave must be active.
template <int D, int H, int N>
__device__ __forceinline__ void wave_halve(float (&v)[N], int lane) {
#pragma clang fp contract(off)
  if constexpr (D == 8) {
    const bool up = (lane & D) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const float keep = up ? v[H + i] : v[i];
      const float send = up ? v[i] : v[H + i];
      v[i] = keep + dpp_f32<0x128>(send);     // row_ror 8: the lane at distance 8 in the row
    }
  } else {
#pragma unroll
    for (int i = 0; i < H; ++i) {
      lane_swap<D>(v[i], v[H + i]);
      v[i] = v[i] + v[H + i];
    }
  }
}

template <int N>
__device__ __forceinline__ void wave_sums(float (&v)[N], int lane) {
#pragma clang fp contract(off)
  wave_halve<32, N / 2>(v, lane);
  wave_halve<16, N / 4>(v, lane);
  wave_halve<8, N / 8>(v, lane);
  // row_ror 4 reads lane l - 4 of the row, so lanes 8m + 4 .. 8m + 7 hold the pair sums of
  // their group; quad swaps by 2 and 1 inside the quad 8m + 4 .. 8m + 7 finish the group
#pragma unroll
  for (int i = 0; i < N / 8; ++i) {
    v[i] = v[i] + dpp_f32<0x124>(v[i]);
    v[i] = v[i] + dpp_f32<0x4E>(v[i]);
    v[i] = v[i] + dpp_f32<0xB1>(v[i]);
  }
}

// Points per wave of the leaf kernel: each wave reads its lines' nu once for them.
constexpr int kLeafTasks = 8;

// Leaf moments of every task: one wave per (leaf, kLeafTasks points), one line per lane, the
// lanes' terms summed by wave_sums.  The radius, centre and the extremes the truncation bounds
// use are FP64; each line's scaled pole offsets w (|w| <= 1) and its weight go to FP32, and
// the powers and the wave sums run in FP32: a moment's rounding error is then ~1e-7 of the
// sum of its terms' magnitudes, and the series weighs moment k by at most eta^(k+1), so the
// far sums keep ~1e-7 relative accuracy (the weights are 0.28 S, S/alpha being FP32 in the
// mixed path already).  The records are stored in FP32 (FarVal) and every reader computes in FP64.
__device__ __forceinline__ void leaf_record(const MixedTask& tk, bool cplx, int l, bool act, double nu, double c,
                                            int lane, FarRec<double> rec) {
  {
    double y = 0.0;
    float amp = 0.0f;
    if (act) {
      y = mixed_y(tk, l, tk.g[l]);
      amp = tk.amp[l];
    }
    const double al = act ? nu * tk.dpwcoef : 0.0;               // alpha: g = 1 / (nu dpwcoef)
    const double s = kPoleShift * al, im = y * al;               // im: the line's gamma
    const double dp = nu + s - c, dm = nu - s - c;
    // the bounds' extremes from FP32 values rounded outward (all are >= 0): the maxima can
    // only grow and the minima shrink, so every test built on them stays conservative
    // alpha = nu dpwcoef rises along the sorted lines when the first one has nu > 1e-6 (a line at
    // nu = 0 counts as 1e-6, above the lines just after it), and the active lanes are a prefix of
    // the wave: then its maximum is the last active lane's
    const float alb = f32_bound<true>(al);
    const bool rising = __builtin_amdgcn_readfirstlane(act && nu > 1e-6 ? 1 : 0) != 0;
    const int last = __popcll(__ballot(act)) - 1;
    const double amax = (double)(rising ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(alb), last))
                                        : wave_extreme_f32<true>(alb));
    const double ymin = (double)wave_extreme_f32<false>(act ? f32_bound<false>(y) : INFINITY);
    double glo = 0.0, ghi = 0.0;                                 // gamma range: complex centres only
    if (cplx) {
      glo = (double)wave_extreme_f32<false>(act ? f32_bound<false>(im) : INFINITY);
      ghi = (double)wave_extreme_f32<true>(act ? f32_bound<true>(im) : 0.0f);
    }
    const double gc = cplx ? (double)(float)(0.5 * (glo + ghi)) : 0.0;   // stored as FP32
    const double wim = im - gc;                                  // Im (zeta - centre)
    const double r2 =
        (double)wave_extreme_f32<true>(act ? f32_bound<true>(fmax(dp * dp, dm * dm) + wim * wim) : 0.0f);
    double r = (double)f32_bound<true>(sqrt(r2));               // stored as FP32, rounded up
    if (!(r > 0.0)) r = 1.0;
    const double ir = rcp_nr(r);
    const float kap = act ? (float)(kKappa * (double)amp * al) : 0.0f;   // amp = S g
    const float wi = (float)(wim * ir);
    if (lane == 0) {
      rec[0] = r;
      rec[1] = amax;
      rec[2] = ymin;
      rec[3] = glo;
      rec[4] = ghi;
      rec[5] = gc;
    }
    // every lane's terms for all moments, then one halving reduction over the wave; lanes
    // 8m (m < 8) write the totals they hold
    const float wrp = (float)(dp * ir), wrm = (float)(dm * ir);
    const int first = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
    if (!cplx) {
      // Im w^k of each pole by the real recurrence z_k = 2 Re(w) z_(k-1) - |w|^2 z_(k-2)
      // (w and its conjugate are the roots of t^2 - 2 Re(w) t + |w|^2), z_0 = 0, z_1 = Im w,
      // each scaled by the line weight kap
      float v[kFarP];
      const float tp = 2.0f * wrp, np = wrp * wrp + wi * wi;
      const float tm = 2.0f * wrm, nm = wrm * wrm + wi * wi;
      float p1 = kap * wi, p0 = 0.0f, m1 = p1, m0 = 0.0f;        // the weight kap carried along
#pragma unroll
      for (int k = 0; k < kFarP; ++k) {
        v[k] = p1 + m1;
        const float p2 = tp * p1 - np * p0, m2 = tm * m1 - nm * m0;
        p0 = p1;
        p1 = p2;
        m0 = m1;
        m1 = m2;
      }
      wave_sums(v, lane);
      if ((lane & 7) == 4) {
#pragma unroll
        for (int t = 0; t < kFarP / 8; ++t) rec[kFarHead + first * (kFarP / 8) + t] = (double)v[t];
      }
    } else {
      float v[2 * kFarP];                                        // (Re, Im) of M_0..M_(P-1)
      float pr = kap, pi = 0.0f, mr = kap, mi = 0.0f;            // kap w+^0, kap w-^0
#pragma unroll
      for (int k = 0; k < kFarP; ++k) {
        v[2 * k] = pr + mr;
        v[2 * k + 1] = pi + mi;
        const float nr = pr * wrp - pi * wi, ni = pr * wi + pi * wrp;
        const float n2r = mr * wrm - mi * wi, n2i = mr * wi + mi * wrm;
        pr = nr;
        pi = ni;
        mr = n2r;
        mi = n2i;
      }
      wave_sums(v, lane);
      if ((lane & 7) == 4) {
#pragma unroll
        for (int t = 0; t < kFarP / 4; ++t) rec[kFarHead + first * (kFarP / 4) + t] = (double)v[t];
      }
    }
  }
}

// Near range U of a leaf at a real-centred point: the bins left to the core pass for all its
// lines, namely every bin of a line's W4 core (|nu - f| g + y < 15 implies |nu - f| < (15 - ymin)
// amax, with the leaf's outward-rounded extremes) and every bin nearer to the leaf's centre than
// r / kNearRho, where its series would not meet the truncation bound.  Elsewhere in the blocks
// whose near lines it holds, every pair of the leaf is a region-I pair and the series converges.
// Empty (not near-eligible) unless the leaf is full, U lies within kmax of every line's centre bin
// (the core pass rows hold it) and inside every line's window (ib, ie rise with nu).
__device__ __forceinline__ int2 near_range(const MixedPlan& P, const MixedTask& tk, int leaf, double r, double amax,
                                           double ymin) {
#pragma clang fp contract(off)
  const int l0 = leaf * kFarLeaf, l1 = l0 + kFarLeaf - 1;
  if (l1 >= P.nlines) return make_int2(0, 0);
  const GridView gv = P.gv;
  const double4 gm = P.far_geom[leaf];                       // centre, lowest nu, highest nu
  // x = |nu - f| / alpha stays below 1e6 over the core pass rows (FP32 region I there)
  if (!(gm.y * tk.dpwcoef * 1e6 >= (double)(P.kmax + 2) * gv.dfreq)) return make_int2(0, 0);
  const double h = ymin < 15.0 ? (15.0 - ymin) * amax : 0.0;
  const double dc = r / kNearRho;
  const double a = fmin(gm.y - h, gm.x - dc), z = fmax(gm.z + h, gm.x + dc);
  // every bin j with a < f_j < z: j > (a - f_0) / dnu and j < (z - f_0) / dnu, with a margin of
  // 1e-3 bin for the rounding of f_j and of these quotients
  const double xa = floor((a - gv.start_nu) / gv.dfreq - 1e-3), xz = ceil((z - gv.start_nu) / gv.dfreq + 1e-3);
  if (!(xa > -1e9 && xz < 1e9)) return make_int2(0, 0);
  const int u0 = (int)xa, u1 = (int)xz;
  const int cb0 = centre_bin(nu_if(P.nu[l0]), gv), cb1 = centre_bin(nu_if(P.nu[l1]), gv);
  if (u0 < cb1 - P.kmax || u1 - 1 > cb0 + P.kmax) return make_int2(0, 0);
  if (u0 < P.ib[l1] || u1 > P.ie[l0]) return make_int2(0, 0);
  return make_int2(u0, u1);
}

// Leaf records of four consecutive leaves (one per wave) are staged in LDS and written field
// by field, four doubles per field; workgroups are mapped so that each XCD takes a contiguous
// run of leaves (workgroups go to the XCDs in turn), and the 128-byte lines of a field, shared
// by 16 leaves, fill in one L2.  Against one store per lane and field straight to HBM this
// saves the per-store address arithmetic of a VALU-bound kernel and the partial-line
// write-backs from four L2s (54 ms against 61.6 ms per 48-point table).
__global__ __launch_bounds__(256) void k_far_leaf(MixedPlan P) {
  __shared__ double s_rec[4][kFarRec + 1];   // +1: the rows start in different banks
  __shared__ double s_head[4][kLeafTasks][3];  // radius, max alpha, min y of each point (near ranges)
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int per = gridDim.x / 8;              // the launch pads gridDim.x to a multiple of 8
  const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int leaf0 = grp * 4;
  const int nleaf = min(4, P.far_n[0] - leaf0);
  if (nleaf <= 0) return;                     // the whole workgroup
  const int leaf = leaf0 + wave;
  const bool valid = wave < nleaf;
  const int l = leaf * kFarLeaf + lane;
  const bool act = valid && l < P.nlines;
  const double c = valid ? P.far_geom[leaf].x : 0.0;
  const double nu = act ? nu_if(P.nu[l]) : c;  // read once for kLeafTasks points
  const size_t ns = (size_t)P.far_nclus;
  const int t1 = min(P.ntasks, (int)(blockIdx.y + 1) * kLeafTasks);
  for (int task = blockIdx.y * kLeafTasks; task < t1; ++task) {
    const MixedTask tk = P.tasks[task];
    const bool cplx = P.far_kind[task] != 0;
    if (valid) leaf_record(tk, cplx, l, act, nu, c, lane, FarRec<double>{&s_rec[wave][0], 1});
    if (P.near && valid && lane == 0) {        // lane 0 wrote the head it copies
      const int t = task - blockIdx.y * kLeafTasks;
      s_head[wave][t][0] = s_rec[wave][0];
      s_head[wave][t][1] = s_rec[wave][1];
      s_head[wave][t][2] = cplx ? -1.0 : s_rec[wave][2];   // -1: complex centres, no near range
    }
    __syncthreads();
    const int nf = cplx ? kFarRec : kFarHead + kFarP;
    FarVal* dst = P.far_rec + (size_t)task * ns * kFarRec + leaf0;
    for (int idx = threadIdx.x; idx < nf * 4; idx += 256) {
      const int f = idx >> 2, lf = idx & 3;
      if (lf < nleaf) dst[(size_t)f * ns + lf] = (FarVal)s_rec[lf][f];   // the head is FP32 already
    }
    __syncthreads();
  }
  // near ranges of the wave's leaf at its points, one lane per point
  const int t = lane, task = blockIdx.y * kLeafTasks + lane;
  if (P.near && valid && t < kLeafTasks && task < t1) {
    const double ymin = s_head[wave][t][2];
    P.near_u[(size_t)task * ns + leaf] =
        ymin < 0.0 ? make_int2(0, 0)
                   : near_range(P, P.tasks[task], leaf, s_head[wave][t][0], s_head[wave][t][1], ymin);
  }
}

// Parents of level lev from their two children.  Real centres: shift a = (c_child - c) / r
// and scale b = r_child / r, Q_k = sum_m C(k, m) a^(k-m) b^m Q_m(child).  Complex centres:
// the same with a complex shift, M_k = sum_m C(k, m) a^(k-m) b^m M_m(child), formed by
// repeated Taylor-shift passes v_k += a v_(k-1) over v_m = b^m M_m(child).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_far_up(MixedPlan P, int lev) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.far_n[lev]) return;
  if ((P.far_kind[blockIdx.y] != 0) != CPLX) return;         // the other kind's launch
  FarVal* const tb = P.far_rec + (size_t)blockIdx.y * P.far_nclus * kFarRec;   // the task's records
  const size_t ns = (size_t)P.far_nclus;
  constexpr bool cplx = CPLX;
  const int self = P.far_off[lev] + i;
  const double c = P.far_geom[self].x;
  const int nch = min(2, P.far_n[lev - 1] - 2 * i);
  double amax = 0.0, ymin = 1e300, glo = 1e300, ghi = -1e300;
  for (int h = 0; h < nch; ++h) {
    const FarRecC cr{tb + P.far_off[lev - 1] + 2 * i + h, ns};
    amax = fmax(amax, cr[1]);
    ymin = fmin(ymin, cr[2]);
    glo = fmin(glo, cr[3]);
    ghi = fmax(ghi, cr[4]);
  }
  const double gc = cplx ? (double)(float)(0.5 * (glo + ghi)) : 0.0;   // the stored FP32 value
  double r = 0.0;
  for (int h = 0; h < nch; ++h) {
    const int ch = P.far_off[lev - 1] + 2 * i + h;
    const FarRecC cr{tb + ch, ns};
    const double dx = P.far_geom[ch].x - c, dy = cr[5] - gc;
    r = fmax(r, (cplx ? sqrt(dx * dx + dy * dy) : fabs(dx)) + cr[0]);
  }
  r = (double)f32_bound<true>(r);                // the stored FP32 value, rounded up
  if (!(r > 0.0)) r = 1.0;
  const double ir = rcp_nr(r);
  if (P.near) {                                  // hull of the children's near ranges
    int2 u = make_int2(0, 0);
    for (int h = 0; h < nch; ++h) {
      const int2 cu = P.near_u[(size_t)blockIdx.y * ns + P.far_off[lev - 1] + 2 * i + h];
      if (cu.y <= cu.x) continue;
      u = u.y > u.x ? make_int2(min(u.x, cu.x), max(u.y, cu.y)) : cu;
    }
    P.near_u[(size_t)blockIdx.y * ns + self] = u;
  }
  const FarRec<FarVal> rec{tb + self, ns};      // every head field is an FP32 value already
  rec[0] = (FarVal)r;
  rec[1] = (FarVal)amax;
  rec[2] = (FarVal)ymin;
  rec[3] = (FarVal)glo;
  rec[4] = (FarVal)ghi;
  rec[5] = (FarVal)gc;
  if constexpr (!cplx) {
    double q[kFarP];
#pragma unroll
    for (int k = 0; k < kFarP; ++k) q[k] = 0.0;
    for (int h = 0; h < nch; ++h) {
      const int ch = P.far_off[lev - 1] + 2 * i + h;
      const FarRecC cr{tb + ch, ns};
      const double a = (P.far_geom[ch].x - c) * ir, b = cr[0] * ir;
      // Q_k (k = 1..P) += sum_{m=1..k} C(k, m) a^(k-m) b^m Q_m: v_m = b^m Q_m (v_0 = Q_0 = 0 for a
      // real centre), then P in-place Taylor-shift passes v_k += a v_(k-1)
      double v[kFarP + 1];
      v[0] = 0.0;
      double bm = b;
#pragma unroll
      for (int m = 1; m <= kFarP; ++m) {
        v[m] = bm * cr[kFarHead + m - 1];
        bm *= b;
      }
#pragma unroll
      for (int j = 1; j <= kFarP; ++j)
#pragma unroll
        for (int k = kFarP; k >= j; --k) v[k] = v[k] + a * v[k - 1];
#pragma unroll
      for (int k = 1; k <= kFarP; ++k) q[k - 1] += v[k];
    }
#pragma unroll
    for (int k = 0; k < kFarP; ++k) rec[kFarHead + k] = (FarVal)q[k];
  } else {
    double qr[kFarP], qi[kFarP];
#pragma unroll
    for (int k = 0; k < kFarP; ++k) qr[k] = qi[k] = 0.0;
    for (int h = 0; h < nch; ++h) {
      const int ch = P.far_off[lev - 1] + 2 * i + h;
      const FarRecC cr{tb + ch, ns};
      const double ar = (P.far_geom[ch].x - c) * ir, ai = (cr[5] - gc) * ir, b = cr[0] * ir;
      double vr[kFarP], vi[kFarP];
      double bm = 1.0;
#pragma unroll
      for (int m = 0; m < kFarP; ++m) {
        vr[m] = bm * cr[kFarHead + 2 * m];
        vi[m] = bm * cr[kFarHead + 2 * m + 1];
        bm *= b;
      }
#pragma unroll
      for (int j = 1; j < kFarP; ++j) {
#pragma unroll
        for (int k = kFarP - 1; k >= j; --k) {
          const double nr = vr[k] + (ar * vr[k - 1] - ai * vi[k - 1]);
          const double ni = vi[k] + (ar * vi[k - 1] + ai * vr[k - 1]);
          vr[k] = nr;
          vi[k] = ni;
        }
      }
#pragma unroll
      for (int k = 0; k < kFarP; ++k) {
        qr[k] += vr[k];
        qi[k] += vi[k];
      }
    }
#pragma unroll
    for (int k = 0; k < kFarP; ++k) {
      rec[kFarHead + 2 * k] = (FarVal)qr[k];
      rec[kFarHead + 2 * k + 1] = (FarVal)qi[k];
    }
  }
}

__global__ void k_far_reset(unsigned long long* c) { *c = 0ull; }

// First index in [a, z) with v[index] > key (v non-decreasing there).
__device__ __forceinline__ int upper_bound_int(const int* __restrict__ v, int a, int z, int key) {
  while (a < z) {
    const int m = a + (z - a) / 2;
    if (v[m] <= key) a = m + 1;
    else z = m;
  }
  return a;
}

// First line of [a, z) with ie > j (IB: ib > j), j in [gv.start_idx - 1, gv.start_idx + gv.points - 1]:
// the per-bin first line over the whole sorted list (k_bin_first), clamped to [a, z) (exact,
// ib and ie being non-decreasing over the list).
template <bool IB>
__device__ __forceinline__ int first_above(const MixedPlan& P, int a, int z, int j) {
  const int* t = IB ? P.far_ib1 : P.far_ie1;
  return min(max(t[j - P.gv.start_idx + 1], a), z);
}

// The per-bin first lines: one thread per bin j = gv.start_idx - 1 + t, t in [0, gv.points].
__global__ __launch_bounds__(256) void k_bin_first(MixedPlan P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > P.gv.points) return;
  const int j = P.gv.start_idx - 1 + t;
  P.far_ie1[t] = upper_bound_int(P.ie, 0, P.nlines, j);
  P.far_ib1[t] = upper_bound_int(P.ib, 0, P.nlines, j);
}

// Whether a series of P orders about centre nu_c (+ i gamma_c), radius r, converges within
// kFarTol relative to the sum at every bin at least dist from nu_c.  With rho = r / |D| the
// tail of a line's series is, relative to the line's own value: real centres (each term is
// gamma times a polynomial of degree k - 1 in w, |w| <= 1)
//     (P + 1) rho^P (1 + rho)^2 / (1 - rho)^2,
// complex centres (the line's value is kappa gamma / |zeta - f|^2 >= kappa gamma /
// (|D| (1 + rho))^2 and the tail at most kappa rho^P / ((1 - rho) |D|))
//     rho^P (1 + rho)^2 |D| / ((1 - rho) gamma_min).
// Both fall as |D| grows, so the nearest bin decides.
template <int P>
__device__ __forceinline__ bool series_ok(double r, double gamma_c, double gamma_min, double dist, bool cplx) {
#pragma clang fp contract(off)
  const double D = cplx ? sqrt(dist * dist + gamma_c * gamma_c) : dist;
  const double rho = r / D;
  if (!(rho <= kFarEta)) return false;
  double rp = 1.0;
#pragma unroll
  for (int k = 0; k < P; ++k) rp *= rho;
  const double w = (1.0 + rho) * (1.0 + rho);
  return cplx ? rp * w * D <= kFarTol * (1.0 - rho) * gamma_min
              : (double)(P + 1) * rp * w <= kFarTol * (1.0 - rho) * (1.0 - rho);
}

// Whether cluster (geometry gm, record rc, near range u) serves every bin of [f_lo, f_hi] =
// bins [j_lo, j_hi]: no line of it has its W4 core there (its lowest y is >= 15, or its lines are
// 15 Doppler widths away), no bin is in the near range of one of its leaves (the core pass has
// those pairs), and its series converges there within the truncation bound of its kind.
__device__ __forceinline__ bool serves(const double4 gm, FarRecC rc, bool cplx, double f_lo,
                                       double f_hi, int2 u, int j_lo, int j_hi) {
#pragma clang fp contract(off)
  if (u.y > u.x && u.x <= j_hi && u.y > j_lo) return false;
  const double dist = fmax(fmax(f_lo - gm.x, gm.x - f_hi), 0.0);      // centre to the bins
  const double edge = fmax(fmax(f_lo - gm.z, gm.y - f_hi), 0.0);      // the lines' nu range to the bins
  if (!(rc[2] >= 15.0 || edge >= 15.001 * rc[1])) return false;
  return series_ok<kFarP>(rc[0], rc[5], rc[3], dist, cplx);
}

// Near range of cluster k of a task (empty when near leaves are off).
__device__ __forceinline__ int2 near_u_of(const MixedPlan& P, int task, int k) {
  return P.near ? P.near_u[(size_t)task * P.far_nclus + k] : make_int2(0, 0);
}

// Sum of a cluster (record rc, centre nu_c) at bin f.
__device__ __forceinline__ double cluster_sum(FarRecC rc, double nu_c, bool cplx, double f) {
  if (!cplx) {
    const double id = rcp_nr(f - nu_c);
    const double t = rc[0] * id;
    double q = rc[kFarHead + kFarP - 1];
#pragma unroll
    for (int m = kFarP - 2; m >= 0; --m) q = q * t + rc[kFarHead + m];
    return q * t * id;
  }
  const double dr = f - nu_c, di = -rc[5];                    // D = f - c
  const double inv = rcp_nr(dr * dr + di * di);
  const double ir = dr * inv, ii = -di * inv;                 // 1/D
  const double tr = rc[0] * ir, ti = rc[0] * ii;              // r/D
  double sr = rc[kFarHead + 2 * (kFarP - 1)], si = rc[kFarHead + 2 * (kFarP - 1) + 1];
#pragma unroll
  for (int m = kFarP - 2; m >= 0; --m) {
    const double nr = (sr * tr - si * ti) + rc[kFarHead + 2 * m];
    const double ni = (sr * ti + si * tr) + rc[kFarHead + 2 * m + 1];
    sr = nr;
    si = ni;
  }
  return sr * ii + si * ir;                                   // Re(-i s / D) = Im(s / D)
}

// The far clusters of one (task, block): one thread each.
__global__ __launch_bounds__(256) void k_far_plan(MixedPlan P) {
#pragma clang fp contract(off)
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= P.nblocks) return;
  const int task = blockIdx.y;
  const size_t slot = (size_t)task * P.nblocks + b;
  int4 rng = make_int4(0, 0, 0, 0), nrng = make_int4(0, 0, 0, 0);
  int cnt = 0;
  unsigned long long far_lines = 0;
  const GridView gv = P.gv;
  const int w0 = gv.start_idx + b * P.far_bins;
  const int w1 = w0 + P.far_bins;
  const int2 rg = P.ranges[b];
  if (w1 <= gv.start_idx + gv.points && rg.y > rg.x) {
    const double f_lo = gv.freq[w0], f_hi = gv.freq[w1 - 1];
    // lines whose windows cover the whole block: ie >= w1 and ib <= w0 (ib, ie rise with nu)
    const int L1 = first_above<false>(P, rg.x, rg.y, w1 - 1);
    const int L2 = first_above<true>(P, rg.x, rg.y, w0);
    const int leafL = (L1 + kFarLeaf - 1) / kFarLeaf, leafR = L2 / kFarLeaf;
    const FarVal* recs = P.far_rec + (size_t)task * P.far_nclus * kFarRec;
    int* list = P.far_list + slot * kFarList;
    const bool cplx = P.far_kind[task] != 0;
    auto usable = [&](int lev, int idx) -> bool {
      const int k = P.far_off[lev] + idx;
      return serves(P.far_geom[k], FarRecC{recs + k, (size_t)P.far_nclus}, cplx, f_lo, f_hi, near_u_of(P, task, k),
                    w0, w1 - 1);
    };
    int i = leafL;
    if (leafL < leafR) {
      // from the left end of the slice towards the block
      while (i < leafR && cnt < kFarList) {
        int lev = min(P.far_nlev - 1, i ? __builtin_ctz(i) : 31);
        bool took = false;
        for (; lev >= 0; --lev) {
          if ((i + (1 << lev)) > leafR) continue;
          if (usable(lev, i >> lev)) {
            list[cnt++] = (lev << 24) | (i >> lev);
            i += 1 << lev;
            took = true;
            break;
          }
        }
        if (!took) break;
      }
      int j = leafR;
      while (j > i && cnt < kFarList) {
        int lev = min(P.far_nlev - 1, j ? __builtin_ctz(j) : 31);
        bool took = false;
        for (; lev >= 0; --lev) {
          if (j - (1 << lev) < i) continue;
          if (usable(lev, (j - (1 << lev)) >> lev)) {
            list[cnt++] = (lev << 24) | ((j - (1 << lev)) >> lev);
            j -= 1 << lev;
            took = true;
            break;
          }
        }
        if (!took) break;
      }
      far_lines = (unsigned long long)(i - leafL) * kFarLeaf + (unsigned long long)(leafR - j) * kFarLeaf;
      // near leaves next to the far clusters on either side: summed per bin by k_near_eval
      int i2 = i, j2 = j;
      if (P.near) {
        while (i2 < j && near_u_of(P, task, i2).y > near_u_of(P, task, i2).x) ++i2;
        while (j2 > i2 && near_u_of(P, task, j2 - 1).y > near_u_of(P, task, j2 - 1).x) --j2;
        nrng = make_int4(i, i2, j2, j);
      }
      rng = make_int4(leafL * kFarLeaf, i2 * kFarLeaf, j2 * kFarLeaf, leafR * kFarLeaf);
    }
  }
  if (P.near) P.near_rng[slot] = nrng;
  P.far_rng[slot] = rng;
  P.far_cnt[slot] = cnt;
  if (far_lines) atomicAdd(P.far_evals, far_lines * (unsigned long long)P.far_bins);
}

// Lines whose windows end inside a block (ib <= w0 < ie < w1, bins j < ie) and lines whose
// windows start inside it (w0 < ib < w1 <= ie, bins j >= ib), taken from moments by
// k_far_edges when every leaf they touch serves the whole block; else left to the
// region-I kernel.  The end of each range that is the same for every bin of the block (the
// last line ending inside it, the first line starting inside it) is moved to a leaf
// boundary, the lines beyond it staying with the region-I kernel, so k_far_edges evaluates
// at most one partial leaf per range and bin.  One wave per (task, block), the lanes testing
// the leaves side by side.
__global__ __launch_bounds__(256) void k_edge_plan(MixedPlan P) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (b >= P.nblocks) return;
  const int task = blockIdx.y;
  const size_t slot = (size_t)task * P.nblocks + b;
  int4 er = make_int4(0, 0, 0, 0);
  const GridView gv = P.gv;
  const int w0 = gv.start_idx + b * P.far_bins;
  const int w1 = w0 + P.far_bins;
  const int2 rg = P.ranges[b];
  if (w1 <= gv.start_idx + gv.points && rg.y > rg.x) {
    const double f_lo = gv.freq[w0], f_hi = gv.freq[w1 - 1];
    const int L1 = first_above<false>(P, rg.x, rg.y, w1 - 1);    // ie >= w1 from here
    const int L2 = first_above<true>(P, rg.x, rg.y, w0);         // ib > w0 from here
    const int Lx = first_above<false>(P, rg.x, rg.y, w0);        // ie > w0 from here
    const int Ly = first_above<true>(P, rg.x, rg.y, w1 - 1);     // ib >= w1 from here
    const int em0 = Lx, em1 = min(L1, L2) / kFarLeaf * kFarLeaf;
    const int ep0 = (max(L1, L2) + kFarLeaf - 1) / kFarLeaf * kFarLeaf, ep1 = Ly;
    const FarVal* recs = P.far_rec + (size_t)task * P.far_nclus * kFarRec;
    const bool cplx = P.far_kind[task] != 0;
    auto leaves_usable = [&](int a, int z) -> bool {       // every lane calls it together
      bool ok = true;
      for (int leaf = a / kFarLeaf + lane; ok && leaf * kFarLeaf < z; leaf += 64)
        ok = serves(P.far_geom[leaf], FarRecC{recs + leaf, (size_t)P.far_nclus}, cplx, f_lo, f_hi,
                    near_u_of(P, task, leaf), w0, w1 - 1);
      return __ballot(!ok) == 0ull;
    };
    if (em1 > em0 && leaves_usable(em0, em1)) {
      er.x = em0;
      er.y = em1;
    }
    if (ep1 > ep0 && leaves_usable(ep0, ep1)) {
      er.z = ep0;
      er.w = ep1;
    }
  }
  if (lane == 0) P.edge_rng[slot] = er;
}

// Far-field sums of one (task, block) per wave, added to the task's grid, one instantiation
// per moment kind.  NL clusters are summed side by side (BL lanes each, Q bins per lane: 8
// for real centres, 4 for complex ones): their records are fetched together (each lane loads
// a few fields of them; the fields lie in different rows of the field-major array), staged in
// LDS and read back moment by moment for all the lane's bins, and each bin's NL partial sums
// are added in a fixed order at the end.
template <int R, bool CPLX>
struct FarEvalShape {
  static constexpr int Q = (CPLX ? 4 : 8) > R ? (CPLX ? 4 : 8) : R;   // bins per lane
  static constexpr int BL = 64 * R / Q;                                // lanes per cluster
  static constexpr int NL = 64 / BL;                                   // clusters at a time
};

template <int R, bool CPLX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_far_eval(MixedPlan P) {
  using S = FarEvalShape<R, CPLX>;
  constexpr int NL = S::NL, BL = S::BL, Q = S::Q;
  constexpr int nf = CPLX ? kFarRec : kFarHead + kFarP;
  __shared__ double s_rc[4][NL][nf];
  __shared__ double s_c[4][NL];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * 4 + wave;
  if (b >= P.nblocks) return;
  const int task = blockIdx.y;
  if ((P.far_kind[task] != 0) != CPLX) return;               // the other kind's launch
  const size_t slot_b = (size_t)task * P.nblocks + b;
  const int cnt = P.far_cnt[slot_b];
  if (cnt == 0) return;
  const MixedTask tk = P.tasks[task];
  const GridView gv = P.gv;
  const int w0 = gv.start_idx + b * 64 * R;
  const int* list = P.far_list + slot_b * kFarList;
  const FarVal* recs = P.far_rec + (size_t)task * P.far_nclus * kFarRec;
  const size_t ns = (size_t)P.far_nclus;
  const int slot = lane / BL, bl = lane % BL;
  double f[Q], acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    f[q] = gv.freq[w0 + bl + BL * q];       // far clusters are planned for whole blocks only
    acc[q] = 0.0;
  }
  for (int e0 = 0; e0 < cnt; e0 += NL) {
    const int ne = min(NL, cnt - e0);
    wave_lds_handover();                     // the previous clusters' reads are done
    for (int idx = lane; idx < ne * nf; idx += 64) {
      const int sl = idx / nf, fld = idx - sl * nf;
      const int ent = list[e0 + sl];
      const int k = P.far_off[ent >> 24] + (ent & 0xffffff);
      s_rc[wave][sl][fld] = recs[(size_t)fld * ns + k];
      if (fld == 0) s_c[wave][sl] = P.far_geom[k].x;
    }
    wave_lds_handover();
    if (slot < ne) {
      const double* rc = &s_rc[wave][slot][0];
      const double c = s_c[wave][slot];
      if (CPLX) {
        // cluster_sum's complex series, moment by moment over the lane's bins
        double ir[Q], ii[Q], tr[Q], ti[Q], sr[Q], si[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const double dr = f[q] - c, di = -rc[5];
          const double inv = rcp_nr(dr * dr + di * di);
          ir[q] = dr * inv;
          ii[q] = -di * inv;
          tr[q] = rc[0] * ir[q];
          ti[q] = rc[0] * ii[q];
          sr[q] = rc[kFarHead + 2 * (kFarP - 1)];
          si[q] = rc[kFarHead + 2 * (kFarP - 1) + 1];
        }
#pragma unroll
        for (int m = kFarP - 2; m >= 0; --m) {
          const double mr = rc[kFarHead + 2 * m], mi = rc[kFarHead + 2 * m + 1];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const double nr = (sr[q] * tr[q] - si[q] * ti[q]) + mr;
            const double ni = (sr[q] * ti[q] + si[q] * tr[q]) + mi;
            sr[q] = nr;
            si[q] = ni;
          }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] += sr[q] * ii[q] + si[q] * ir[q];
      } else {
        // cluster_sum's real series, moment by moment over the lane's bins
        double id[Q], t[Q], v[Q];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          id[q] = rcp_nr(f[q] - c);
          t[q] = rc[0] * id[q];
          v[q] = rc[kFarHead + kFarP - 1];
        }
#pragma unroll
        for (int m = kFarP - 2; m >= 0; --m) {
          const double a = rc[kFarHead + m];
#pragma unroll
          for (int q = 0; q < Q; ++q) v[q] = v[q] * t[q] + a;
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[q] += v[q] * t[q] * id[q];
      }
    }
  }
#pragma unroll
  for (int d = BL; d < 64; d <<= 1)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] += __shfl_xor(acc[q], d);
  if (slot == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) tk.out[w0 + bl + BL * q] += acc[q];
  }
}

// Leaves staged per pass of k_near_eval.
constexpr int kNearBatch = 32;

// Lane layout of k_near_eval for blocks of 64 R bins: 8 / R leaves side by side, 8 R lanes per
// leaf, 8 bins per lane (bin lane b takes bins b + 8 R q, q < 8), so one LDS read of a moment
// serves eight series steps of its lane.
template <int R>
struct NearShape {
  static constexpr int NL = 8 / R;          // leaves at a time
  static constexpr int BL = 64 / NL;        // lanes per leaf
  static constexpr int Q = 64 * R / BL;     // bins per lane (8)
};

// Near leaves of one (task, block) per wave (the runs k_far_plan left between the far clusters
// and the block's direct lines): bin j adds a leaf's real series wherever j is outside the
// leaf's near range (the core pass evaluates the leaf's lines at the bins inside it).  The
// leaves of a run are consecutive, so each field of kNearBatch leaves is one coalesced load
// (field-major records), staged in LDS; NL leaves are summed side by side and their lanes'
// sums added in a fixed order at the end.
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_near_eval(MixedPlan P) {
  using S = NearShape<R>;
  constexpr int NL = S::NL, BL = S::BL, Q = S::Q;
  __shared__ double s_q[4][kFarP][kNearBatch];
  __shared__ double s_c[4][kNearBatch], s_r[4][kNearBatch];
  __shared__ int2 s_u[4][kNearBatch];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * 4 + wave;
  if (b >= P.nblocks) return;
  const int task = blockIdx.y;
  int4 nr = P.near_rng[(size_t)task * P.nblocks + b];
  nr.x = __builtin_amdgcn_readfirstlane(nr.x);
  nr.y = __builtin_amdgcn_readfirstlane(nr.y);
  nr.z = __builtin_amdgcn_readfirstlane(nr.z);
  nr.w = __builtin_amdgcn_readfirstlane(nr.w);
  if (nr.x >= nr.y && nr.z >= nr.w) return;
  const MixedTask tk = P.tasks[task];
  const GridView gv = P.gv;
  const FarVal* recs = P.far_rec + (size_t)task * P.far_nclus * kFarRec;
  const int2* nu_t = P.near_u + (size_t)task * P.far_nclus;
  const size_t ns = (size_t)P.far_nclus;
  const int slot = lane / BL, bl = lane % BL;
  const int w0 = gv.start_idx + b * 64 * R;
  double f[Q], acc[Q];
  int jj[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    jj[q] = w0 + bl + BL * q;
    f[q] = gv.freq[jj[q]];                   // near runs are planned for whole blocks only
    acc[q] = 0.0;
  }
  unsigned long long pairs = 0;
  for (int run = 0; run < 2; ++run) {
    const int L0 = run ? nr.z : nr.x, L1 = run ? nr.w : nr.y;
    for (int Lb = L0; Lb < L1; Lb += kNearBatch) {
      const int nb = min(kNearBatch, L1 - Lb);
      wave_lds_handover();                   // the previous batch's reads are done
      {
        const int e = lane & (kNearBatch - 1), half = lane / kNearBatch;
        if (e < nb) {
          const int L = Lb + e;
#pragma unroll
          for (int k = half; k < kFarP; k += 2) s_q[wave][k][e] = recs[(size_t)(kFarHead + k) * ns + L];
          if (half == 0) {
            s_c[wave][e] = P.far_geom[L].x;
            s_r[wave][e] = recs[L];
          } else {
            s_u[wave][e] = nu_t[L];
          }
        }
      }
      wave_lds_handover();
      for (int e0 = 0; e0 < nb; e0 += NL) {
        const int e = e0 + slot;
        if (e < nb) {
          const double c = s_c[wave][e], r = s_r[wave][e];
          const int2 u = s_u[wave][e];
          // the leaf's series at the lane's bins outside its near range (zero weight inside),
          // moment by moment over all of them
          double id[Q], t[Q], v[Q];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const bool out = jj[q] < u.x || jj[q] >= u.y;
            pairs += out ? kFarLeaf : 0;
            id[q] = out ? rcp_nr(f[q] - c) : 0.0;
            t[q] = r * id[q];
            v[q] = s_q[wave][kFarP - 1][e];
          }
#pragma unroll
          for (int m = kFarP - 2; m >= 0; --m) {
            const double a = s_q[wave][m][e];
#pragma unroll
            for (int q = 0; q < Q; ++q) v[q] = __builtin_fma(v[q], t[q], a);
          }
#pragma unroll
          for (int q = 0; q < Q; ++q) acc[q] += v[q] * t[q] * id[q];
        }
      }
    }
  }
  // the NL leaf slots of each bin, added in a fixed order
#pragma unroll
  for (int d = BL; d < 64; d <<= 1)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] += __shfl_xor(acc[q], d);
  if (slot == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) tk.out[jj[q]] += acc[q];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o);
  if (lane == 0 && pairs) atomicAdd(P.far_evals, pairs);
}

// Region I of lines [a, z) at bin f in FP64, line by line (the same form as the region-I
// kernel's: c y (u + 1) / (u^2 + 2 y^2), u = x^2 + y^2 - 1/2, scaled by amp = S g).
__device__ __forceinline__ double region1_line(double nu, double g, double y, float amp, double f) {
  const double x = (nu_if(nu) - f) * g;
  const double y2 = y * y;
  const double u = x * x + y2 - 0.5;
  return (double)amp * 0.5641896 * y * (u + 1.0) / (u * u + 2.0 * y2);
}

__device__ __forceinline__ double region1_lines(const MixedPlan& P, const MixedTask& tk, int a, int z, double f) {
  double s = 0.0;
  int l = a;
  // eight lines' loads issued together (a thread walking its own lines is bound by their
  // latency), then their terms added in line order; the fallback of leaves_sum when a range's
  // series misses its bound
  for (; l + 8 <= z; l += 8) {
    double nu[8], g[8], y[8];
    float amp[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      nu[i] = P.nu[l + i];
      g[i] = tk.g[l + i];
      y[i] = tk.gamma[l + i] * tk.pscale;
      y[i] *= tk.ts[l + i];
      amp[i] = tk.amp[l + i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) s += region1_line(nu[i], g[i], y[i] * g[i], amp[i], f);
  }
  for (; l < z; ++l) {
    const double g = tk.g[l];
    s += region1_line(P.nu[l], g, mixed_y(tk, l, g), tk.amp[l], f);
  }
  return s;
}

// Region I of the lines [a, z) (every one contributing) at bin f: leaves of the range
// from the largest aligned clusters that serve f, partial leaves line by line.
__device__ __forceinline__ double range_sum(const MixedPlan& P, const MixedTask& tk, const FarVal* __restrict__ recs, bool cplx,
                            int a, int z, double f, int task, int jb) {
  if (z <= a) return 0.0;
  const int A = (a + kFarLeaf - 1) / kFarLeaf, B = z / kFarLeaf;
  if (A >= B) return region1_lines(P, tk, a, z, f);
  double s = region1_lines(P, tk, a, A * kFarLeaf, f);
  int i = A;
  while (i < B) {
    int lev = min(P.far_nlev - 1, i ? __builtin_ctz(i) : 31);
    bool took = false;
    for (; lev >= 0; --lev) {
      if (i + (1 << lev) > B) continue;
      const int k = P.far_off[lev] + (i >> lev);
      const double4 gm = P.far_geom[k];
      const FarRecC rc{recs + k, (size_t)P.far_nclus};
      if (!serves(gm, rc, cplx, f, f, near_u_of(P, task, k), jb, jb)) continue;
      s += cluster_sum(rc, gm.x, cplx, f);
      i += 1 << lev;
      took = true;
      break;
    }
    if (!took) {                 // a leaf that does not serve f: its lines one by one
      s += region1_lines(P, tk, i * kFarLeaf, min((i + 1) * kFarLeaf, z), f);
      ++i;
    }
  }
  return s + region1_lines(P, tk, max(B * kFarLeaf, a), z, f);
}

// ---- window-end ranges from one expansion per (point, block, range) ----
//
// A window-end range's lines lie together, a cutoff away from the block, so one series of
// kRangeP orders about the range's own centre converges fast at every bin of the block.
// Its moments come from the leaves' moments shifted to that centre (exact up to rounding:
// order k of the shifted series takes orders <= k of a leaf's).  Bin j of the range ending
// inside the block needs the whole leaves [A_j, B) (A_j rises with j): thread j shifts the
// leaves [A_j, A_j+1) and a suffix sum over the threads, in a fixed order, gives each its
// [A_j, B); the range starting inside the block is the mirror image (prefix sums).

constexpr int kRangeP = 16;

template <bool CPLX>
struct RangeMoments {
  static constexpr int N = CPLX ? 2 * kRangeP : kRangeP;   // real: orders 1..16; complex: (Re, Im) of 0..15
};

// Leaf moments (record rc, centre nu_l) shifted to the centre cE + i gE with radius rE,
// the first kRangeP orders added to m.
template <bool CPLX>
__device__ __forceinline__ void leaf_to_range(FarRecC rc, double nu_l, double cE, double gE,
                                              double rE, double (&m)[RangeMoments<CPLX>::N]) {
  const double irE = rcp_nr(rE);
  const double b = rc[0] * irE;
  if (!CPLX) {
    const double a = (nu_l - cE) * irE;
    double v[kRangeP + 1];
    v[0] = 0.0;                                                // Q_0 of a real centre is 0
    double bm = b;
#pragma unroll
    for (int k = 1; k <= kRangeP; ++k) {
      v[k] = bm * rc[kFarHead + k - 1];
      bm *= b;
    }
#pragma unroll
    for (int j = 1; j <= kRangeP; ++j)
#pragma unroll
      for (int k = kRangeP; k >= j; --k) v[k] = v[k] + a * v[k - 1];
#pragma unroll
    for (int k = 1; k <= kRangeP; ++k) m[k - 1] += v[k];
  } else {
    const double ar = (nu_l - cE) * irE, ai = (rc[5] - gE) * irE;
    double vr[kRangeP], vi[kRangeP];
    double bm = 1.0;
#pragma unroll
    for (int k = 0; k < kRangeP; ++k) {
      vr[k] = bm * rc[kFarHead + 2 * k];
      vi[k] = bm * rc[kFarHead + 2 * k + 1];
      bm *= b;
    }
#pragma unroll
    for (int j = 1; j < kRangeP; ++j) {
#pragma unroll
      for (int k = kRangeP - 1; k >= j; --k) {
        const double nr = vr[k] + (ar * vr[k - 1] - ai * vi[k - 1]);
        const double ni = vi[k] + (ar * vi[k - 1] + ai * vr[k - 1]);
        vr[k] = nr;
        vi[k] = ni;
      }
    }
#pragma unroll
    for (int k = 0; k < kRangeP; ++k) {
      m[2 * k] += vr[k];
      m[2 * k + 1] += vi[k];
    }
  }
}

// The series of m (centre cE + i gE, radius rE) at bin f.
template <bool CPLX>
__device__ __forceinline__ double range_value(const double (&m)[RangeMoments<CPLX>::N], double cE, double gE,
                                              double rE, double f) {
  if (!CPLX) {
    const double id = 1.0 / (f - cE);
    const double t = rE * id;
    double q = m[kRangeP - 1];
#pragma unroll
    for (int k = kRangeP - 2; k >= 0; --k) q = q * t + m[k];
    return q * t * id;
  }
  const double dr = f - cE, di = -gE;
  const double inv = 1.0 / (dr * dr + di * di);
  const double ir = dr * inv, ii = -di * inv;
  const double tr = rE * ir, ti = rE * ii;
  double sr = m[2 * (kRangeP - 1)], si = m[2 * (kRangeP - 1) + 1];
#pragma unroll
  for (int k = kRangeP - 2; k >= 0; --k) {
    const double nr = (sr * tr - si * ti) + m[2 * k];
    const double ni = (sr * ti + si * tr) + m[2 * k + 1];
    sr = nr;
    si = ni;
  }
  return sr * ii + si * ir;
}

// Maximum (MAX) or minimum of v over the workgroup of NW waves; every thread gets it.
template <int NW, bool MAX>
__device__ __forceinline__ double block_extreme(double v, double* s_red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o);
    v = MAX ? fmax(v, w) : fmin(v, w);
  }
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  v = s_red[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) v = MAX ? fmax(v, s_red[w]) : fmin(v, s_red[w]);
  __syncthreads();
  return v;
}

// Suffix (SUFFIX) or prefix sums of N values over the threads of the workgroup, in a fixed
// order: within each wave by doubling distances, then the totals of the other waves.
template <int N, int NW, bool SUFFIX>
__device__ __forceinline__ void block_scan(double (&v)[N], double* s_tot) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const bool take = SUFFIX ? lane + d < 64 : lane >= d;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double x = SUFFIX ? __shfl_down(v[i], d) : __shfl_up(v[i], d);
      if (take) v[i] += x;
    }
  }
  if (lane == (SUFFIX ? 0 : 63)) {
#pragma unroll
    for (int i = 0; i < N; ++i) s_tot[wave * N + i] = v[i];
  }
  __syncthreads();
  double add[N];
#pragma unroll
  for (int i = 0; i < N; ++i) add[i] = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    if (SUFFIX ? w > wave : w < wave) {
#pragma unroll
      for (int i = 0; i < N; ++i) add[i] += s_tot[w * N + i];
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] += add[i];
  __syncthreads();
}

// Whole leaves [lo, hi) of thread's range at bin f (every thread of the workgroup calls
// this together): SUFFIX, hi is the same for all threads and lo rises with the thread;
// else lo is the same and hi rises.  From one shifted series when it meets the truncation
// bound at every bin of the block, else through the tree bin by bin.
template <int NW, bool CPLX, bool SUFFIX>
__device__ __forceinline__ double leaves_sum(const MixedPlan& P, const MixedTask& tk, const FarVal* __restrict__ recs, int lo,
                             int hi, double f, double f_lo, double f_hi, int* s_cut, double* s_tot, double* s_red,
                             int task, int jb) {
  constexpr int NB = 64 * NW;
  const int tid = threadIdx.x;
  __syncthreads();                                             // s_cut of an earlier call read
  s_cut[tid] = SUFFIX ? lo : hi;
  __syncthreads();
  const int U0 = SUFFIX ? s_cut[0] : lo;                       // the union of the threads' leaves
  const int U1 = SUFFIX ? hi : s_cut[NB - 1];
  const int own0 = SUFFIX ? lo : (tid ? s_cut[tid - 1] : lo);  // the leaves this thread shifts
  const int own1 = SUFFIX ? (tid + 1 < NB ? s_cut[tid + 1] : hi) : hi;
  if (U0 >= U1) return 0.0;
  const double cE = 0.5 * (P.far_geom[U0].y + P.far_geom[U1 - 1].z);
  double glo = 1e300, ghi = -1e300;                            // Lorentz widths: complex centres only
  if (CPLX) {
    for (int L = U0 + tid; L < U1; L += NB) {
      const FarRecC rc{recs + L, (size_t)P.far_nclus};
      glo = fmin(glo, rc[3]);
      ghi = fmax(ghi, rc[4]);
    }
    glo = block_extreme<NW, false>(glo, s_red);
    ghi = block_extreme<NW, true>(ghi, s_red);
  }
  const double gE = CPLX ? 0.5 * (glo + ghi) : 0.0;
  double rE = 0.0;
  for (int L = U0 + tid; L < U1; L += NB) {
    const FarRecC rc{recs + L, (size_t)P.far_nclus};
    const double dx = P.far_geom[L].x - cE, dy = rc[5] - gE;
    rE = fmax(rE, (CPLX ? sqrt(dx * dx + dy * dy) : fabs(dx)) + rc[0]);
  }
  rE = block_extreme<NW, true>(rE, s_red);
  if (!(rE > 0.0)) rE = 1.0;
  const double dist = fmax(fmax(f_lo - cE, cE - f_hi), 0.0);
  if (!series_ok<kRangeP>(rE, gE, glo, dist, CPLX))
    return range_sum(P, tk, recs, CPLX, lo * kFarLeaf, hi * kFarLeaf, f, task, jb);
  double m[RangeMoments<CPLX>::N];
#pragma unroll
  for (int i = 0; i < RangeMoments<CPLX>::N; ++i) m[i] = 0.0;
  for (int L = own0; L < own1; ++L)
    leaf_to_range<CPLX>(FarRecC{recs + L, (size_t)P.far_nclus}, P.far_geom[L].x, cE, gE, rE, m);
  block_scan<RangeMoments<CPLX>::N, NW, SUFFIX>(m, s_tot);
  return range_value<CPLX>(m, cE, gE, rE, f);
}

// Window-end lines, one workgroup per (task, block), one bin per thread: bin j gets the
// lines ending inside the block with ie > j and those starting inside it with ib <= j,
// each range as at most one partial leaf line by line plus whole leaves (leaves_sum).  The
// whole leaves' series are added here, for the tasks of one moment kind (block scans, many
// registers); the partial leaves by k_far_edge_parts, with the same split.
template <int R, bool CPLX>
__global__ __launch_bounds__(64 * R) void k_far_edges(MixedPlan P) {
  constexpr int NB = 64 * R;
  __shared__ int s_cut[NB];
  __shared__ double s_tot[R * RangeMoments<CPLX>::N];
  __shared__ double s_red[R];
  // (block, point) pairs in block-major order, dealt to the XCDs (workgroups go to them in
  // turn) in contiguous runs: an XCD takes the points of one block one after another, so the
  // block's window-end lines (nu for every point, 1/alpha and amp for each temperature) come
  // from its L2 instead of HBM; the launch pads gridDim.x to a multiple of 8
  const int per = gridDim.x / 8;
  const int pair = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (pair >= P.nblocks * P.ntasks) return;
  const int b = pair / P.ntasks;
  const int task = pair - b * P.ntasks;
  const size_t slot = (size_t)task * P.nblocks + b;
  const int4 er = P.edge_rng[slot];
  if (er.x >= er.y && er.z >= er.w) return;                    // set for whole blocks only
  if ((P.far_kind[task] != 0) != CPLX) return;                 // the other kind's launch
  const MixedTask tk = P.tasks[task];
  const GridView gv = P.gv;
  const FarVal* recs = P.far_rec + (size_t)task * P.far_nclus * kFarRec;
  const int w0 = gv.start_idx + b * NB;
  const int j = w0 + threadIdx.x;
  const double f = gv.freq[j], f_lo = gv.freq[w0], f_hi = gv.freq[w0 + NB - 1];
  double acc = 0.0;
  // The partial leaf of a range at bin j: its lines in the range one by one, or, when they are
  // more than half of it, the leaf's series (it joins the whole leaves; every leaf the range
  // touches serves the block) less its other lines one by one.  The threads' first (last) whole
  // leaf still rises with j, as leaves_sum needs: the partial part shrinks (grows) along j.
  if (er.x < er.y) {                                           // er.y is a leaf boundary
    const int s0 = first_above<false>(P, er.x, er.y, j);       // ie > j from here
    int A = (s0 + kFarLeaf - 1) / kFarLeaf;
    if (2 * (A * kFarLeaf - s0) > kFarLeaf) --A;
    acc += leaves_sum<R, CPLX, true>(P, tk, recs, A, er.y / kFarLeaf, f, f_lo, f_hi, s_cut, s_tot, s_red, task, j);
  }
  if (er.z < er.w) {                                           // er.z is a leaf boundary
    const int p1 = first_above<true>(P, er.z, er.w, j);        // ib <= j before here
    int D = p1 / kFarLeaf;
    const int leaf_end = min((D + 1) * kFarLeaf, P.nlines);
    if (2 * (p1 - D * kFarLeaf) > leaf_end - D * kFarLeaf) ++D;
    acc += leaves_sum<R, CPLX, false>(P, tk, recs, er.z / kFarLeaf, D, f, f_lo, f_hi, s_cut, s_tot, s_red, task, j);
  }
  tk.out[j] += acc;
}

// A lane's double from lane i of the wave (wave-uniform result).
__device__ __forceinline__ double readlane_f64(double v, int i) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), i);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), i);
  return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {   // every source lane valid for these controls
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sums over each half (32 lanes) of the wave, fixed order: quad swaps by 1 and 2, row rotations
// by 4 and 8 (every lane of a row of 16 then holds a row total), then rows 0 + 1 and 2 + 3 by the
// 16-lane swap.  Lanes 0 and 32 hold the halves' totals.  Every lane of the wave must be active.
__device__ __forceinline__ double half_sums(double v) {
#pragma clang fp contract(off)
  v = v + dpp_f64<0xB1>(v);               // quad_perm [1, 0, 3, 2]
  v = v + dpp_f64<0x4E>(v);               // quad_perm [2, 3, 0, 1]
  v = v + dpp_f64<0x124>(v);              // row_ror 4
  v = v + dpp_f64<0x128>(v);              // row_ror 8
  float alo = __int_as_float(__double2loint(v)), ahi = __int_as_float(__double2hiint(v));
  float blo = alo, bhi = ahi;
  lane_swap<16>(alo, blo);
  lane_swap<16>(ahi, bhi);
  return __hiloint2double(__float_as_int(ahi), __float_as_int(alo)) +
         __hiloint2double(__float_as_int(bhi), __float_as_int(blo));
}

// Region I of one line at bin f, scaled by amp = S g: c y (u + 1) / (u^2 + 2 y^2) with
// u = x^2 + y^2 - 1/2 (the region-I kernel's form).  x is formed in FP64 and the rational in
// FP32 (u > 100 in region I, so no cancellation: ~1e-7 relative), in FP64 where |x| or y
// reaches 1e9 (there u^2 could leave the FP32 range).
__device__ __forceinline__ double region1_line_mixed(double nu, double g, double y, float amp, double f) {
  const double x = (nu_if(nu) - f) * g;
  if (fabs(x) < 1e9 && y < 1e9) {
    const float xf = (float)x, yf = (float)y;
    const float u = __builtin_fmaf(xf, xf, __builtin_fmaf(yf, yf, -0.5f));
    const float q = (u + 1.0f) * __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, 2.0f * (yf * yf)));
    return (double)((amp * (0.5641896f * yf)) * q);
  }
  return region1_line(nu, g, y, amp, f);
}

// Partial leaves of the window-end ranges.  Bin j's part of the range ending inside the block
// is the line set [a0, z0) of one leaf, of the range starting inside it [a1, z1), each at most
// 32 lines (the complement form keeps it so) and added or, for the complement, subtracted (the
// split of far_edges).  The wave takes its bins one by one: lanes 0-31 evaluate the first set
// of bin i, lanes 32-63 the second, with coalesced loads of consecutive lines; the two halves
// are summed over their lanes in a fixed order and added to bin i's thread.  (One bin per
// thread walking its own lines made every load a 64-line gather: 16 ms per 48-point table,
// bound by address processing.)
template <int R>
__global__ __launch_bounds__(64 * R) void k_far_edge_parts(MixedPlan P) {
  constexpr int NB = 64 * R;
  const int per = gridDim.x / 8;                               // as far_edges maps them
  const int pair = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (pair >= P.nblocks * P.ntasks) return;
  const int b = pair / P.ntasks;
  const int task = pair - b * P.ntasks;
  const int4 er = P.edge_rng[(size_t)task * P.nblocks + b];
  if (er.x >= er.y && er.z >= er.w) return;                    // set for whole blocks only
  const MixedTask tk = P.tasks[task];
  const int lane = threadIdx.x & 63;
  const int j = P.gv.start_idx + b * NB + threadIdx.x;
  int a0 = 0, z0 = 0, a1 = 0, z1 = 0;
  bool n0 = false, n1 = false;
  unsigned long long pairs = 0;
  if (er.x < er.y) {                                           // er.y is a leaf boundary
    const int s0 = first_above<false>(P, er.x, er.y, j);       // ie > j from here
    const int A = (s0 + kFarLeaf - 1) / kFarLeaf;
    pairs += (unsigned long long)(er.y - s0);
    n0 = 2 * (A * kFarLeaf - s0) > kFarLeaf;
    a0 = n0 ? (A - 1) * kFarLeaf : s0;
    z0 = n0 ? s0 : A * kFarLeaf;
  }
  if (er.z < er.w) {                                           // er.z is a leaf boundary
    const int p1 = first_above<true>(P, er.z, er.w, j);        // ib <= j before here
    const int D = p1 / kFarLeaf;
    const int leaf_end = min((D + 1) * kFarLeaf, P.nlines);
    pairs += (unsigned long long)(p1 - er.z);
    n1 = 2 * (p1 - D * kFarLeaf) > leaf_end - D * kFarLeaf;
    a1 = n1 ? p1 : D * kFarLeaf;
    z1 = n1 ? leaf_end : p1;
  }
  const double f = P.gv.freq[j];
  const bool second = lane >= 32;
  const int hl = lane & 31;
  double acc = 0.0;
  for (int i = 0; i < 64; ++i) {
    const int sa0 = __builtin_amdgcn_readlane(a0, i), sz0 = __builtin_amdgcn_readlane(z0, i);
    const int sa1 = __builtin_amdgcn_readlane(a1, i), sz1 = __builtin_amdgcn_readlane(z1, i);
    if (sa0 >= sz0 && sa1 >= sz1) continue;                    // wave-uniform
    const int l = (second ? sa1 : sa0) + hl;
    double t = 0.0;
    if (l < (second ? sz1 : sz0)) {
      const double fi = readlane_f64(f, i);
      const double g = tk.g[l];
      t = region1_line_mixed(P.nu[l], g, mixed_y(tk, l, g), tk.amp[l], fi);
    }
    t = half_sums(t);
    const double t0 = readlane_f64(t, 0), t1 = readlane_f64(t, 32);
    if (lane == i) acc = (n0 ? -t0 : t0) + (n1 ? -t1 : t1);    // each bin's sets once
  }
  P.tasks[task].out[j] += acc;
  // pairs of this wave, counted once
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) pairs += __shfl_xor(pairs, o);
  if (lane == 0 && pairs) atomicAdd(P.far_evals, pairs);
}

}  // namespace

hipError_t launch_far_setup(const MixedPlan& plan, hipStream_t s) {
  if (!plan.far || plan.ntasks <= 0) return hipSuccess;
  if (plan.far_bin_cap < plan.gv.points + 1) return hipErrorInvalidValue;   // sized for another grid
  k_far_geom<<<(plan.far_nclus + 255) / 256, 256, 0, s>>>(plan, plan.far_geom);
  k_far_reset<<<1, 1, 0, s>>>(plan.far_evals);
  k_far_kind<<<plan.ntasks, 256, 0, s>>>(plan);
  k_far_leaf<<<dim3(((plan.far_n[0] + 3) / 4 + 7) / 8 * 8, (plan.ntasks + kLeafTasks - 1) / kLeafTasks), 256, 0, s>>>(plan);
  for (int lev = 1; lev < plan.far_nlev; ++lev) {   // one launch per moment kind: the complex
    const dim3 g((plan.far_n[lev] + 255) / 256, plan.ntasks);   // shifts hold far more registers
    k_far_up<false><<<g, 256, 0, s>>>(plan, lev);
    k_far_up<true><<<g, 256, 0, s>>>(plan, lev);
  }
  k_bin_first<<<(plan.gv.points + 1 + 255) / 256, 256, 0, s>>>(plan);
  k_far_plan<<<dim3((plan.nblocks + 255) / 256, plan.ntasks), 256, 0, s>>>(plan);
  k_edge_plan<<<dim3((plan.nblocks + 3) / 4, plan.ntasks), 256, 0, s>>>(plan);
  return hipGetLastError();
}

template <int R>
dim3 near_grid(const MixedPlan& plan) {
  return dim3((plan.nblocks + 3) / 4, plan.ntasks);
}

hipError_t launch_far_eval(const MixedPlan& plan, int R, hipStream_t s) {
  if (!plan.far || plan.ntasks <= 0) return hipSuccess;
  const dim3 grid((plan.nblocks + 3) / 4, plan.ntasks);
  const dim3 grid_blocks(((unsigned)plan.nblocks * (unsigned)plan.ntasks + 7u) / 8u * 8u);   // k_far_edges
  switch (R) {
    case 1:
      k_far_eval<1, false><<<grid, 256, 0, s>>>(plan);
      k_far_eval<1, true><<<grid, 256, 0, s>>>(plan);
      if (plan.near) k_near_eval<1><<<near_grid<1>(plan), 256, 0, s>>>(plan);
      k_far_edge_parts<1><<<grid_blocks, 64, 0, s>>>(plan);
      k_far_edges<1, false><<<grid_blocks, 64, 0, s>>>(plan);
      k_far_edges<1, true><<<grid_blocks, 64, 0, s>>>(plan);
      break;
    case 2:
      k_far_eval<2, false><<<grid, 256, 0, s>>>(plan);
      k_far_eval<2, true><<<grid, 256, 0, s>>>(plan);
      if (plan.near) k_near_eval<2><<<near_grid<2>(plan), 256, 0, s>>>(plan);
      k_far_edge_parts<2><<<grid_blocks, 128, 0, s>>>(plan);
      k_far_edges<2, false><<<grid_blocks, 128, 0, s>>>(plan);
      k_far_edges<2, true><<<grid_blocks, 128, 0, s>>>(plan);
      break;
    case 4:
      k_far_eval<4, false><<<grid, 256, 0, s>>>(plan);
      k_far_eval<4, true><<<grid, 256, 0, s>>>(plan);
      if (plan.near) k_near_eval<4><<<near_grid<4>(plan), 256, 0, s>>>(plan);
      k_far_edge_parts<4><<<grid_blocks, 256, 0, s>>>(plan);
      k_far_edges<4, false><<<grid_blocks, 256, 0, s>>>(plan);
      k_far_edges<4, true><<<grid_blocks, 256, 0, s>>>(plan);
      break;
    default:
      k_far_eval<8, false><<<grid, 256, 0, s>>>(plan);
      k_far_eval<8, true><<<grid, 256, 0, s>>>(plan);
      if (plan.near) k_near_eval<8><<<near_grid<8>(plan), 256, 0, s>>>(plan);
      k_far_edge_parts<8><<<grid_blocks, 512, 0, s>>>(plan);
      k_far_edges<8, false><<<grid_blocks, 512, 0, s>>>(plan);
      k_far_edges<8, true><<<grid_blocks, 512, 0, s>>>(plan);
      break;
  }
  return hipGetLastError();
}

}  // namespace gx
