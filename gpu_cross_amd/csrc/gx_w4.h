// gx_w4.h -- device evaluation of Humlicek's W4 approximation to Re w(x+iy)
// (J. Humlicek, JQSRT 27 (1982) 437), the reference's Voigt function
// (profiles.cpp:9-42, cross_kernal.cu:149-181).
//
// The region tests (|x|+y >= 15, >= 5.5, y >= 0.195|x| - 0.176) and the
// published coefficients are the algorithm itself and are kept as they are.
// The arithmetic is written for CDNA: only the real part is formed, region I
// uses its closed real form
//     Re = c y (0.5 + x^2 + y^2) / ((0.5 + y^2 - x^2)^2 + 4 x^2 y^2),
// and regions II-IV use explicit real/imaginary pairs with a single real
// division for the final quotient.
#pragma once

#include <hip/hip_runtime.h>

namespace gx {

struct cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cadd(double r, cplx a) { return {r + a.re, a.im}; }
__device__ __forceinline__ cplx csub(double r, cplx a) { return {r - a.re, -a.im}; }
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ double cdiv_re(cplx a, cplx b) {
  return (a.re * b.re + a.im * b.im) / (b.re * b.re + b.im * b.im);
}

// num / den for den > 0 and normal: the hardware reciprocal refined by two Newton steps
// (relative error squared each step), the quotient by one residual correction
// q + (num - den q) / den; within an ulp or two of the IEEE quotient, without the scaling
// steps an IEEE division spends on operands that cannot occur here.
__device__ __forceinline__ double div_pos(double num, double den) {
  double r = __builtin_amdgcn_rcp(den);
  r = __builtin_fma(r, __builtin_fma(-den, r, 1.0), r);
  r = __builtin_fma(r, __builtin_fma(-den, r, 1.0), r);
  const double q = num * r;
  return __builtin_fma(__builtin_fma(-den, q, num), r, q);
}

// Region I (|x| + y >= 15) in closed real form, FP64.  The denominator is positive there
// ((x^2 - y^2 - 1/2)^2 + 4 x^2 y^2 with |x| + y >= 15).
__device__ __forceinline__ double w4_region1(double x, double y) {
  const double x2 = x * x, y2 = y * y;
  const double dr = 0.5 + y2 - x2;
  return div_pos(0.5641896 * y * (0.5 + x2 + y2), dr * dr + 4.0 * x2 * y2);
}

// The region tests of W4, each operand rounded as the reference's plain C rounds it (no
// multiply-add fused into them): x + y against 15 and 5.5, y against 0.195 x - 0.176.
__device__ __forceinline__ double w4_sum(double x, double y) {   // x + y: regions I and II
#pragma clang fp contract(off)
  return x + y;
}

__device__ __forceinline__ bool w4_region3(double x, double y) {   // region III (given not II)
#pragma clang fp contract(off)
  const double b = 0.195 * x - 0.176;
  return y >= b;
}

// Regions II-IV, FP64 (x >= 0 assumed by the callers; only |x| enters).
__device__ __noinline__ double w4_inner(double x, double y) {
  const cplx t{y, -x};
  if (w4_sum(x, y) >= 5.5) {  // region II
    const cplx u = cmul(t, t);
    const cplx num = cmul(t, cadd(1.410474, cscale(u, 0.5641896)));
    const cplx den = cadd(0.75, cmul(u, cadd(3.0, u)));
    return cdiv_re(num, den);
  }
  if (w4_region3(x, y)) {  // region III
    cplx n = cscale(t, 0.5642236);
    n = cmul(t, cadd(3.778987, n));
    n = cmul(t, cadd(11.96482, n));
    n = cmul(t, cadd(20.20933, n));
    n = cadd(16.4955, n);
    cplx d = cadd(6.699398, t);
    d = cadd(21.69274, cmul(t, d));
    d = cadd(39.27121, cmul(t, d));
    d = cadd(38.82363, cmul(t, d));
    d = cadd(16.4955, cmul(t, d));
    return cdiv_re(n, d);
  }
  // region IV: exp(u) - t P(u) / Q(u)
  const cplx u = cmul(t, t);
  cplx p = cscale(u, 0.56419);
  p = csub(1.320522, p);
  p = csub(35.76683, cmul(u, p));
  p = csub(219.0313, cmul(u, p));
  p = csub(1540.787, cmul(u, p));
  p = csub(3321.9905, cmul(u, p));
  p = csub(36183.31, cmul(u, p));
  cplx q = csub(1.841439, u);
  q = csub(61.57037, cmul(u, q));
  q = csub(364.2191, cmul(u, q));
  q = csub(2186.181, cmul(u, q));
  q = csub(9022.228, cmul(u, q));
  q = csub(24322.84, cmul(u, q));
  q = csub(32066.6, cmul(u, q));
  double sn, cs;
  sincos(u.im, &sn, &cs);
  return exp(u.re) * cs - cdiv_re(cmul(t, p), q);
}

// Full W4, FP64.  x >= 0.
__device__ __forceinline__ double w4(double x, double y) {
  if (w4_sum(x, y) >= 15.0) return w4_region1(x, y);
  return w4_inner(x, y);
}

// ---------------------------------------------------------------------------
// FP32 forms used by the mixed-precision kernels.
// ---------------------------------------------------------------------------

// Region II of W4, Re of t (a + b u) / (0.75 + u (3 + u)) with t = y - i x, u = t^2,
// a = 1.410474, b = 0.5641896, as a rational in s = x^2 (the real part is even in x).
// With Y = y^2, N = t (a + b u) and D = 0.75 + u (3 + u):
//   Re D = s^2 + (-3 - 6Y) s + (0.75 + 3Y + Y^2),   Im D / x = 4y s - 2y (3 + 2Y),
//   Re N = y (a + bY) - 3by s,                      Im N / x = b s - (a + 3bY),
//   Re(N / D) = (Re N Re D + s (Im N / x)(Im D / x)) / (Re D^2 + s (Im D / x)^2).
// The coefficients depend on the line only: formed in FP64 once per line and rounded
// to FP32; per bin 14 FP32 operations (max 1.9e-6 relative against FP64 over region II,
// the complex form in FP32 1.7e-6).
struct W4Region2 {
  float c0, c1, d0, d1, n0, n1, m0;
  W4Region2() = default;
  __device__ __forceinline__ explicit W4Region2(double y) {
    const double a = 1.410474, b = 0.5641896, Y = y * y;
    c0 = (float)(0.75 + 3.0 * Y + Y * Y);
    c1 = (float)(-3.0 - 6.0 * Y);
    d0 = (float)(-2.0 * y * (3.0 + 2.0 * Y));
    d1 = (float)(4.0 * y);
    n0 = (float)(y * (a + b * Y));
    n1 = (float)(-3.0 * b * y);
    m0 = (float)(-(a + 3.0 * b * Y));
  }
  __device__ __forceinline__ float operator()(float x) const {
    const float s = x * x;
    const float dr = (s + c1) * s + c0;
    const float e = d1 * s + d0;                    // Im D / x
    const float nr = n1 * s + n0;
    const float ni = 0.5641896f * s + m0;           // Im N / x
    const float se = s * e;
    return (nr * dr + se * ni) * __builtin_amdgcn_rcpf(dr * dr + se * e);
  }
};

// Region III of W4, Re N(t)/D(t) with t = y - i x and the real-coefficient polynomials
//   N = 16.4955 + 20.20933 t + 11.96482 t^2 + 3.778987 t^3 + 0.5642236 t^4,
//   D = 16.4955 + 38.82363 t + 39.27121 t^2 + 21.69274 t^3 + 6.699398 t^4 + t^5,
// in FP32.  t and conj t are the roots of v^2 - 2y v + (x^2 + y^2), so each polynomial
// at t equals its remainder b1 t + b0 of the division by that quadratic (the
// recurrence below, 2 FMA per coefficient); then N = (b1 y + b0) - i b1 x and likewise
// D, and Re N/D = (Re N Re D + b1 d1 x^2) / (Re D^2 + d1^2 x^2).  About 26 operations
// instead of the complex Horner form's 39; max 2.3e-6 relative against FP64 over region III.
__device__ __forceinline__ float w4_region3_f32(float x, float y) {
  const float x2 = x * x, s = 2.0f * y, m = __builtin_fmaf(y, y, x2);
  auto rem = [&](const float* c, int n, float& b1, float& b0) {
    float p1 = c[n], p2 = 0.0f;
#pragma unroll
    for (int k = n - 1; k >= 1; --k) {
      const float q = __builtin_fmaf(s, p1, __builtin_fmaf(-m, p2, c[k]));
      p2 = p1;
      p1 = q;
    }
    b1 = p1;
    b0 = __builtin_fmaf(-m, p2, c[0]);
  };
  constexpr float kN[5] = {16.4955f, 20.20933f, 11.96482f, 3.778987f, 0.5642236f};
  constexpr float kD[6] = {16.4955f, 38.82363f, 39.27121f, 21.69274f, 6.699398f, 1.0f};
  float n1, n0, d1, d0;
  rem(kN, 4, n1, n0);
  rem(kD, 5, d1, d0);
  const float rn = __builtin_fmaf(n1, y, n0), rd = __builtin_fmaf(d1, y, d0);
  const float num = __builtin_fmaf(rn, rd, (n1 * d1) * x2);
  const float den = __builtin_fmaf(rd, rd, (d1 * d1) * x2);
  return num * __builtin_amdgcn_rcpf(den);
}

// W4 regions II-IV for the mixed-precision core bins (|x| + y < 15), x >= 0, of the
// line whose region-II coefficients are r2.  Regions II and III are rational in t with
// no cancellation and run in FP32 (max 2.3e-6 relative against FP64 over their domains);
// region IV keeps the FP64 t P(u)/Q(u) term, whose real part cancels in FP32 near
// |x| ~ 5, and takes exp(u) = e^(y^2-x^2)(cos 2xy - i sin 2xy) from the FP32 hardware
// exp/cos (2.7e-6 max).  The region is decided in FP64 by the caller's x, y.
__device__ __forceinline__ float w4_core_mixed(double xd, double yd, const W4Region2& r2) {
  const float x = (float)xd, y = (float)yd;
  if (w4_sum(xd, yd) >= 5.5) return r2(x);        // region II
  if (w4_region3(xd, yd)) return w4_region3_f32(x, y);   // region III
  // region IV: exp(u) - t P(u) / Q(u), u = t^2.  P and Q have real coefficients, so at
  // complex u each is the remainder of its division by (v - u)(v - conj u) = v^2 - s v + m,
  // s = 2 Re u, m = |u|^2: b_k = c_k + s b_(k+1) - m b_(k+2) down to k = 1, then
  // A(u) = b_1 u + (c_0 - m b_2); 2 FMA per coefficient instead of a complex multiply
  // (max 7.9e-14 relative over region IV against an extended-precision evaluation).
  const double ur = yd * yd - xd * xd, ui = -2.0 * xd * yd;
  const double s = 2.0 * ur, m = ur * ur + ui * ui;
  auto rem = [&](const double* c, int n, double& re, double& im) {
    double b1 = c[n], b2 = 0.0;
#pragma unroll
    for (int k = n - 1; k >= 1; --k) {
      const double b0 = __builtin_fma(s, b1, __builtin_fma(-m, b2, c[k]));
      b2 = b1;
      b1 = b0;
    }
    re = __builtin_fma(b1, ur, __builtin_fma(-m, b2, c[0]));
    im = b1 * ui;
  };
  constexpr double kP[7] = {36183.31, -3321.9905, 1540.787, -219.0313, 35.76683, -1.320522, 0.56419};
  constexpr double kQ[8] = {32066.6, -24322.84, 9022.228, -2186.181, 364.2191, -61.57037, 1.841439, -1.0};
  double pr, pi, qr, qi;
  rem(kP, 6, pr, pi);
  rem(kQ, 7, qr, qi);
  const double nr = yd * pr + xd * pi, ni = yd * pi - xd * pr;    // t P, t = y - i x
  const double qq = qr * qr + qi * qi;
  double rq = __builtin_amdgcn_rcp(qq);
  rq = rq * __builtin_fma(-qq, rq, 2.0);          // one Newton step on the hardware reciprocal
  const double tpq = (nr * qr + ni * qi) * rq;
  const float ex = __expf((float)ur) * __cosf((float)ui);
  return (float)((double)ex - tpq);
}

// W4 at a bin of a line's near range in the mixed core pass: region I (|x| + y >= 15, the FP64
// decision) in its closed real form in FP32, c y (u + 1) / (u^2 + 2 y^2) with u = x^2 + y^2 - 1/2
// (no cancellation there: u > 100; x stays below 1e6 in a near range, so u^2 stays finite), and
// regions II-IV as w4_core_mixed.
__device__ __forceinline__ float w4_near_mixed(double xd, double yd, const W4Region2& r2) {
  if (w4_sum(xd, yd) >= 15.0) {
    const float x = (float)xd, y = (float)yd;
    const float u = __builtin_fmaf(x, x, __builtin_fmaf(y, y, -0.5f));
    return (0.5641896f * y) * (u + 1.0f) * __builtin_amdgcn_rcpf(__builtin_fmaf(u, u, 2.0f * (y * y)));
  }
  return w4_core_mixed(xd, yd, r2);
}

}  // namespace gx
