// gx_plain.h -- device profile functions of the two sums the reference has only in its
// CUDA manager: the fixed-half-width Gaussian (cross_kernal.cu:567-625) and the Voigt sum
// of a line list uploaded without per-line Lorentz widths (cross_kernal.cu:860-922), whose
// profile is Zaghloul & Ali's Algorithm 916 (ACM TOMS 38 (2011) 15) with the truncation the
// reference uses, or a 3-point Gauss-Hermite expression far from the centre.
//
// FP64 throughout, with contraction off where the reference's plain arithmetic rounds each
// product, so that the evaluation order is the reference's.  erf, erfc, exp, sin and cos are
// the device library's.
#pragma once

#include <hip/hip_runtime.h>

namespace gx {

// defines.h:4-7,14 literals
constexpr double kPlainLn2 = 0.69314718056;
constexpr double kPlainSqrtLn2 = 0.83255461115;
constexpr double kPlainSqrtLn2Pi = 0.46971863935;
constexpr double kPlainPi = 3.14159265359;
// Lorentz half-width and window of the sum without per-line widths: gamma_L = 0.05 cm^-1,
// window 500 gamma_L (cross_kernal.cu:870,885-888)
constexpr double kNoGammaL = 0.05;

// erfcx(y) = exp(y^2) erfc(y), y >= 0.  Below 26: y^2 = h + l exactly (Dekker's product via
// fma), exp(y^2) = exp(h) (1 + l) to first order in |l| <= ulp(h)/2, times erfc(y) (still a
// normal number at 26).  From 26 up: Laplace's continued fraction
//   sqrt(pi) erfcx(y) = 1/(y + (1/2)/(y + 1/(y + (3/2)/(y + 2/(y + ...)))))
// (Abramowitz & Stegun 7.1.14) from 40 levels down, converged to double precision there.
__device__ __noinline__ double plain_erfcx(double y) {
#pragma clang fp contract(off)
  if (y < 26.0) {
    const double h = y * y;
    const double l = __builtin_fma(y, y, -h);
    return exp(h) * (1.0 + l) * erfc(y);
  }
  double t = y;
  for (int k = 40; k >= 1; --k) t = y + (0.5 * k) / t;
  return 1.0 / (1.7724538509055159 * t);   // sqrt(pi)
}

// V(x, y) of Alg. 916 (equation 13), with the sigma series truncated as cross_kernal.cu:61-147
// truncates it: n = 1..4, the third series centred on n0 = ceil(x/a).  Callers pass x >= 0.
__device__ __noinline__ double plain_voigt916(double x, double y, double a) {
#pragma clang fp contract(off)
  const double ex2 = exp(-x * x);
  if (x == 0) return plain_erfcx(y);
  if (y == 0) return ex2;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  const double yy = y * y;
  const int n0 = (int)ceil(x / a);
  for (int n = 1; n < 5; ++n) {
    const int n3p = n0 + n - 1, n3n = n0 - n;
    const double an = a * n, an3p = a * n3p, an3n = a * n3n;
    const double f = 1.0 / (an * an + yy);
    const double f3p = 1.0 / (an3p * an3p + yy);
    const double f3n = 1.0 / (an3n * an3n + yy);
    s1 += f * exp(-(an * an + x * x));
    s2 += f * exp(-(an + x) * (an + x));
    s3 += f3p * exp(-(an3p - x) * (an3p - x));
    if (n3n >= 1) s3 += f3n * exp(-(an3n - x) * (an3n - x));
  }
  const double xy = x * y;
  const double a2ipi = 2.0 * a / 3.141592653589793;   // M_PI
  const double cos2xy = cos(2.0 * xy);
  const double sinxy = sin(xy);
  double t1 = ex2 * plain_erfcx(y) * cos2xy;
  t1 += a2ipi * x * sinxy * ex2 * sinxy / xy;
  t1 += a2ipi * y * (-cos2xy * s1 + 0.5 * (s2 + s3));
  return t1;
}

// The 3-point Gauss-Hermite expression of cross_kernal.cu:183-186, constants as written.
__device__ __forceinline__ double plain_gh3(double x, double y, double xxyy) {
#pragma clang fp contract(off)
  return 1.1181635900 * y / (kPlainPi * xxyy) +
         2.0 * 0.2954089751 * (xxyy + 1.5) / ((xxyy + 1.5) * (xxyy + 1.5) - 4 * x * x * 1.5);
}

// One (line, bin) term of the no-gamma Voigt sum (cross_kernal.cu:884-904): dnu = nu - f,
// gG = nu sqrt(2 ln2 k T / m) / c (per line), y = sqrt(ln2) gamma_L / gG (per line).  The
// series is evaluated at |x| (the reference's signed x adds n <= 0 terms above the centre).
__device__ __forceinline__ double plain_nogamma_term(double S, double dnu, double gG, double y) {
#pragma clang fp contract(off)
  const double x = kPlainSqrtLn2 * dnu / gG;
  const double xxyy = x * x + y * y;
  if (xxyy < 100.0) return S * kPlainSqrtLn2Pi / (gG) * y * plain_voigt916(fabs(x), y, 1.0);
  return S * (kPlainSqrtLn2Pi / gG) * plain_gh3(x, y, xxyy);
}

// One (line, bin) term of the Gaussian sum (cross_kernal.cu:594-604): dfr = f - nu,
// x_ = sqrt(ln2) / HW, x0 = x_ dnu_bin / 2.  Both erf saturate to +-1 beyond |x| = 6.
__device__ __forceinline__ double plain_gauss_term(double S, double dfr, double x_, double x0) {
#pragma clang fp contract(off)
  const double xp = x_ * (dfr) + x0;
  const double xm = x_ * (dfr) - x0;
  if (xm >= 6.0 || xp <= -6.0) return 0.0;
  return S * (erf(xp) - erf(xm));
}

}  // namespace gx
