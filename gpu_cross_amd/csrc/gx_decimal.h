// gx_decimal.h -- whitespace-separated tokens of ExoMol .trans text parsed the way
// `std::istream >> int >> int >> double` parses them (ExomolStateReader.cpp:145), on the
// GPU and on the host.
//
// Only the common, unambiguous forms are converted here; every other token is handed
// back (GX_TOK_HOST) so the caller parses that record with an istream itself:
//   int     [+-]?[0-9]{1,9}, then whitespace or the end of the text
//   double  [+-]?([0-9]+(\.[0-9]*)?|\.[0-9]+)([eE][+-]?[0-9]+)?, then whitespace or the
//           end, at most 19 significant digits, a normal binary64 result
// A clean token reads exactly as the stream reads it; the double is the correctly
// rounded value strtod returns (the stream converts with strtod).
//
// Decimal -> binary64: value = w 10^q with w < 10^19.  w (normalised to 64 bits) times
// the 128-bit truncation t_q of 5^q (gx_pow5.h) gives the top 128 bits of the exact
// product to within 2 units of the last kept bit, below it (both truncations err low).
// Rounding to 53 bits then only depends on where the discarded tail lies relative to
// one half: a tail within 2 units of the half point is handed back to the host, every
// other tail rounds the same way as the exact value (Lemire, "Number parsing at a
// gigabyte per second", Softw. Pract. Exp. 51 (2021), with this conservative test in
// place of its exact-halfway analysis).
#pragma once

#include <stdint.h>

#include "gx_pow5.h"

#if defined(__HIPCC__)
#define GX_HD __host__ __device__
#else
#define GX_HD
#endif

namespace gx {

enum { GX_TOK_OK = 0, GX_TOK_HOST = 1 };

GX_HD inline bool dec_is_space(unsigned char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
GX_HD inline bool dec_is_digit(unsigned char c) { return c >= '0' && c <= '9'; }

GX_HD inline void dec_mul64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  *lo = a * b;
  *hi = __umul64hi(a, b);
#else
  const unsigned __int128 p = (unsigned __int128)a * b;
  *lo = (uint64_t)p;
  *hi = (uint64_t)(p >> 64);
#endif
}

// w 10^q -> binary64 bits (positive); GX_TOK_HOST outside the normal range or when the
// rounding cannot be decided from the 128-bit product.
GX_HD inline int dec_to_binary64(uint64_t w, int q, const uint64_t* words, const int* shifts, uint64_t* bits) {
  if (w == 0) {
    *bits = 0;
    return GX_TOK_OK;
  }
  if (q < GX_POW5_MIN_Q || q > GX_POW5_MAX_Q) return GX_TOK_HOST;
  const int lz = __builtin_clzll(w);
  const uint64_t wn = w << lz;
  const int k = q - GX_POW5_MIN_Q;
  uint64_t a_hi, a_lo, b_hi, b_lo;
  dec_mul64(wn, words[2 * k], &a_hi, &a_lo);
  dec_mul64(wn, words[2 * k + 1], &b_hi, &b_lo);
  const uint64_t plo = a_lo + b_hi;
  const uint64_t phi = a_hi + (plo < a_lo ? 1u : 0u);
  // (phi:plo) = 1.f 2^(126 + up); value = (phi:plo) 2^(64 + q - lz - s_q)
  const int up = (int)(phi >> 63);
  const int sh = 10 + up;                            // phi bits below the 53-bit significand
  uint64_t m = phi >> sh;
  const uint64_t tail = phi & ((1ull << sh) - 1);
  const uint64_t half = 1ull << (sh - 1);
  if (tail == half && plo <= 2) return GX_TOK_HOST;
  if (tail == half - 1 && plo >= ~0ull - 2) return GX_TOK_HOST;
  int e2 = 190 + up + q - lz - shifts[k];            // unbiased exponent of 1.f
  if (tail >= half) {
    m += 1;
    if (m == (1ull << 53)) {
      m >>= 1;
      e2 += 1;
    }
  }
  const int be = e2 + 1023;
  if (be <= 0 || be >= 2047) return GX_TOK_HOST;     // subnormal or overflow: the stream decides
  *bits = ((uint64_t)be << 52) | (m & ((1ull << 52) - 1));
  return GX_TOK_OK;
}

// One int token starting at p (non-space), text ends at end.
GX_HD inline int dec_parse_int(const char* p, const char* end, int* v) {
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  int n = 0;
  int64_t val = 0;
  while (p < end && dec_is_digit((unsigned char)*p)) {
    if (n < 10) val = val * 10 + (*p - '0');
    ++n;
    ++p;
  }
  if (n == 0 || n > 9) return GX_TOK_HOST;
  if (p < end && !dec_is_space((unsigned char)*p)) return GX_TOK_HOST;
  *v = (int)(neg ? -val : val);
  return GX_TOK_OK;
}

// One double token starting at p (non-space), text ends at end.
GX_HD inline int dec_parse_double(const char* p, const char* end, const uint64_t* words, const int* shifts,
                                  double* v) {
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) {
    neg = *p == '-';
    ++p;
  }
  uint64_t w = 0;
  int nd = 0, q = 0;
  bool any = false, lost = false;
  while (p < end && dec_is_digit((unsigned char)*p)) {
    const int d = *p - '0';
    any = true;
    if (w == 0 && d == 0) {
    } else if (nd < 19) {
      w = w * 10 + d;
      ++nd;
    } else {
      ++q;
      lost |= d != 0;
    }
    ++p;
  }
  if (p < end && *p == '.') {
    ++p;
    while (p < end && dec_is_digit((unsigned char)*p)) {
      const int d = *p - '0';
      any = true;
      if (w == 0 && d == 0) {
        --q;
      } else if (nd < 19) {
        w = w * 10 + d;
        ++nd;
        --q;
      } else {
        lost |= d != 0;
      }
      ++p;
    }
  }
  if (!any) return GX_TOK_HOST;
  if (p < end && (*p == 'e' || *p == 'E')) {
    ++p;
    bool eneg = false;
    if (p < end && (*p == '+' || *p == '-')) {
      eneg = *p == '-';
      ++p;
    }
    if (!(p < end && dec_is_digit((unsigned char)*p))) return GX_TOK_HOST;
    int e = 0;
    while (p < end && dec_is_digit((unsigned char)*p)) {
      if (e < 100000) e = e * 10 + (*p - '0');
      ++p;
    }
    q += eneg ? -e : e;
  }
  if (p < end && !dec_is_space((unsigned char)*p)) return GX_TOK_HOST;
  if (lost) return GX_TOK_HOST;
  uint64_t bits;
  if (dec_to_binary64(w, q, words, shifts, &bits) != GX_TOK_OK) return GX_TOK_HOST;
  if (neg) bits |= 1ull << 63;
  union {
    uint64_t u;
    double d;
  } cv;
  cv.u = bits;
  *v = cv.d;
  return GX_TOK_OK;
}

}  // namespace gx
