// decimal_check -- the token parsers of gx_decimal.h (the ones the ingest kernels run)
// against the C library and the C++ stream on millions of generated tokens.
//
// Every token the parser accepts must read to the same bits as strtod / std::istream
// (the reference reads .trans files with `stream >> id_f >> id_i >> aif`,
// ExomolStateReader.cpp:145); tokens it hands back are counted, and their share must stay
// small for the formats ExoMol writes.  Exit status 0 = no mismatch.
//
//   decimal_check [count] [seed]
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <sstream>
#include <string>

#include "gx_decimal.h"

static const uint64_t kWords[] = {GX_POW5_WORDS};
static const int kShifts[] = {GX_POW5_SHIFTS};

static uint64_t bits_of(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
}

struct Stats {
  long long n = 0, ok = 0, host = 0, bad = 0;
};

static void check_double(const std::string& s, Stats& st, const char* kind) {
  double v = 0;
  ++st.n;
  const int rc = gx::dec_parse_double(s.data(), s.data() + s.size(), kWords, kShifts, &v);
  if (rc != gx::GX_TOK_OK) {
    ++st.host;
    return;
  }
  ++st.ok;
  std::istringstream is(s);
  double r = -1;
  is >> r;
  const double c = std::strtod(s.c_str(), nullptr);
  if (is.fail() || bits_of(r) != bits_of(v) || bits_of(c) != bits_of(v)) {
    if (st.bad < 20)
      std::printf("MISMATCH %s token '%s': parsed %.17g (%016" PRIx64 ") stream %.17g strtod %.17g\n", kind,
                  s.c_str(), v, bits_of(v), r, c);
    ++st.bad;
  }
}

static void check_int(const std::string& s, Stats& st) {
  int v = 0;
  ++st.n;
  if (gx::dec_parse_int(s.data(), s.data() + s.size(), &v) != gx::GX_TOK_OK) {
    ++st.host;
    return;
  }
  ++st.ok;
  std::istringstream is(s);
  int r = -1;
  is >> r;
  if (is.fail() || r != v) {
    if (st.bad < 20) std::printf("MISMATCH int token '%s': parsed %d stream %d\n", s.c_str(), v, r);
    ++st.bad;
  }
}

int main(int argc, char** argv) {
  const long long count = argc > 1 ? std::atoll(argv[1]) : 2000000;
  const unsigned seed = argc > 2 ? (unsigned)std::atoi(argv[2]) : 12345u;
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  char buf[512];
  Stats exo, fmt, digits, near, ints;

  for (long long i = 0; i < count; ++i) {
    // ExoMol Einstein A columns: %10.4E / %10.3e of log-uniform values
    const double a = std::pow(10.0, -45.0 + 50.0 * u01(rng));
    std::snprintf(buf, sizeof buf, (i & 1) ? "%10.4E" : "%10.3e", a);
    const char* p = buf;
    while (*p == ' ') ++p;
    check_double(p, exo, "exomol");

    // random binary64 values printed with 1..17 significant digits, both styles
    uint64_t u = rng();
    double d;
    std::memcpy(&d, &u, 8);
    if (std::isfinite(d)) {
      const int prec = 1 + (int)(rng() % 17);
      std::snprintf(buf, sizeof buf, (rng() & 1) ? "%.*e" : "%.*g", prec - 1 + ((rng() & 1) ? 0 : 1), d);
      check_double(buf, fmt, "binary64");
    }

    // random digit strings: 1..25 digits, a decimal point anywhere, an exponent
    {
      std::string s;
      if (rng() % 4 == 0) s += (rng() & 1) ? '-' : '+';
      const int nd = 1 + (int)(rng() % 25);
      const int dot = (int)(rng() % (nd + 2)) - 1;
      for (int k = 0; k < nd; ++k) {
        if (k == dot) s += '.';
        s += (char)('0' + rng() % 10);
      }
      if (rng() & 1) {
        s += (rng() & 1) ? 'e' : 'E';
        if (rng() & 1) s += (rng() & 1) ? '-' : '+';
        s += std::to_string((int)(rng() % 340));
      }
      check_double(s, digits, "digits");
    }

    // near-halfway: the midpoint of two neighbouring doubles cut to 17-19 digits
    {
      const double x = std::pow(10.0, -300.0 + 600.0 * u01(rng));
      const double y = std::nextafter(x, INFINITY);
      const long double mid = ((long double)x + (long double)y) / 2;
      std::snprintf(buf, sizeof buf, "%.*Le", 16 + (int)(rng() % 3), mid);
      check_double(buf, near, "near-half");
    }

    // ints of the .trans id columns
    {
      const long long v = (long long)(rng() % 2000000000ull) - (long long)((rng() & 7) == 0 ? 1000000000 : 0);
      std::snprintf(buf, sizeof buf, (rng() & 3) ? "%lld" : "%012lld", v);
      check_int(buf, ints);
    }
  }
  // fixed cases: zero forms, trailing garbage, forms the stream alone decides
  const char* fixed[] = {"0", "-0", "0.0", ".5", "5.", "1e5", "1E-5", "+1.5e+10", "1e", "1e+", ".", "-", "1.2.3",
                         "1.0D-05", "inf", "nan", "0x1p3", "1e400", "1e-400", "4.9406564584124654e-324",
                         "2.2250738585072011e-308", "2.2250738585072014e-308", "1.7976931348623157e308",
                         "9007199254740993", "9007199254740992.5", "123456789012345678901234567890"};
  Stats fx;
  for (const char* s : fixed) check_double(s, fx, "fixed");

  Stats* all[] = {&exo, &fmt, &digits, &near, &ints, &fx};
  const char* names[] = {"exomol %10.4E", "binary64 %.Ne/%.Ng", "digit strings", "near-halfway", "int ids", "fixed"};
  long long bad = 0;
  for (int k = 0; k < 6; ++k) {
    std::printf("%-20s tokens %10lld  parsed %10lld  handed back %9lld  mismatches %lld\n", names[k], all[k]->n,
                all[k]->ok, all[k]->host, all[k]->bad);
    bad += all[k]->bad;
  }
  std::printf("%s\n", bad ? "FAIL" : "OK");
  return bad ? 1 : 0;
}
