/*
 * ingest_oracle.cpp -- TEST INFRASTRUCTURE ONLY.  A CPU restatement of the reference's
 * ExoMol reading path, used by tests/ to check the GPU reader of libgx.so; nothing in the
 * product links or calls it.
 *
 * Follows, step by step and with the same library calls:
 *   initialize_states   exomol_functions.cpp:8-48  (fgets 1024, strtol base 0, strtod)
 *   InitializeBroadener ExomolStateReader.cpp:39-87 (split, atoi/atof, key = code + Ji + Jf)
 *   GetGammaN           ExomolStateReader.cpp:7-37  (key "a1"JiJf, then the same stream
 *                       after clear(): "a1"JiJf"a0"Ji"-1"; gamma *= pressure)
 *   ReadNextState       ExomolStateReader.cpp:130-167 (stream >> id_f >> id_i >> aif)
 *   the read loop       VoigtProfile.cpp:56-68 (continue below lo, break above hi, per file)
 * and the HITRAN .par reader:
 *   GetField            HITRANStateReader.cpp:3-40  (fixed widths, 13-byte zeroed copy cut
 *                       at 12, atof / atoi)
 *   ReadNextState       HITRANStateReader.cpp:136-250 (fgets 1024 into one buffer kept for
 *                       the file; gamma mixing, A from the intensity when A == 0)
 *   the read loop       VoigtProfile.cpp:61-75 (the same filter; g stored as int)
 * gamma / n before the first key match are 0 here (uninitialised in the reference,
 * VoigtProfile.cpp:41); a blank .broad line is skipped (the reference indexes an empty
 * split, undefined).
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <istream>
#include <iterator>
#include <map>
#include <sstream>
#include <streambuf>
#include <string>
#include <vector>

namespace {

struct MemBuf : std::streambuf {
  MemBuf(const char* p, size_t n) {
    char* b = const_cast<char*>(p);
    setg(b, b, b + n);
  }
};

struct Broad {
  double gamma, n;
};

}  // namespace

extern "C" {

/* exomol_functions.cpp:8-48 over text in memory (fmemopen stands in for fopen). */
long orc_states(const char* text, long len, long cap, int* id, double* energy, int* gns, int* J) {
  FILE* f = fmemopen(const_cast<char*>(text), (size_t)len, "r");
  if (!f) return -1;
  char buffer[1024];
  long n = 0;
  char* ln_ptr;
  while (fgets(buffer, 1024, f)) {
    if (n < cap) {
      id[n] = strtol(buffer, &ln_ptr, 0) - 1;
      energy[n] = strtod(ln_ptr, &ln_ptr);
      gns[n] = strtol(ln_ptr, &ln_ptr, 0);
      J[n] = strtol(ln_ptr, &ln_ptr, 0);
    }
    n++;
  }
  fclose(f);
  return n;
}

/* The read loop over nfiles .trans texts.  Returns the number of kept lines (at most
 * cap are stored), *records = ReadNextState calls that returned true. */
long orc_ingest(const double* energy, const int* gns, const int* J, long nstates, const char* broad, long broad_len,
                int use_broad, double pressure, double lo, double hi, const char* const* files, const long* lens,
                int nfiles, long cap, double* elow, double* nu, double* aif, int* g_out, double* gamma_out,
                double* n_out, long* records) {
  std::map<std::string, Broad> broad_map;
  if (use_broad) {
    std::istringstream ib(std::string(broad, (size_t)broad_len));
    std::string line;
    while (std::getline(ib, line)) {
      std::istringstream buffer(line);
      std::vector<std::string> split_line((std::istream_iterator<std::string>(buffer)),
                                          std::istream_iterator<std::string>());
      if (split_line.empty()) continue;
      std::string code = "a0";
      int Ji = atoi(split_line[3].c_str());
      int Jf = -1;
      double gam = atof(split_line[1].c_str());
      double n = atof(split_line[2].c_str());
      if (split_line[0] == "a1") {
        code = "a1";
        Jf = atoi(split_line[4].c_str());
      }
      std::stringstream key;
      key << code << Ji << Jf;
      broad_map[key.str()].gamma = gam;
      broad_map[key.str()].n = n;
    }
  }
  double t_gammaL = 0.0, t_n = 0.0;
  long kept = 0, nrec = 0;
  for (int fi = 0; fi < nfiles; ++fi) {
    MemBuf mb(files[fi], (size_t)lens[fi]);
    std::istream stream(&mb);
    int id_f, id_i;
    double t_aif;
    while (true) {
      if (!(stream >> id_f >> id_i >> t_aif)) break;
      id_f -= 1;
      id_i -= 1;
      if (id_f < 0 || id_f >= nstates || id_i < 0 || id_i >= nstates) return -2;
      const double t_nu = energy[id_f] - energy[id_i];
      const double t_ei = energy[id_i];
      const int t_gns = gns[id_f];
      if (use_broad && !broad_map.empty()) {
        const int Ji = J[id_i], Jf = J[id_f];
        std::stringstream key;
        key << "a1" << Ji << Jf;
        if (broad_map.count(key.str()) > 0) {
          t_gammaL = broad_map[key.str()].gamma * pressure;
          t_n = broad_map[key.str()].n;
        } else {
          key.clear();
          key << "a0" << Ji << -1;
          if (broad_map.count(key.str()) > 0) {
            t_gammaL = broad_map[key.str()].gamma * pressure;
            t_n = broad_map[key.str()].n;
          }
        }
      }
      ++nrec;
      if (t_nu < lo) continue;
      if (t_nu > hi) break;
      if (kept < cap) {
        elow[kept] = t_ei;
        nu[kept] = t_nu;
        aif[kept] = t_aif;
        g_out[kept] = t_gns;
        gamma_out[kept] = t_gammaL;
        n_out[kept] = t_n;
      }
      ++kept;
    }
  }
  *records = nrec;
  return kept;
}

/* HITRANStateReader over one .par text in memory: the constructor's first record gives the
 * molecule and isotope, then ReadNextState record by record with the driver's filter.  The
 * reader's buffer starts zeroed (uninitialised in the reference). */
long orc_hitran(const char* text, long len, double pressure, double mixture, double ref_partition, double lo,
                double hi, long cap, double* elow, double* nu, double* aif, int* gns, double* gamma, double* nexp,
                long* records, int* mol_id, int* iso) {
  static const int width[19] = {2, 1, 12, 10, 10, 5, 5, 10, 4, 8, 15, 15, 15, 15, 6, 12, 1, 7, 7};
  char buffer[1024];
  std::memset(buffer, 0, sizeof buffer);
  auto copy_field = [&](int f, char* c) {
    int at = 0;
    for (int i = 0; i < f; ++i) at += width[i];
    std::memset(c, 0, 13);
    std::memcpy(c, buffer + at, (size_t)width[f]);
    c[12] = '\0';
  };
  auto as_double = [&](int f) {
    char c[13];
    copy_field(f, c);
    return std::atof(c);
  };
  auto as_int = [&](int f) {
    char c[13];
    copy_field(f, c);
    return std::atoi(c);
  };
  *records = 0;
  *mol_id = *iso = 0;
  if (len <= 0) return 0;
  FILE* fp = fmemopen(const_cast<char*>(text), (size_t)len, "r");
  if (!fp) return -1;
  if (std::fgets(buffer, 1024, fp)) {
    *mol_id = as_int(0);
    *iso = as_int(1);
  }
  std::rewind(fp);
  const double planck = 6.6260693e-27, vellgt = 2.99792458e10, boltz = 1.380658e-16;
  const double pi = std::acos(-1.0);
  long kept = 0, nrec = 0;
  while (std::fgets(buffer, 1024, fp)) {
    ++nrec;
    double v = as_double(2), e = as_double(7), a = as_double(4);
    const double gair = as_double(5), gself = as_double(6), n = as_double(8);
    double g = as_double(17);
    const double s = as_double(3);
    const double gam = (gair * mixture + gself * (1.0 - mixture)) * pressure;
    if (a == 0.0) {
      const double beta = planck * vellgt / (boltz * 296.0);
      a = s * 8.0 * pi * vellgt * ref_partition * v * v / (std::exp(-beta * e) * (1.0 - std::exp(-beta * v)));
      g = 1.0;
    }
    if (v < lo) continue;
    if (v > hi) break;
    if (kept < cap) {
      elow[kept] = e;
      nu[kept] = v;
      aif[kept] = a;
      gns[kept] = (int)g;
      gamma[kept] = gam;
      nexp[kept] = n;
    }
    ++kept;
  }
  std::fclose(fp);
  *records = nrec;
  return kept;
}

}  // extern "C"
