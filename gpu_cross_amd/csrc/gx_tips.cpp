// gx_tips.cpp -- TIPS-2011 partition sums read from the caller's copy of the reference's
// Fortran source (HITRAN_files/BD_TIPS_2011_v1p0.for) and interpolated as its AtoB does.
//
// The file is read as fixed-form Fortran: comment lines (c, C, * or ! in column 1) are
// dropped, a non-blank, non-zero column 6 continues the previous statement, statement text
// is columns 7-72, '!' starts a trailing comment and blanks carry no meaning.  From the
// statements the reader takes
//   * the temperature grid:            DATA Tdat/ ... /
//   * the molecule -> routine mapping: IF(MOL.EQ.n) THEN followed by CALL QT_name(...), or by
//                                      gi=0 (molecules with no table, Q = 0)
//   * per routine QT_name:             DIMENSION xgj(k), QofT(k,NT); DATA xgj/ ... /;
//                                      DATA (QofT(i,J),J=1,NT)/ ... /  for i = 1..k
// DATA values may carry a repeat count (r*value).  A default-REAL literal (no D exponent) is
// a single-precision constant, so a value is rounded to float and widened, as the Fortran
// compiler stores it into the DOUBLE PRECISION tables; a D-exponent literal keeps double.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/gx.h"
#include "../../include/gx_tips.h"
#include "gx_common.h"

namespace {

struct Routine {
  int niso = 0;
  std::vector<double> gi;                  // xgj, [niso]
  std::vector<std::vector<double>> q;      // QofT rows, [niso][nt]
};

std::string upper_nospace(const std::string& s) {
  std::string o;
  o.reserve(s.size());
  for (char c : s)
    if (!std::isspace((unsigned char)c)) o.push_back((char)std::toupper((unsigned char)c));
  return o;
}

// Statements of a fixed-form source, blanks removed and upper-cased.
std::vector<std::string> statements(const char* text, size_t n) {
  std::vector<std::string> out;
  std::string cur;
  bool have = false;
  size_t p = 0;
  while (p < n) {
    size_t e = p;
    while (e < n && text[e] != '\n') ++e;
    std::string line(text + p, e - p);
    p = e + 1;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty()) continue;
    const char c0 = line[0];
    if (c0 == 'c' || c0 == 'C' || c0 == '*' || c0 == '!') continue;
    bool cont = false;
    std::string body;
    if (c0 == '\t') {                        // tab format: a digit 1-9 after the tab continues
      if (line.size() > 1 && line[1] >= '1' && line[1] <= '9') {
        cont = true;
        body = line.substr(2);
      } else {
        body = line.substr(1);
      }
    } else {
      if (line.size() > 72) line.resize(72);
      if (line.size() > 5 && line[5] != ' ' && line[5] != '0') cont = true;
      body = line.size() > 6 ? line.substr(6) : std::string();
    }
    const size_t bang = body.find('!');
    if (bang != std::string::npos) body.resize(bang);
    const std::string t = upper_nospace(body);
    if (cont && have) {
      cur += t;
    } else {
      if (have && !cur.empty()) out.push_back(cur);
      cur = t;
      have = true;
    }
  }
  if (have && !cur.empty()) out.push_back(cur);
  return out;
}

// Values of a DATA list "v1,v2,r*v3,...".
bool data_values(const std::string& list, std::vector<double>& out) {
  size_t p = 0;
  while (p < list.size()) {
    size_t e = list.find(',', p);
    if (e == std::string::npos) e = list.size();
    std::string tok = list.substr(p, e - p);
    p = e + 1;
    if (tok.empty()) return false;
    long rep = 1;
    const size_t star = tok.find('*');
    if (star != std::string::npos) {
      rep = std::strtol(tok.substr(0, star).c_str(), nullptr, 10);
      tok = tok.substr(star + 1);
      if (rep <= 0) return false;
    }
    bool dbl = false;
    for (char& c : tok)
      if (c == 'D') {
        c = 'E';
        dbl = true;
      }
    const bool real = dbl || tok.find('.') != std::string::npos || tok.find('E') != std::string::npos;
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (!end || *end != '\0') return false;
    const double stored = (real && !dbl) ? (double)(float)v : v;   // default REAL: single precision
    for (long k = 0; k < rep; ++k) out.push_back(stored);
  }
  return true;
}

// "DATA<name>/<list>/" -> list (name already matched by the caller)
bool slash_list(const std::string& st, size_t from, std::string& list) {
  const size_t a = st.find('/', from);
  const size_t b = st.rfind('/');
  if (a == std::string::npos || b == std::string::npos || b <= a) return false;
  list = st.substr(a + 1, b - a - 1);
  return true;
}

bool starts(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

}  // namespace

struct gx_tips {
  std::vector<double> tdat;                // temperature grid
  std::map<int, Routine> mols;             // molecules with tables
  std::map<int, bool> zero;                // molecules the dispatch answers with Q = 0
};

extern "C" {

int gx_tips_lagrange(const double* A, const double* B, int npt, double aa, double* out) {
  using gx::fail;
  if (!A || !B || !out || npt < 3) return fail(GX_INVALID_ARG, "tips_lagrange: bad arguments");
  *out = 0.0;
  // AtoB (BD_TIPS_2011_v1p0.for:5830-5922), 1-based indices as written there
  auto a = [&](int i) { return A[i - 1]; };
  auto b = [&](int i) { return B[i - 1]; };
  auto nz = [](double d) { return d == 0.0 ? (double)0.0001f : d; };   // "do not divide by zero"
  for (int I = 2; I <= npt; ++I) {
    if (!(a(I) >= aa)) continue;
    if (I < 3 || I == npt) {               // three-point
      int J = I;
      if (I < 3) J = 3;
      if (I == npt) J = npt;
      const double A0D1 = nz(a(J - 2) - a(J - 1)), A0D2 = nz(a(J - 2) - a(J));
      const double A1D1 = nz(a(J - 1) - a(J - 2)), A1D2 = nz(a(J - 1) - a(J));
      const double A2D1 = nz(a(J) - a(J - 2)), A2D2 = nz(a(J) - a(J - 1));
      const double A0 = (aa - a(J - 1)) * (aa - a(J)) / (A0D1 * A0D2);
      const double A1 = (aa - a(J - 2)) * (aa - a(J)) / (A1D1 * A1D2);
      const double A2 = (aa - a(J - 2)) * (aa - a(J - 1)) / (A2D1 * A2D2);
      *out = A0 * b(J - 2) + A1 * b(J - 1) + A2 * b(J);
    } else {                               // four-point
      const int J = I;
      const double A0D1 = nz(a(J - 2) - a(J - 1)), A0D2 = nz(a(J - 2) - a(J)), A0D3 = nz(a(J - 2) - a(J + 1));
      const double A1D1 = nz(a(J - 1) - a(J - 2)), A1D2 = nz(a(J - 1) - a(J)), A1D3 = nz(a(J - 1) - a(J + 1));
      const double A2D1 = nz(a(J) - a(J - 2)), A2D2 = nz(a(J) - a(J - 1)), A2D3 = nz(a(J) - a(J + 1));
      const double A3D1 = nz(a(J + 1) - a(J - 2)), A3D2 = nz(a(J + 1) - a(J - 1)), A3D3 = nz(a(J + 1) - a(J));
      double A0 = (aa - a(J - 1)) * (aa - a(J)) * (aa - a(J + 1));
      A0 = A0 / (A0D1 * A0D2 * A0D3);
      double A1 = (aa - a(J - 2)) * (aa - a(J)) * (aa - a(J + 1));
      A1 = A1 / (A1D1 * A1D2 * A1D3);
      double A2 = (aa - a(J - 2)) * (aa - a(J - 1)) * (aa - a(J + 1));
      A2 = A2 / (A2D1 * A2D2 * A2D3);
      double A3 = (aa - a(J - 2)) * (aa - a(J - 1)) * (aa - a(J));
      A3 = A3 / (A3D1 * A3D2 * A3D3);
      *out = A0 * b(J - 2) + A1 * b(J - 1) + A2 * b(J) + A3 * b(J + 1);
    }
    return GX_OK;
  }
  return fail(GX_INVALID_ARG, "tips_lagrange: x above the last node");
}

int gx_tips_load_text(const char* text, size_t n, gx_tips** out) {
  using gx::fail;
  if (!text || !out) return fail(GX_INVALID_ARG, "tips_load: NULL argument");
  *out = nullptr;
  const std::vector<std::string> st = statements(text, n);
  auto* T = new gx_tips();
  std::map<int, std::string> dispatch;
  std::map<std::string, Routine> routines;
  std::map<std::string, std::vector<std::vector<double>>> rows_by;   // temporarily sparse rows
  std::string sub;
  int pending_mol = -1;
  auto bad = [&](const std::string& m) {
    delete T;
    return fail(GX_INVALID_ARG, "tips_load: " + m);
  };
  for (size_t i = 0; i < st.size(); ++i) {
    const std::string& s = st[i];
    if (pending_mol >= 0) {
      if (starts(s, "CALLQT_")) {
        dispatch[pending_mol] = s.substr(4, s.find('(') - 4);
      } else if (starts(s, "GI=0")) {
        T->zero[pending_mol] = true;
      }
      pending_mol = -1;
    }
    if (starts(s, "IF(MOL.EQ.")) {
      const int mol = std::atoi(s.c_str() + 10);
      if (s.find(")THEN") != std::string::npos && mol > 0) pending_mol = mol;
      continue;
    }
    if (starts(s, "SUBROUTINEQT_")) {
      sub = s.substr(10, s.find('(') - 10);
      routines[sub];
      continue;
    }
    if (starts(s, "DATATDAT/")) {
      std::string list;
      if (!slash_list(s, 0, list) || !data_values(list, T->tdat)) return bad("bad Tdat list");
      continue;
    }
    if (sub.empty()) continue;
    Routine& r = routines[sub];
    if (starts(s, "DIMENSIONXGJ(")) {
      r.niso = std::atoi(s.c_str() + 13);
    } else if (starts(s, "DATAXGJ/")) {
      std::string list;
      r.gi.clear();
      if (!slash_list(s, 0, list) || !data_values(list, r.gi)) return bad("bad xgj list in " + sub);
    } else if (starts(s, "DATA(QOFT(")) {
      const int iso = std::atoi(s.c_str() + 10);
      std::string list;
      std::vector<double> v;
      const size_t close = s.find(")/");
      if (iso <= 0 || close == std::string::npos || !slash_list(s, close, list) || !data_values(list, v))
        return bad("bad QofT list in " + sub);
      auto& rows = rows_by[sub];
      if ((int)rows.size() < iso) rows.resize(iso);
      rows[iso - 1] = std::move(v);
    }
  }
  const size_t nt = T->tdat.size();
  if (nt < 3) return bad("no temperature grid (DATA Tdat)");
  for (size_t k = 1; k < nt; ++k)
    if (!(T->tdat[k] > T->tdat[k - 1])) return bad("temperature grid not increasing");
  for (const auto& d : dispatch) {
    auto it = routines.find(d.second);
    if (it == routines.end()) return bad("molecule " + std::to_string(d.first) + ": no routine " + d.second);
    Routine r = it->second;
    r.q = rows_by[d.second];
    if (r.niso <= 0 || (int)r.gi.size() != r.niso || (int)r.q.size() != r.niso)
      return bad("molecule " + std::to_string(d.first) + " (" + d.second + "): incomplete tables");
    for (const auto& row : r.q)
      if (row.size() != nt) return bad("molecule " + std::to_string(d.first) + ": table row of wrong length");
    T->mols[d.first] = std::move(r);
  }
  if (T->mols.empty()) return bad("no molecule tables found");
  *out = T;
  return GX_OK;
}

int gx_tips_load(const char* path, gx_tips** out) {
  using gx::fail;
  if (!path || !out) return fail(GX_INVALID_ARG, "tips_load: NULL argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GX_NOT_FOUND, std::string("tips_load: cannot open ") + path);
  std::string text;
  char buf[1 << 16];
  size_t k;
  while ((k = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, k);
  std::fclose(f);
  return gx_tips_load_text(text.data(), text.size(), out);
}

void gx_tips_free(gx_tips* t) { delete t; }

int gx_tips_isotopes(const gx_tips* t, int mol) {
  if (!t) return 0;
  auto it = t->mols.find(mol);
  return it == t->mols.end() ? 0 : it->second.niso;
}

int gx_tips_q(const gx_tips* t, int mol, int iso, double temp, double* gi, double* q) {
  using gx::fail;
  if (!t || !gi || !q) return fail(GX_INVALID_ARG, "tips_q: NULL argument");
  if (t->zero.count(mol)) {                  // BD_TIPS_2011_v1p0.for:266-270, 290-294
    *gi = 0.0;
    *q = 0.0;
    return GX_OK;
  }
  auto it = t->mols.find(mol);
  if (it == t->mols.end()) return fail(GX_INVALID_ARG, "tips_q: no table for molecule " + std::to_string(mol));
  const Routine& r = it->second;
  if (iso < 1 || iso > r.niso)
    return fail(GX_INVALID_ARG, "tips_q: molecule " + std::to_string(mol) + " has no isotopologue " +
                                    std::to_string(iso));
  *gi = r.gi[iso - 1];
  if (temp < 70.0 || temp > 3000.0) {        // the routines' range test, Qt = -1
    *q = -1.0;
    return fail(GX_INVALID_ARG, "tips_q: temperature outside 70-3000 K");
  }
  return gx_tips_lagrange(t->tdat.data(), r.q[iso - 1].data(), (int)t->tdat.size(), temp, q);
}

}  // extern "C"
