// gx_output.cpp -- output writers (C ABI in include/gx_output.h): the reference's per-bin
// text lines and the opacity-table file of the batched (T,P) tables.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gx_output.h"
#include "gx_common.h"

using namespace gx;

namespace {

constexpr char kMagic[8] = {'G', 'X', 'O', 'P', 'A', 'C', '0', '1'};
constexpr uint32_t kHeaderBytes = 128;
constexpr uint32_t kVersion = 1;

// The file's header, field by field at the offsets include/gx_output.h lists.
void put_header(unsigned char* h, const gx_opacity_info& in, uint64_t data_offset) {
  std::memset(h, 0, kHeaderBytes);
  auto put = [&](size_t off, const void* v, size_t n) { std::memcpy(h + off, v, n); };
  put(0, kMagic, 8);
  put(8, &kHeaderBytes, 4);
  put(12, &kVersion, 4);
  put(16, &in.npoints, 8);
  put(24, &in.npts, 8);
  put(32, &in.start_nu, 8);
  put(40, &in.end_nu, 8);
  put(48, &in.dfreq, 8);
  put(56, &in.profile, 4);
  put(60, &in.precision, 4);
  put(64, &in.mean_mass, 8);
  put(72, &in.ref_temp, 8);
  put(80, &in.cutoff, 8);
  put(88, &data_offset, 8);
}

int get_header(const unsigned char* h, gx_opacity_info* out) {
  if (std::memcmp(h, kMagic, 8) != 0) return fail(GX_INVALID_ARG, "opacity file: not a GXOPAC01 file");
  uint32_t hb = 0, ver = 0;
  std::memcpy(&hb, h + 8, 4);
  std::memcpy(&ver, h + 12, 4);
  if (hb != kHeaderBytes || ver != kVersion) return fail(GX_INVALID_ARG, "opacity file: unknown header or version");
  auto get = [&](size_t off, void* v, size_t n) { std::memcpy(v, h + off, n); };
  get(16, &out->npoints, 8);
  get(24, &out->npts, 8);
  get(32, &out->start_nu, 8);
  get(40, &out->end_nu, 8);
  get(48, &out->dfreq, 8);
  get(56, &out->profile, 4);
  get(60, &out->precision, 4);
  get(64, &out->mean_mass, 8);
  get(72, &out->ref_temp, 8);
  get(80, &out->cutoff, 8);
  get(88, &out->data_offset, 8);
  return GX_OK;
}

uint64_t data_offset_for(uint64_t npts) {
  const uint64_t end = kHeaderBytes + 24 * npts;
  return (end + 4095) / 4096 * 4096;
}

}  // namespace

extern "C" {

int gx_xsec_text(const double* freq, const double* intens, int n, char* out, size_t cap, size_t* written) {
  if (n < 0 || !written || (n > 0 && (!freq || !intens))) return fail(GX_INVALID_ARG, "gx_xsec_text: bad arguments");
  const unsigned nt = (unsigned)std::min<long>(std::max(1u, std::thread::hardware_concurrency()),
                                               std::max<long>(1, n / 65536));
  std::vector<std::string> part(nt);
  auto work = [&](unsigned t) {
    const int a = (int)((long long)n * t / nt), z = (int)((long long)n * (t + 1) / nt);
    std::string& o = part[t];
    o.reserve((size_t)(z - a) * 28);
    char buf[128];
    for (int i = a; i < z; ++i) {
      const int m = std::snprintf(buf, sizeof buf, "%12.6f %13.8E\n", freq[i], intens[i]);   // BaseProfile.cpp:70
      o.append(buf, (size_t)m);
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& x : th) x.join();
  size_t total = 0;
  for (auto& p : part) total += p.size();
  *written = total;
  if (!out) return GX_OK;
  if (total > cap) return fail(GX_INVALID_ARG, "gx_xsec_text: output buffer too small");
  size_t o = 0;
  for (auto& p : part) {
    std::memcpy(out + o, p.data(), p.size());
    o += p.size();
  }
  return GX_OK;
}

int gx_xsec_text_write(const char* path, const double* freq, const double* intens, int n) {
  size_t need = 0;
  int rc;
  if ((rc = gx_xsec_text(freq, intens, n, nullptr, 0, &need))) return rc;
  std::vector<char> buf(need + 1);
  if ((rc = gx_xsec_text(freq, intens, n, buf.data(), buf.size(), &need))) return rc;
  FILE* f = path ? std::fopen(path, "wb") : stdout;
  if (!f) return fail(GX_INVALID_ARG, std::string("gx_xsec_text_write: cannot open ") + path);
  const size_t w = std::fwrite(buf.data(), 1, need, f);
  const bool ok = w == need && std::fflush(f) == 0;
  if (path) std::fclose(f);
  return ok ? GX_OK : fail(GX_DEVICE_ERROR, "gx_xsec_text_write: write failed");
}

int gx_opacity_write(const char* path, const gx_opacity_info* info, const gx_point* pts, const double* grids) {
  if (!path || !info || (info->npts && (!pts || (info->npoints && !grids))))
    return fail(GX_INVALID_ARG, "gx_opacity_write: null pointer");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(GX_INVALID_ARG, std::string("gx_opacity_write: cannot open ") + path);
  const uint64_t doff = data_offset_for(info->npts);
  unsigned char h[kHeaderBytes];
  put_header(h, *info, doff);
  bool ok = std::fwrite(h, 1, kHeaderBytes, f) == kHeaderBytes;
  for (uint64_t k = 0; ok && k < info->npts; ++k) {
    const double v[3] = {pts[k].temperature, pts[k].pressure, pts[k].partition};
    ok = std::fwrite(v, 8, 3, f) == 3;
  }
  const uint64_t pad = doff - (kHeaderBytes + 24 * info->npts);
  std::vector<unsigned char> zeros(pad, 0);
  if (ok && pad) ok = std::fwrite(zeros.data(), 1, pad, f) == pad;
  const uint64_t nv = info->npts * info->npoints;
  if (ok && nv) ok = std::fwrite(grids, 8, nv, f) == nv;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? GX_OK : fail(GX_DEVICE_ERROR, "gx_opacity_write: write failed");
}

int gx_opacity_read_info(const char* path, gx_opacity_info* info) {
  if (!path || !info) return fail(GX_INVALID_ARG, "gx_opacity_read_info: null pointer");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GX_INVALID_ARG, std::string("gx_opacity_read_info: cannot open ") + path);
  unsigned char h[kHeaderBytes];
  const bool ok = std::fread(h, 1, kHeaderBytes, f) == kHeaderBytes;
  std::fclose(f);
  if (!ok) return fail(GX_INVALID_ARG, "gx_opacity_read_info: short file");
  return get_header(h, info);
}

int gx_opacity_read(const char* path, gx_point* pts, size_t npts, double* grids, size_t nvalues) {
  gx_opacity_info in;
  int rc;
  if ((rc = gx_opacity_read_info(path, &in))) return rc;
  if (npts != in.npts || nvalues != in.npts * in.npoints || (npts && !pts) || (nvalues && !grids))
    return fail(GX_INVALID_ARG, "gx_opacity_read: sizes do not match the file");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(GX_INVALID_ARG, std::string("gx_opacity_read: cannot open ") + path);
  bool ok = std::fseek(f, kHeaderBytes, SEEK_SET) == 0;
  for (size_t k = 0; ok && k < npts; ++k) {
    double v[3];
    ok = std::fread(v, 8, 3, f) == 3;
    pts[k].temperature = v[0];
    pts[k].pressure = v[1];
    pts[k].partition = v[2];
  }
  if (ok) ok = std::fseek(f, (long)in.data_offset, SEEK_SET) == 0;
  if (ok && nvalues) ok = std::fread(grids, 8, nvalues, f) == nvalues;
  std::fclose(f);
  return ok ? GX_OK : fail(GX_INVALID_ARG, "gx_opacity_read: short file");
}

}  // extern "C"
