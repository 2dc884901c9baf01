/*
 * vw_host.cpp -- host-only utilities of the C-ABI: seeded synthetic skies and
 * the reference's file formats.  No device code.
 *
 *   sky loader   common.hpp:89-142  (binary float32 exact size, or text with
 *                an optional first line equal to npix; auto order bin, txt)
 *   map writer   main_recon_online.cu:516-538 and
 *                mirror_tile_vissfast/main_recon_online_mirror_tile_vissfast.cu:606-640
 *                (float32 binary or one value per line, 6 significant digits)
 *   Viss writer  main_recon_online.cu:358-362 ("re im" per full row)
 *   3-col writer common.hpp:145-153 (scientific, 15 digits)
 */
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/vissws.h"
#include "vw_shared.h"

#define VW_API extern "C" __attribute__((visibility("default")))

namespace {

/* uniform in (0, 1] from the top 53 bits */
double u01_open0(std::mt19937_64& g) { return ((double)(g() >> 11) + 1.0) * (1.0 / 9007199254740992.0); }
double u01(std::mt19937_64& g) { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }

bool load_bin_exact(const std::string& path, int64_t n, double* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::vector<float> buf(1 << 20);
  int64_t got = 0;
  while (got < n) {
    size_t want = (size_t)std::min<int64_t>((int64_t)buf.size(), n - got);
    size_t r = std::fread(buf.data(), sizeof(float), want, f);
    for (size_t i = 0; i < r; ++i) out[got + (int64_t)i] = (double)buf[i];
    got += (int64_t)r;
    if (r < want) break;
  }
  std::fclose(f);
  return got == n;
}

bool load_txt_auto(const std::string& path, int64_t n, double* out) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return false;
  char line[256];
  if (!std::fgets(line, sizeof line, f)) {
    std::fclose(f);
    return false;
  }
  int64_t i = 0;
  float first = std::strtof(line, nullptr);
  long long maybe = std::llround((double)first);
  if (maybe != (long long)n) out[i++] = std::strtod(line, nullptr);
  while (i < n && std::fgets(line, sizeof line, f)) out[i++] = std::strtod(line, nullptr);
  std::fclose(f);
  return i == n;
}

}  // namespace

VW_API int vw_make_sky(uint64_t seed, int64_t npix, int32_t kind, double* out) {
  if (!out || npix <= 0 || kind < 0 || kind > 2) return VW_ERR_ARG;
  std::mt19937_64 g(seed);
  if (kind == 2) {
    /* smooth band sky: (1 + 9 exp(-|lat| / 8 deg)) times a log-normal
     * large-scale field exp(0.5 g - 0.125), g ~ N(0, 1) drawn per Nside-16
     * pixel (the NESTED parent of each pixel); lat = asin(z) of the pixel
     * centre.  Needs npix = 12 Nside^2 with Nside a power of two >= 16. */
    int64_t nside = 1;
    while (12 * nside * nside < npix) nside *= 2;
    if (12 * nside * nside != npix || nside < 16) return VW_ERR_ARG;
    const int64_t ncoarse = 12 * 16 * 16, per = (nside / 16) * (nside / 16);
    std::vector<double> field((size_t)ncoarse);
    for (int64_t q = 0; q < ncoarse; q += 2) {
      double r = std::sqrt(-2.0 * std::log(u01_open0(g)));
      double th = 2.0 * M_PI * u01(g);
      field[(size_t)q] = std::exp(0.5 * r * std::cos(th) - 0.125);
      field[(size_t)q + 1] = std::exp(0.5 * r * std::sin(th) - 0.125);
    }
    const double scale = 180.0 / (M_PI * 8.0);
    for (int64_t p = 0; p < npix; ++p) {
      double l, m, z;
      vw_pix2lmn_nest(nside, p, &l, &m, &z);
      out[p] = (1.0 + 9.0 * std::exp(-std::fabs(std::asin(z)) * scale)) * field[(size_t)(p / per)];
    }
    return VW_OK;
  }
  /* log-normal(0,1) diffuse field, Box-Muller pairs */
  for (int64_t p = 0; p < npix; p += 2) {
    double r = std::sqrt(-2.0 * std::log(u01_open0(g)));
    double th = 2.0 * M_PI * u01(g);
    out[p] = std::exp(r * std::cos(th));
    if (p + 1 < npix) out[p + 1] = std::exp(r * std::sin(th));
  }
  if (kind == 1) {
    /* Poisson(100) point sources (Knuth), capped at 200; Pareto(1.5) flux
     * x_m = 1 at uniform random pixels */
    const double L = std::exp(-100.0);
    int k = 0;
    double prod = u01_open0(g);
    while (prod > L) {
      ++k;
      prod *= u01_open0(g);
    }
    if (k > 200) k = 200;
    for (int s = 0; s < k; ++s) {
      int64_t pix = (int64_t)(u01(g) * (double)npix);
      if (pix >= npix) pix = npix - 1;
      double flux = std::pow(u01_open0(g), -1.0 / 1.5);
      out[pix] += flux;
    }
  }
  return VW_OK;
}

VW_API int vw_io_load_sky(const char* sky_dir, const char* btag, const char* mode, int64_t npix, double* out,
                          char* used_path, int32_t used_cap) {
  if (!sky_dir || !btag || !mode || !out || npix <= 0) return VW_ERR_ARG;
  std::string d(sky_dir);
  while (!d.empty() && d.back() == '/') d.pop_back();
  const std::string b1 = d + "/B_" + btag + ".bin", b2 = d + "/B.bin";
  const std::string t1 = d + "/B_" + btag + ".txt", t2 = d + "/B.txt";
  std::string m(mode);
  std::vector<std::pair<std::string, bool>> order; /* path, is_bin */
  if (m == "bin") order = {{b1, true}, {b2, true}};
  else if (m == "txt") order = {{t1, false}, {t2, false}};
  else if (m == "auto") order = {{b1, true}, {b2, true}, {t1, false}, {t2, false}};
  else return VW_ERR_ARG;
  for (auto& o : order) {
    bool ok = o.second ? load_bin_exact(o.first, npix, out) : load_txt_auto(o.first, npix, out);
    if (ok) {
      if (used_path && used_cap > 0) std::snprintf(used_path, (size_t)used_cap, "%s", o.first.c_str());
      return VW_OK;
    }
  }
  return VW_ERR_IO;
}

VW_API int vw_io_write_map(const char* path, const double* C, int64_t n, int32_t as_text) {
  if (!path || (!C && n > 0)) return VW_ERR_ARG;
  FILE* f = std::fopen(path, as_text ? "w" : "wb");
  if (!f) return VW_ERR_IO;
  bool ok = true;
  if (as_text) {
    for (int64_t i = 0; i < n && ok; ++i) ok = std::fprintf(f, "%g\n", (double)(float)C[i]) > 0;
  } else {
    std::vector<float> buf(1 << 20);
    for (int64_t off = 0; off < n && ok; off += (int64_t)buf.size()) {
      size_t cur = (size_t)std::min<int64_t>((int64_t)buf.size(), n - off);
      for (size_t i = 0; i < cur; ++i) buf[i] = (float)C[off + (int64_t)i];
      ok = std::fwrite(buf.data(), sizeof(float), cur, f) == cur;
    }
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? VW_OK : VW_ERR_IO;
}

VW_API int vw_io_write_viss(const char* path, const vw_c128* V, int64_t n, int32_t append) {
  if (!path || (!V && n > 0)) return VW_ERR_ARG;
  FILE* f = std::fopen(path, append ? "a" : "w");
  if (!f) return VW_ERR_IO;
  bool ok = true;
  for (int64_t i = 0; i < n && ok; ++i)
    ok = std::fprintf(f, "%g %g\n", (double)(float)V[i].re, (double)(float)V[i].im) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? VW_OK : VW_ERR_IO;
}

VW_API int vw_io_write_cols3(const char* path, const double* a, const double* b, const double* c, int64_t n,
                             int32_t append) {
  if (!path || ((!a || !b || !c) && n > 0)) return VW_ERR_ARG;
  FILE* f = std::fopen(path, append ? "a" : "w");
  if (!f) return VW_ERR_IO;
  bool ok = true;
  for (int64_t i = 0; i < n && ok; ++i) ok = std::fprintf(f, "%.15e %.15e %.15e\n", a[i], b[i], c[i]) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? VW_OK : VW_ERR_IO;
}
