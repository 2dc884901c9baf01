/*
 * vissws_combine -- offline accumulation of per-day maps (combine.py:42-126).
 *
 *   vissws_combine --dir=DIR --pattern='C*day10M.bin' [--previous_sum=FILE]
 *                  [--output=FILE]
 *
 * Finds the files in DIR matching the shell pattern, orders them by the day
 * number n of "C<n>day..." (names without one go last, in name order), and
 * adds them element by element in float32 in that order, starting from zero
 * or from the float32 map in --previous_sum.  Writes the float32 sum to
 * --output (default DIR/accumulated_sum.bin) and a text summary next to it
 * (<output without .bin>_metadata.txt: file count, file list, shape, min,
 * max, mean).  Exit status 1 when nothing matches or the sizes differ.
 */
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <glob.h>
#include <string>
#include <vector>

namespace {

std::string get_arg(int argc, char** argv, const std::string& key, const std::string& defv) {
  for (int i = 1; i < argc; ++i) {
    std::string s(argv[i]);
    if (s.rfind(key + "=", 0) == 0) return s.substr(key.size() + 1);
  }
  return defv;
}

std::string base_name(const std::string& p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? p : p.substr(k + 1);
}

/* day number of "C<n>day...", or -1 */
long long day_number(const std::string& path) {
  std::string b = base_name(path);
  if (b.size() < 5 || b[0] != 'C') return -1;
  size_t i = 1;
  while (i < b.size() && std::isdigit((unsigned char)b[i])) ++i;
  if (i == 1 || b.compare(i, 3, "day") != 0) return -1;
  return std::stoll(b.substr(1, i - 1));
}

bool read_f32(const std::string& path, std::vector<float>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0 || n % 4) {
    std::fclose(f);
    return false;
  }
  out.resize((size_t)n / 4);
  bool ok = std::fread(out.data(), 4, out.size(), f) == out.size();
  std::fclose(f);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  std::string dir = get_arg(argc, argv, "--dir", ".");
  std::string pattern = get_arg(argc, argv, "--pattern", "C*day30M.bin");
  std::string prev = get_arg(argc, argv, "--previous_sum", "");
  if (!dir.empty() && dir.back() != '/') dir.push_back('/');
  std::string output = get_arg(argc, argv, "--output", dir + "accumulated_sum.bin");

  glob_t g;
  std::vector<std::string> files;
  if (glob((dir + pattern).c_str(), 0, nullptr, &g) == 0) {
    for (size_t i = 0; i < g.gl_pathc; ++i) files.emplace_back(g.gl_pathv[i]);
    globfree(&g);
  }
  if (files.empty()) {
    std::fprintf(stderr, "no file matches %s in %s\n", pattern.c_str(), dir.c_str());
    return 1;
  }
  std::stable_sort(files.begin(), files.end(), [](const std::string& a, const std::string& b) {
    long long na = day_number(a), nb = day_number(b);
    if ((na < 0) != (nb < 0)) return nb < 0;
    return na < nb;
  });
  std::printf("%zu files:\n", files.size());
  for (const auto& f : files) std::printf("  - %s\n", base_name(f).c_str());

  std::vector<float> acc, data;
  if (!prev.empty()) {
    if (!read_f32(prev, acc)) {
      std::fprintf(stderr, "cannot read previous sum %s\n", prev.c_str());
      return 1;
    }
    std::printf("previous sum %s: %zu values\n", prev.c_str(), acc.size());
  } else {
    if (!read_f32(files[0], data)) {
      std::fprintf(stderr, "cannot read %s\n", files[0].c_str());
      return 1;
    }
    acc.assign(data.size(), 0.0f);
  }
  for (size_t k = 0; k < files.size(); ++k) {
    if (!read_f32(files[k], data)) {
      std::fprintf(stderr, "cannot read %s\n", files[k].c_str());
      return 1;
    }
    if (data.size() != acc.size()) {
      std::fprintf(stderr, "size mismatch: %s has %zu values, expected %zu\n", files[k].c_str(), data.size(),
                   acc.size());
      return 1;
    }
    for (size_t i = 0; i < acc.size(); ++i) acc[i] += data[i];
  }
  FILE* f = std::fopen(output.c_str(), "wb");
  if (!f || std::fwrite(acc.data(), 4, acc.size(), f) != acc.size()) {
    std::fprintf(stderr, "cannot write %s\n", output.c_str());
    if (f) std::fclose(f);
    return 1;
  }
  std::fclose(f);
  double mn = INFINITY, mx = -INFINITY, sum = 0.0;
  for (float v : acc) {
    mn = std::fmin(mn, (double)v);
    mx = std::fmax(mx, (double)v);
    sum += v;
  }
  double mean = acc.empty() ? 0.0 : sum / (double)acc.size();
  std::string meta = output;
  if (meta.size() > 4 && meta.compare(meta.size() - 4, 4, ".bin") == 0) meta.resize(meta.size() - 4);
  meta += "_metadata.txt";
  FILE* m = std::fopen(meta.c_str(), "w");
  if (!m) {
    std::fprintf(stderr, "cannot write %s\n", meta.c_str());
    return 1;
  }
  std::fprintf(m, "Number of files processed: %zu\nFiles:\n", files.size());
  for (const auto& p : files) std::fprintf(m, "  %s\n", p.c_str());
  std::fprintf(m, "\nOutput shape: (%zu,)\nData type: float32\nMin: %.9g\nMax: %.9g\nMean: %.9g\n", acc.size(), mn,
               mx, mean);
  std::fclose(m);
  std::printf("sum of %zu files -> %s (min %g, max %g, mean %g)\n", files.size(), output.c_str(), mn, mx, mean);
  return 0;
}
