/*
 * vissws_recon_online -- drop-in driver for main_recon_online
 * (main_recon_online.cu:5-586) and its multi-day mirror variant
 * (mirror_tile_vissfast/main_recon_online_mirror_tile_vissfast.cu:6-700),
 * running the gfx950 library through its C-ABI on one GPU.
 *
 * Reference flags (same names, defaults and fall-back-on-parse-error
 * behaviour, common.hpp:58-66):
 *   --btag={1M,10M,30M} --nside --day --day_start --day_count --segs
 *   --blockage --occ_mode --write_viss --write_baseline_txt --B_mode
 *   --C_mode={txt,bin} --gen_gpu_index --orbit_seed --sky_dir --out_dir
 *   --gpus --print_stats --viss_tile_pix={256,512}
 * Additional flags:
 *   --geometry={lunar8,circular} --n_ant --n_epochs --dt_s --freq_hz
 *   --precision={plain,mixed} --recon={full,mirror} --img_n --bl_max_m
 *   --sky={file,lognormal,lognormal_ps} --sky_seed --occ_radius_km
 *   --occ_altitude_km
 *   --viss_form={auto,rows,ant}  Viss summation form (vw_config.viss_form)
 *   --checkpoint=PATH  segment-level checkpoint (raw accumulators, position
 *                      in the schedule, text-output sizes) after every
 *                      segment and day; --resume=1 continues from it;
 *                      --stop_after_segs=N ends this invocation after N
 *                      segments (the checkpoint is kept)
 * Outputs (out_dir): C{day}day{btag}.{txt|bin} (npix float32, NESTED),
 * Viss{day}day{btag}.txt, uvw/xyza/xyzb{day}day{btag}.txt, and with
 * --img_n>0 img{day}day{btag}.bin (img_n^2 float64, row = m, col = l).
 */
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>

#include "../../include/vissws.h"

namespace {

std::string get_arg(int argc, char** argv, const std::string& key, const std::string& defv) {
  for (int i = 1; i < argc; ++i) {
    std::string s(argv[i]);
    if (s.rfind(key + "=", 0) == 0) return s.substr(key.size() + 1);
  }
  return defv;
}
int to_int(const std::string& s, int d) {
  try { return std::stoi(s); } catch (...) { return d; }
}
double to_dbl(const std::string& s, double d) {
  try { return std::stod(s); } catch (...) { return d; }
}
uint64_t to_u64(const std::string& s, uint64_t d) {
  try { return (uint64_t)std::stoull(s); } catch (...) { return d; }
}

struct Timer {
  std::chrono::high_resolution_clock::time_point t0;
  void tic() { t0 = std::chrono::high_resolution_clock::now(); }
  double toc() const {
    return std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
  }
};

int mkdirs(const std::string& path) {
  std::string cur;
  for (size_t i = 0; i < path.size(); ++i) {
    cur.push_back(path[i]);
    if (path[i] == '/' || i + 1 == path.size()) mkdir(cur.c_str(), 0755);
  }
  return 0;
}

int die(void* ctx, const char* what) {
  std::fprintf(stderr, "ERROR %s: %s\n", what, vw_last_error(ctx));
  return 1;
}

/* ---- segment-level checkpoint ----
 * Layout: Ckpt header, then (when next_seg > 0) C[npix] f64, W[npix] u32,
 * G[ngp] f64, GW[ngp] u32.  Written to PATH.tmp and renamed over PATH, so an
 * interrupted write leaves the previous checkpoint. */
struct Ckpt {
  char magic[8];
  uint64_t fingerprint; /* run parameters that must match on resume */
  int32_t day, next_seg; /* next work item: segment next_seg of day      */
  int32_t day_start, day_count; /* day range of the run that wrote it    */
  int64_t npix, ngp, wtotal;
  int64_t fsize[4]; /* Viss, uvw, xyza, xyzb text sizes (-1: absent)     */
};
const char kMagic[8] = {'V', 'W', 'C', 'K', 'P', 'T', '0', '2'};

struct Fnv {
  uint64_t h = 1469598103934665603ULL;
  void add(const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ULL;
  }
  template <class T> void v(T x) { add(&x, sizeof x); }
  void str(const std::string& s) { add(s.data(), s.size()); v((uint64_t)s.size()); }
};

int64_t file_size(const std::string& f) {
  struct stat st;
  return stat(f.c_str(), &st) == 0 ? (int64_t)st.st_size : -1;
}

bool write_ckpt(const std::string& path, const Ckpt& h, const std::vector<double>& C, const std::vector<uint32_t>& W,
                const std::vector<double>& G, const std::vector<uint32_t>& GW) {
  std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  if (ok && h.next_seg > 0)
    ok = std::fwrite(C.data(), sizeof(double), (size_t)h.npix, f) == (size_t)h.npix &&
         std::fwrite(W.data(), sizeof(uint32_t), (size_t)h.npix, f) == (size_t)h.npix &&
         std::fwrite(G.data(), sizeof(double), (size_t)h.ngp, f) == (size_t)h.ngp &&
         std::fwrite(GW.data(), sizeof(uint32_t), (size_t)h.ngp, f) == (size_t)h.ngp;
  ok = (std::fclose(f) == 0) && ok;
  return ok && std::rename(tmp.c_str(), path.c_str()) == 0;
}

/* 1 = loaded, 0 = absent, -1 = unreadable */
int read_ckpt(const std::string& path, Ckpt& h, std::vector<double>& C, std::vector<uint32_t>& W,
              std::vector<double>& G, std::vector<uint32_t>& GW) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return 0;
  bool ok = std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, kMagic, 8) == 0 && h.npix >= 0 &&
            h.ngp >= 0;
  if (ok && h.next_seg > 0) {
    C.resize((size_t)h.npix);
    W.resize((size_t)h.npix);
    G.resize((size_t)h.ngp);
    GW.resize((size_t)h.ngp);
    ok = std::fread(C.data(), sizeof(double), (size_t)h.npix, f) == (size_t)h.npix &&
         std::fread(W.data(), sizeof(uint32_t), (size_t)h.npix, f) == (size_t)h.npix &&
         std::fread(G.data(), sizeof(double), (size_t)h.ngp, f) == (size_t)h.ngp &&
         std::fread(GW.data(), sizeof(uint32_t), (size_t)h.ngp, f) == (size_t)h.ngp;
  }
  std::fclose(f);
  return ok ? 1 : -1;
}

}  // namespace

int main(int argc, char** argv) {
  std::string btag = get_arg(argc, argv, "--btag", "10M");
  std::string geom_s = get_arg(argc, argv, "--geometry", "lunar8");
  int geometry = geom_s == "circular" ? VW_GEOM_CIRCULAR : VW_GEOM_LUNAR8;
  if (geom_s != "circular" && geom_s != "lunar8") {
    std::fprintf(stderr, "ERROR --geometry\n");
    return 1;
  }
  vw_config cfg;
  vw_config_init(&cfg, geometry);
  cfg.nside = to_int(get_arg(argc, argv, "--nside", "4096"), 4096);
  int day = to_int(get_arg(argc, argv, "--day", "1"), 1);
  int day_start = to_int(get_arg(argc, argv, "--day_start", std::to_string(day)), day);
  int day_count = to_int(get_arg(argc, argv, "--day_count", "1"), 1);
  int segs = to_int(get_arg(argc, argv, "--segs", "0"), 0);
  cfg.blockage = to_int(get_arg(argc, argv, "--blockage", "1"), 1);
  cfg.occ_mode = to_int(get_arg(argc, argv, "--occ_mode", "2"), 2);
  int write_viss = to_int(get_arg(argc, argv, "--write_viss", "0"), 1); /* 1 on a parse error, as main_recon_online.cu:13 */
  int write_baseline_txt = to_int(get_arg(argc, argv, "--write_baseline_txt", "0"), 0);
  std::string B_mode = get_arg(argc, argv, "--B_mode", "auto");
  std::string C_mode = get_arg(argc, argv, "--C_mode", "txt");
  cfg.orbit_seed = to_u64(get_arg(argc, argv, "--orbit_seed", "42"), 12345ULL);
  std::string sky_dir = get_arg(argc, argv, "--sky_dir", "");
  std::string out_dir = get_arg(argc, argv, "--out_dir", "./out/");
  std::string gpus_s = get_arg(argc, argv, "--gpus", "0");
  int print_stats = to_int(get_arg(argc, argv, "--print_stats", "1"), 1);
  cfg.tile_pix = to_int(get_arg(argc, argv, "--viss_tile_pix", "256"), 256);
  std::string prec = get_arg(argc, argv, "--precision", "plain");
  std::string recon = get_arg(argc, argv, "--recon", "full");
  std::string vform = get_arg(argc, argv, "--viss_form", "auto");
  cfg.img_n = to_int(get_arg(argc, argv, "--img_n", "0"), 0);
  cfg.bl_max_m = to_dbl(get_arg(argc, argv, "--bl_max_m", "0"), 0.0);
  std::string sky_kind = get_arg(argc, argv, "--sky", "file");
  uint64_t sky_seed = to_u64(get_arg(argc, argv, "--sky_seed", "1"), 1);
  std::string ckpt_path = get_arg(argc, argv, "--checkpoint", "");
  int resume = to_int(get_arg(argc, argv, "--resume", "0"), 0);
  int stop_after = to_int(get_arg(argc, argv, "--stop_after_segs", "0"), 0);
  if (geometry == VW_GEOM_CIRCULAR) {
    cfg.n_ant = to_int(get_arg(argc, argv, "--n_ant", "8"), 8);
    cfg.n_epochs = to_int(get_arg(argc, argv, "--n_epochs", "64"), 64);
    cfg.dt_s = to_dbl(get_arg(argc, argv, "--dt_s", "10"), 10.0);
  }
  cfg.occ_radius_km = to_dbl(get_arg(argc, argv, "--occ_radius_km", std::to_string(cfg.occ_radius_km)), cfg.occ_radius_km);
  cfg.occ_altitude_km =
      to_dbl(get_arg(argc, argv, "--occ_altitude_km", std::to_string(cfg.occ_altitude_km)), cfg.occ_altitude_km);

  if (!gpus_s.empty()) setenv("HIP_VISIBLE_DEVICES", gpus_s.c_str(), 1);
  if (gpus_s.find(',') != std::string::npos)
    std::fprintf(stderr, "WARNING --gpus=%s lists several devices; this build runs on the first visible one\n",
                 gpus_s.c_str());

  if (btag != "1M" && btag != "10M" && btag != "30M") {
    std::fprintf(stderr, "ERROR --btag\n");
    return 1;
  }
  if (cfg.nside <= 0) cfg.nside = btag == "1M" ? 512 : (btag == "10M" ? 4096 : 16384);
  if (segs <= 0) segs = btag == "1M" ? 1 : (btag == "10M" ? 10 : 30);
  cfg.segs = segs;
  cfg.freq_hz = btag == "1M" ? 1e6 : (btag == "10M" ? 1e7 : 3e7);
  cfg.freq_hz = to_dbl(get_arg(argc, argv, "--freq_hz", "0"), 0.0) > 0 ? to_dbl(get_arg(argc, argv, "--freq_hz", "0"), 0.0)
                                                                       : cfg.freq_hz;
  if (prec != "plain" && prec != "mixed") {
    std::fprintf(stderr, "ERROR --precision\n");
    return 1;
  }
  cfg.precision = prec == "mixed" ? VW_PREC_MIXED : VW_PREC_PLAIN;
  if (recon != "full" && recon != "mirror") {
    std::fprintf(stderr, "ERROR --recon\n");
    return 1;
  }
  cfg.recon_mode = recon == "mirror" ? VW_RECON_MIRROR : VW_RECON_FULL;
  if (vform != "auto" && vform != "rows" && vform != "ant") {
    std::fprintf(stderr, "ERROR --viss_form\n");
    return 1;
  }
  cfg.viss_form = vform == "rows" ? VW_VISS_ROWS : (vform == "ant" ? VW_VISS_ANT : VW_VISS_AUTO);
  if (C_mode != "txt" && C_mode != "bin") {
    std::fprintf(stderr, "ERROR --C_mode\n");
    return 1;
  }
  if (cfg.tile_pix != 256 && cfg.tile_pix != 512) {
    std::fprintf(stderr, "ERROR --viss_tile_pix supports 256 or 512\n");
    return 1;
  }
  if (sky_dir.empty()) sky_dir = btag == "1M" ? "./earth_1Mhz" : (btag == "10M" ? "./earth_10Mhz" : "./earth_30Mhz");
  if (!out_dir.empty() && out_dir.back() != '/') out_dir.push_back('/');
  mkdirs(out_dir);
  cfg.day = day_start;

  int64_t npix = 12LL * cfg.nside * cfg.nside;
  std::printf("HIP_VISIBLE_DEVICES=%s\n", gpus_s.c_str());
  std::printf("btag=%s nside=%d npix=%lld day_start=%d day_count=%d segs=%d\n", btag.c_str(), cfg.nside,
              (long long)npix, day_start, day_count, segs);
  std::printf("blockage=%d occ_mode=%d precision=%s recon=%s  [Viss stage: blockage=ON, phase_correct=ON]\n",
              cfg.blockage, cfg.occ_mode, prec.c_str(), recon.c_str());
  std::printf(" sky_dir=%s out_dir=%s gpus=%s\n", sky_dir.c_str(), out_dir.c_str(), gpus_s.c_str());

  Timer t_io;
  t_io.tic();
  std::vector<double> B((size_t)npix);
  if (sky_kind == "file") {
    char used[1024] = {0};
    if (vw_io_load_sky(sky_dir.c_str(), btag.c_str(), B_mode.c_str(), npix, B.data(), used, sizeof used) != VW_OK) {
      std::fprintf(stderr, "ERROR reading B in mode=%s from %s\n", B_mode.c_str(), sky_dir.c_str());
      return 1;
    }
    std::printf("Loaded B from %s in %g s\n", used, t_io.toc());
  } else if (sky_kind == "lognormal" || sky_kind == "lognormal_ps") {
    vw_make_sky(sky_seed, npix, sky_kind == "lognormal" ? 0 : 1, B.data());
    std::printf("Synthetic sky %s seed=%llu\n", sky_kind.c_str(), (unsigned long long)sky_seed);
  } else {
    std::fprintf(stderr, "ERROR --sky\n");
    return 1;
  }

  void* ctx = nullptr;
  Timer t_init;
  t_init.tic();
  if (vw_create(&cfg, &ctx) != VW_OK) return die(nullptr, "vw_create");
  vw_plan plan;
  vw_get_plan(ctx, &plan);
  std::printf(" pix_area=%.17g inv_pix_area=%.17g\n", plan.pix_area, 1.0 / plan.pix_area);
  std::printf("cosphi=%.17g lamda=%g bl_max=%g\n", plan.cosphi, plan.lambda_m, plan.bl_max_m);
  if (geometry == VW_GEOM_LUNAR8)
    std::printf("Online orbit plan: OrbitRes=%lld segLen=%lld T=%d N=%lld\n", (long long)plan.orbit_res,
                (long long)plan.seg_len, plan.n_epochs, (long long)plan.n_epochs * 2 * plan.n_pairs);
  std::printf("Segment plan: segs=%d max_tlen=%d max_segN=%lld max_segN_half=%lld\n", plan.segs, plan.max_tlen,
              (long long)plan.max_tlen * 2 * plan.n_pairs, (long long)plan.max_tlen * plan.n_pairs);
  if (vw_set_sky(ctx, B.data(), npix, 1) != VW_OK) return die(ctx, "vw_set_sky");
  std::printf("Init GPU in %g s (device memory of the context %.3f GB)\n", t_init.toc(), plan.device_bytes / 1e9);

  std::vector<int32_t> seg_t0(plan.segs), seg_tlen(plan.segs);
  vw_get_schedule(ctx, seg_t0.data(), seg_tlen.data());

  /* checkpoint identity: everything that changes the accumulated values */
  Fnv fp;
  fp.v(cfg.nside); fp.v(cfg.freq_hz); fp.v((int32_t)geometry); fp.v(cfg.n_ant); fp.v(plan.n_epochs);
  fp.v(cfg.dt_s); fp.v(cfg.orbit_seed); fp.v(plan.segs); fp.v(cfg.blockage); fp.v(cfg.occ_mode);
  fp.v(cfg.recon_mode); fp.v(cfg.precision); fp.v(cfg.tile_pix); fp.v(cfg.img_n); fp.v(plan.bl_max_m);
  fp.v(cfg.viss_form);
  fp.v(cfg.occ_radius_km); fp.v(cfg.occ_altitude_km); fp.v(cfg.r_min_km); fp.v(cfg.r_max_km);
  fp.str(btag); fp.str(sky_kind); fp.v(sky_seed); fp.str(sky_kind == "file" ? sky_dir : std::string());
  fp.str(sky_kind == "file" ? B_mode : std::string());
  /* the loaded sky itself: another file (B_10M.bin vs B.txt) or an edited one
   * in the same directory must not resume into these accumulators */
  fp.add(B.data(), B.size() * sizeof(double));
  fp.v(write_viss); fp.v(write_baseline_txt);
  for (int k = 0; k < plan.segs; ++k) { fp.v(seg_t0[k]); fp.v(seg_tlen[k]); }
  std::vector<double> ckC, ckG;
  std::vector<uint32_t> ckW, ckGW;
  int d_begin = 0, s_begin = 0;
  int64_t ck_wtotal = 0;
  bool restore = false;
  if (resume && !ckpt_path.empty()) {
    Ckpt h;
    int r = read_ckpt(ckpt_path, h, ckC, ckW, ckG, ckGW);
    if (r < 0) {
      std::fprintf(stderr, "ERROR unreadable checkpoint %s\n", ckpt_path.c_str());
      return 1;
    }
    if (r > 0) {
      if (h.fingerprint != fp.h || h.npix != npix || h.ngp != plan.n_img_pts) {
        std::fprintf(stderr, "ERROR checkpoint %s was written by a run with other parameters\n", ckpt_path.c_str());
        return 1;
      }
      if (h.day_start != day_start) {
        std::fprintf(stderr, "ERROR checkpoint %s was written by a run starting at day %d, not %d\n", ckpt_path.c_str(),
                     h.day_start, day_start);
        return 1;
      }
      if (h.day < day_start || h.day > day_start + day_count || h.next_seg < 0 || h.next_seg > plan.segs) {
        std::fprintf(stderr, "ERROR checkpoint %s is outside days [%d, %d]\n", ckpt_path.c_str(), day_start,
                     day_start + day_count - 1);
        return 1;
      }
      d_begin = h.day - day_start;
      s_begin = h.next_seg;
      restore = s_begin > 0;
      if (restore) {
        /* drop text rows written after the checkpoint */
        std::string tagday = std::to_string(h.day) + "day" + btag;
        const std::string names[4] = {out_dir + "Viss" + tagday + ".txt", out_dir + "uvw" + tagday + ".txt",
                                      out_dir + "xyza" + tagday + ".txt", out_dir + "xyzb" + tagday + ".txt"};
        for (int k = 0; k < 4; ++k)
          if (h.fsize[k] >= 0 && truncate(names[k].c_str(), (off_t)h.fsize[k]) != 0) {
            std::fprintf(stderr, "ERROR cannot restore %s to its checkpointed size\n", names[k].c_str());
            return 1;
          }
        ck_wtotal = h.wtotal;
      }
      std::printf("Resuming from %s: day %d segment %d\n", ckpt_path.c_str(), h.day, s_begin + 1);
    }
  }
  auto save_ckpt = [&](int dayid, int next_seg) -> bool {
    if (ckpt_path.empty()) return true;
    Ckpt h;
    std::memcpy(h.magic, kMagic, 8);
    h.fingerprint = fp.h;
    h.day = dayid;
    h.next_seg = next_seg;
    h.day_start = day_start;
    h.day_count = day_count;
    h.npix = npix;
    h.ngp = plan.n_img_pts;
    h.wtotal = 0;
    std::string tagday = std::to_string(dayid) + "day" + btag;
    const std::string names[4] = {out_dir + "Viss" + tagday + ".txt", out_dir + "uvw" + tagday + ".txt",
                                  out_dir + "xyza" + tagday + ".txt", out_dir + "xyzb" + tagday + ".txt"};
    for (int k = 0; k < 4; ++k) h.fsize[k] = next_seg > 0 ? file_size(names[k]) : -1;
    if (next_seg > 0) {
      ckC.resize((size_t)npix);
      ckW.resize((size_t)npix);
      ckG.resize((size_t)plan.n_img_pts);
      ckGW.resize((size_t)plan.n_img_pts);
      if (vw_get_accumulators(ctx, ckC.data(), ckW.data(), ckG.data(), ckGW.data(), &h.wtotal) != VW_OK) return false;
    }
    return write_ckpt(ckpt_path, h, ckC, ckW, ckG, ckGW);
  };
  int segs_done = 0;
  const int64_t max_nf = 2LL * plan.max_tlen * plan.n_pairs;
  std::vector<double> gbuf(9 * (size_t)max_nf);
  std::vector<vw_c128> vfull((size_t)max_nf);
  std::vector<double> Cmap((size_t)npix);
  std::vector<uint32_t> Wmap((size_t)npix);
  std::vector<double> img((size_t)cfg.img_n * cfg.img_n);

  for (int d = d_begin; d < day_count; ++d) {
    int dayid = day_start + d;
    std::printf("===== Day %d / %d =====\n", dayid, day_start + day_count - 1);
    if (d > 0 && vw_set_day(ctx, dayid) != VW_OK) return die(ctx, "vw_set_day");
    const bool resumed_day = restore && d == d_begin;
    if (resumed_day) {
      if (vw_set_accumulators(ctx, ckC.data(), ckW.data(), ckG.data(), ckGW.data(), ck_wtotal) != VW_OK)
        return die(ctx, "vw_set_accumulators");
    } else if (vw_reset_accumulators(ctx) != VW_OK) {
      return die(ctx, "vw_reset_accumulators");
    }
    std::string tagday = std::to_string(dayid) + "day" + btag;
    std::string fviss = out_dir + "Viss" + tagday + ".txt";
    std::string fuvw = out_dir + "uvw" + tagday + ".txt", fxa = out_dir + "xyza" + tagday + ".txt",
                fxb = out_dir + "xyzb" + tagday + ".txt";
    Timer t_rec;
    t_rec.tic();
    for (int s = (d == d_begin ? s_begin : 0); s < plan.segs; ++s) {
      int t0 = seg_t0[s], tlen = seg_tlen[s];
      int64_t segN = 2LL * tlen * plan.n_pairs, segN_half = (int64_t)tlen * plan.n_pairs;
      bool need_geom = write_baseline_txt || print_stats;
      vw_geom_host gh;
      double* gp = gbuf.data();
      gh.u = gp; gh.v = gp + segN; gh.w = gp + 2 * segN;
      gh.xa = gp + 3 * segN; gh.ya = gp + 4 * segN; gh.za = gp + 5 * segN;
      gh.xb = gp + 6 * segN; gh.yb = gp + 7 * segN; gh.zb = gp + 8 * segN;
      int rc = vw_gen_segment(ctx, t0, tlen, need_geom ? &gh : nullptr);
      if (rc == VW_ERR_HALFSYM) {
        std::fprintf(stderr, "ERROR: generated segment uvw does not satisfy half-sym at seg %d\n", s);
        return 1;
      }
      if (rc != VW_OK) return die(ctx, "vw_gen_segment");
      if (write_baseline_txt) {
        int app = s > 0;
        if (vw_io_write_cols3(fuvw.c_str(), gh.u, gh.v, gh.w, segN, app) ||
            vw_io_write_cols3(fxa.c_str(), gh.xa, gh.ya, gh.za, segN, app) ||
            vw_io_write_cols3(fxb.c_str(), gh.xb, gh.yb, gh.zb, segN, app)) {
          std::fprintf(stderr, "ERROR open generated baseline txt outputs\n");
          return 1;
        }
      }
      Timer tv;
      tv.tic();
      if (vw_viss_half(ctx, nullptr) != VW_OK) return die(ctx, "vw_viss_half");
      if (vw_conj_fill(ctx, nullptr, write_viss ? vfull.data() : nullptr) != VW_OK) return die(ctx, "vw_conj_fill");
      double tvs = tv.toc();
      if (write_viss && vw_io_write_viss(fviss.c_str(), vfull.data(), segN, s > 0) != VW_OK) {
        std::fprintf(stderr, "ERROR open %s\n", fviss.c_str());
        return 1;
      }
      Timer tr;
      tr.tic();
      vw_seg_info si;
      if (vw_recon_segment(ctx, &si) != VW_OK) return die(ctx, "vw_recon_segment");
      double trs = tr.toc();
      std::printf("[day %d seg %d/%d] t0=%d segN=%lld segN_half=%lld VissTime=%gs ReconSegTime=%gs fa=%g fb=%g "
                  "denom=%g nuniq=%lld RES=%lld",
                  dayid, s + 1, plan.segs, t0, (long long)segN, (long long)segN_half, tvs, trs, si.fa, si.fb,
                  si.denom, (long long)si.nuniq, (long long)si.res);
      if (print_stats) {
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; ++k)
          for (int64_t i = 0; i < segN; ++i) {
            double x = gp[k * segN + i];
            mn[k] = std::fmin(mn[k], x);
            mx[k] = std::fmax(mx[k], x);
          }
        std::printf(" u=[%g,%g] v=[%g,%g] w=[%g,%g]", mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]);
      }
      std::printf("\n");
      std::fflush(stdout);
      ++segs_done;
      if (s + 1 < plan.segs && !save_ckpt(dayid, s + 1)) {
        std::fprintf(stderr, "ERROR writing checkpoint %s\n", ckpt_path.c_str());
        return 1;
      }
      if (stop_after > 0 && segs_done >= stop_after && s + 1 < plan.segs) {
        std::printf("Stopped after %d segments%s%s\n", segs_done, ckpt_path.empty() ? "" : "; checkpoint ",
                    ckpt_path.c_str());
        vw_destroy(ctx);
        return 0;
      }
    }
    if (vw_finalize(ctx, Cmap.data(), Wmap.data(), cfg.img_n ? img.data() : nullptr, nullptr) != VW_OK)
      return die(ctx, "vw_finalize");
    std::printf("Recon done in %g s\n", t_rec.toc());
    Timer tw;
    tw.tic();
    std::string outC = out_dir + "C" + tagday + (C_mode == "txt" ? ".txt" : ".bin");
    if (vw_io_write_map(outC.c_str(), Cmap.data(), npix, C_mode == "txt") != VW_OK) {
      std::fprintf(stderr, "ERROR open %s\n", outC.c_str());
      return 1;
    }
    if (cfg.img_n) {
      std::string fi = out_dir + "img" + tagday + ".bin";
      FILE* f = std::fopen(fi.c_str(), "wb");
      if (!f || std::fwrite(img.data(), sizeof(double), img.size(), f) != img.size()) {
        std::fprintf(stderr, "ERROR open %s\n", fi.c_str());
        if (f) std::fclose(f);
        return 1;
      }
      std::fclose(f);
    }
    std::printf("Wrote %s in %g s\n", outC.c_str(), tw.toc());
    if (!save_ckpt(dayid + 1, 0)) {
      std::fprintf(stderr, "ERROR writing checkpoint %s\n", ckpt_path.c_str());
      return 1;
    }
    if (stop_after > 0 && segs_done >= stop_after && d + 1 < day_count) {
      std::printf("Stopped after %d segments%s%s\n", segs_done, ckpt_path.empty() ? "" : "; checkpoint ",
                  ckpt_path.c_str());
      break;
    }
  }
  vw_destroy(ctx);
  return 0;
}
