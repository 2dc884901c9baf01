/*
 * vissws.hip -- host orchestration and C-ABI of the gfx950 path
 * (include/vissws.h).  Three HIP streams per context, all at default
 * priority (geometry + Viss; grouping + HEALPix recon; lm-image recon); all
 * device memory is allocated once in vw_create and sized for the longest
 * segment, so the segment loop performs no allocation.  Stage functions
 * enqueue work; the public entry points synchronise before returning.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/vissws.h"
#include "vw_batch.h"
#include "vw_kernels.h"
#include "vw_shared.h"

#define VW_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_create_err;

struct EvPair {
  hipEvent_t a = nullptr, b = nullptr;
};

enum Stage { ST_GEN = 0, ST_VISS, ST_CONJ, ST_GROUP, ST_RECON, ST_FIN, ST_VISSK, ST_RECONK, ST_N };

struct Ctx {
  vw_config cfg{};
  std::string err;
  /* plan */
  int64_t nside = 0, npix = 0, TP = 256, ntile = 0;
  int A = 0, P = 0, T = 0, segs = 0, max_tlen = 0;
  std::vector<int32_t> seg_t0, seg_tlen;
  double lambda = 0, bl_max = 0, cosphi_glob = NAN, pix_area = 0;
  bool per_ant = false;
  int64_t orbit_res = 0, seg_len = 0, tw = 0, day_offset = 0, pc = 0;
  std::vector<int64_t> t1s;
  std::vector<double> circ; /* [7][A] */
  std::vector<int32_t> pa, pb;
  int64_t ngp = 0, ngt = 0;
  std::vector<int64_t> gidx;
  int64_t max_nh = 0, nh_pad = 0, nchunk_max = 1;
  /* device */
  hipStream_t st = nullptr;
  /* streams: st carries the segment geometry, Viss and conjugate fill; stB
   * the grouping and HEALPix recon; st2 the lm-image recon.  The recon of
   * segment s thus runs beside the Viss of segment s+1 (the segment
   * buffers alternate between two sets) */
  hipStream_t stB = nullptr;
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_conj = nullptr;
  hipEvent_t ev_recon[2] = {nullptr, nullptr}; /* recon done with segment buffer set b */
  int cur = 1;                                  /* buffer set of the current segment   */
  std::vector<void*> allocs;
  int64_t dev_bytes = 0; /* bytes of the allocations in allocs */
  double *d_l = nullptr, *d_m = nullptr, *d_n = nullptr, *d_nm1 = nullptr, *d_cone = nullptr;
  double* d_B = nullptr;
  float* d_Bf = nullptr;
  int64_t* d_gidx = nullptr;
  double *d_gl = nullptr, *d_gm = nullptr, *d_gn = nullptr, *d_gcone = nullptr;
  double* d_pos = nullptr;
  int64_t* d_t1s = nullptr;
  double* d_circ = nullptr;
  int32_t *d_pa = nullptr, *d_pb = nullptr;
  /* segment buffers of the current set (pointers into seg[cur]) */
  RowRec* d_rows = nullptr;
  double4* d_rowt = nullptr; /* half-row (u, v, w) * 1024 for the plain Viss phase */
  double2* d_tab = nullptr;  /* the plain kernels' sine table (VW_TAB_N entries) */
  double* d_ant8 = nullptr;
  struct SegSet {
    RowRec* rows;
    double4* rowt;
    double* ant8;
    double2 *vraw, *vh, *vfull;
  } seg[2] = {};
  uint8_t *d_cls_v = nullptr, *d_bits_v = nullptr; /* Viss class scratch */
  uint8_t* d_cls = nullptr; /* class scratch shared by all tile passes (cls_cap bytes) */
  int64_t cls_cap = 0;       /* bytes                                                    */
  int64_t vtpp = 0;          /* sky tiles per Viss pass                                  */
  int64_t rpart_cap = 0;     /* doubles in each recon partial buffer                     */
  uint8_t* d_bits = nullptr; /* cone bits per (tile of a pass, antenna frame)            */
  int64_t nfr_pad = 0;       /* antenna frames of the longest segment, padded            */
  int64_t bits_cap = 0;      /* bytes of d_bits                                          */
  double2* d_part = nullptr;
  int64_t* d_cpart = nullptr;
  /* antenna-factorized Viss: per-frame phase coefficients and the partial
   * product blocks [tlen][chunk][block][256] */
  bool vfact = false;
  double4* d_aph = nullptr;
  double* d_fpart = nullptr;
  double* d_fesum = nullptr; /* [tlen][block][256] epoch sums of the partial blocks */
  int64_t fpart_cap = 0;
  int64_t* d_counts = nullptr;
  double2 *d_vraw = nullptr, *d_vh = nullptr, *d_vfull = nullptr;
  int64_t *d_keys = nullptr, *d_keys_sorted = nullptr, *d_gkeys = nullptr;
  int32_t *d_idx_sorted = nullptr, *d_goff = nullptr, *d_rank = nullptr;
  GroupSoA gs{};
  double* d_rpart = nullptr;
  uint32_t* d_rwpart = nullptr;
  /* scratch of the image recon set (the HEALPix set uses d_cls, d_bits,
   * d_rpart, d_rwpart) so the two sets can run concurrently */
  uint8_t *d_cls2 = nullptr, *d_bits2 = nullptr;
  double* d_rpart2 = nullptr;
  uint32_t* d_rwpart2 = nullptr;
  int64_t nq_pad = 0;
  bool generic = false; /* recon needs per-row classes of the image tiles */
  DevInfo* d_info = nullptr;
  /* sky batch (nsky > 1): packed sky operand, Viss chunk partials, packed
   * group visibilities, recon partials of both output sets */
  bool batch = false;
  int64_t nsky = 1, nsky_pad = 16, nsg = 1;
  int nbw = 1;                          /* 16-sky blocks per wave (1..4)         */
  int64_t vb_target = 8192, rb_target = 16384; /* waves per launch (autotuned)  */
  float* d_skyf = nullptr;
  __bf16 *d_skyh = nullptr, *d_skyl = nullptr;
  double2* d_bpart = nullptr;
  int64_t bpart_cap = 0; /* double2 entries */
  __bf16* d_gpk[4] = {};  /* re hi, re lo, im hi, im lo: [nq_pad / 32][nsky_pad][32] */
  double *d_brpart = nullptr, *d_brpart2 = nullptr;
  uint32_t *d_bwpart = nullptr, *d_bwpart2 = nullptr;
  int64_t brpart_cap = 0, bwpart_cap = 0;
  /* work orders of the batched kernels: keys, indices, sorted keys, order
   * (rows on the Viss stream, groups on the grouping stream) */
  int32_t *d_okey[2] = {}, *d_oidx[2] = {}, *d_okey2[2] = {}, *d_order[2] = {};
  double2* d_orec = nullptr; /* (ugu, vgu) of the group at each work-order position */
  int32_t *d_oia = nullptr, *d_oib = nullptr; /* its representative's antenna frames */
  double* d_oant = nullptr;  /* its representative's antenna frames [position][10]    */
  int32_t* d_rshist[2] = {};  /* radix-sort digit tables per stream (Viss, grouping)   */
  int64_t rshist_cap = 0;
  /* per-sky segment schedules (vw_set_sky_schedules): the batched Viss of
   * the base schedule's segments is shared by every sky and kept for the
   * whole day; each sky is then grouped and reconstructed over its own
   * segments, with its own weights */
  bool sky_sched = false;
  struct SkySeg {
    int t0, tlen;
    std::vector<int32_t> skies; /* skies whose schedule holds this segment */
  };
  std::vector<SkySeg> sky_segs;  /* distinct (t0, tlen) of all skies, by t0 then tlen */
  double2* d_vday = nullptr;      /* [nsky][T P] phase-corrected half visibilities of the day */
  uint32_t *d_Wsky = nullptr, *d_GWsky = nullptr; /* per-sky weights [nsky][npix], [nsky][ngp] */
  uint32_t* d_iwout_sky = nullptr; /* per-sky image weights [nsky][img_n^2] (finalize) */
  /* adjoint Viss (vw_viss_adjoint): residual rows, chunk partials, output */
  double2* d_adj_res = nullptr;
  double *d_adj_part = nullptr, *d_adj_out = nullptr;
  int64_t adj_part_cap = 0;
  /* graph of a whole batched step (reset + every segment), captured on first use */
  hipGraphExec_t graph_exec = nullptr;
  bool capturing = false;
  double *d_Cacc = nullptr, *d_Gacc = nullptr;
  uint32_t *d_Wacc = nullptr, *d_GWacc = nullptr;
  double *d_out = nullptr, *d_iout = nullptr;
  uint32_t *d_wout = nullptr, *d_iwout = nullptr;
  double* d_full = nullptr;
  /* segment state */
  int t0 = -1, tlen = 0;
  int64_t nh = 0;
  bool have_sky = false, have_seg = false, have_viss = false, have_conj = false, have_groups = false;
  /* timing */
  std::vector<EvPair> ev[ST_N];
  size_t ev_used[ST_N] = {};
  vw_timings tim{};

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

#define HC(x)                                                                              \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define KC() HC(hipGetLastError())

template <typename T>
int dalloc(Ctx* c, T** p, size_t n) {
  void* q = nullptr;
  HC(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
  c->allocs.push_back(q);
  c->dev_bytes += (int64_t)((n ? n : 1) * sizeof(T));
  *p = (T*)q;
  return VW_OK;
}

int round_away_f(float x) { return (x >= 0.0f) ? (int)floorf(x + 0.5f) : (int)ceilf(x - 0.5f); }

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

/* event bracket for one stage launch group */
hipEvent_t ev_begin(Ctx* c, Stage s, hipStream_t stream = nullptr) {
  if (c->capturing) return nullptr; /* no timing inside a captured graph */
  auto& v = c->ev[s];
  if (c->ev_used[s] == v.size()) {
    EvPair p;
    (void)hipEventCreate(&p.a);
    (void)hipEventCreate(&p.b);
    v.push_back(p);
  }
  EvPair& p = v[c->ev_used[s]];
  (void)hipEventRecord(p.a, stream ? stream : c->st);
  return p.b;
}
void ev_end(Ctx* c, Stage s, hipStream_t stream = nullptr) {
  if (c->capturing) return;
  (void)hipEventRecord(c->ev[s][c->ev_used[s]].b, stream ? stream : c->st);
  c->ev_used[s]++;
}

/* after a synchronise: fold recorded brackets into the timings */
void collect_timings(Ctx* c) {
  double* dst[ST_N] = {&c->tim.gen_ms, &c->tim.viss_ms, &c->tim.conj_ms, &c->tim.group_ms,
                       &c->tim.recon_ms, &c->tim.finalize_ms, &c->tim.viss_kernel_ms, &c->tim.recon_kernel_ms};
  for (int s = 0; s < ST_N; ++s) {
    for (size_t k = 0; k < c->ev_used[s]; ++k) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, c->ev[s][k].a, c->ev[s][k].b);
      *dst[s] += ms;
    }
    if (s == ST_VISSK) c->tim.n_viss_launch += (int32_t)c->ev_used[s];
    if (s == ST_RECONK) c->tim.n_recon_launch += (int32_t)c->ev_used[s];
    c->ev_used[s] = 0;
  }
}

int sync(Ctx* c) {
  HC(hipStreamSynchronize(c->st));
  if (c->stB && c->stB != c->st) HC(hipStreamSynchronize(c->stB));
  if (c->st2 && c->st2 != c->st) HC(hipStreamSynchronize(c->st2));
  collect_timings(c);
  return VW_OK;
}

/* ---------------- plan ---------------- */
int make_plan(Ctx* c, const vw_config* cfg) {
  c->cfg = *cfg;
  c->cfg.seg_tlen = nullptr;
  if (cfg->nside <= 0 || cfg->nside > (1 << 20) || !(cfg->freq_hz > 0)) return c->fail(VW_ERR_ARG, "bad nside/freq");
  if (cfg->tile_pix != 256 && cfg->tile_pix != 512) return c->fail(VW_ERR_ARG, "tile_pix must be 256 or 512");
  if (cfg->geometry != VW_GEOM_LUNAR8 && cfg->geometry != VW_GEOM_CIRCULAR) return c->fail(VW_ERR_ARG, "geometry");
  if (cfg->geometry == VW_GEOM_LUNAR8 && cfg->n_ant != 8) return c->fail(VW_ERR_ARG, "lunar geometry has 8 antennas");
  if (cfg->n_ant < 2 || cfg->n_ant > 256) return c->fail(VW_ERR_ARG, "n_ant out of range");
  if (cfg->precision != VW_PREC_PLAIN && cfg->precision != VW_PREC_MIXED) return c->fail(VW_ERR_ARG, "precision");
  if (cfg->recon_mode != VW_RECON_MIRROR && cfg->recon_mode != VW_RECON_FULL) return c->fail(VW_ERR_ARG, "recon_mode");
  if (cfg->blockage != 0 && cfg->blockage != 1) return c->fail(VW_ERR_ARG, "blockage must be 0 or 1");
  if (cfg->img_n < 0 || cfg->img_n > 16384) return c->fail(VW_ERR_ARG, "img_n out of range");
  if (cfg->viss_form < VW_VISS_AUTO || cfg->viss_form > VW_VISS_ANT_CHECK) return c->fail(VW_ERR_ARG, "viss_form");
  if ((cfg->viss_form == VW_VISS_ANT || cfg->viss_form == VW_VISS_ANT_CHECK) && cfg->n_ant > 32)
    return c->fail(VW_ERR_ARG, "the antenna-factorized Viss form takes at most 32 antennas");
  c->nside = cfg->nside;
  c->npix = 12 * c->nside * c->nside;
  c->TP = cfg->tile_pix;
  c->ntile = (c->npix + c->TP - 1) / c->TP;
  c->A = cfg->n_ant;
  c->P = c->A * (c->A - 1) / 2;
  /* automatic choice: the factorized form for the mixed precision (per
   * segment, Nside 128 / 16 antennas 0.16 against 0.30 ms, Nside 256 / 32
   * antennas 3.0 against 7.0 ms); the plain form keeps the per-row phases */
  c->vfact = c->A <= 32 && (cfg->viss_form == VW_VISS_ANT || cfg->viss_form == VW_VISS_ANT_CHECK ||
                            (cfg->viss_form == VW_VISS_AUTO && cfg->precision == VW_PREC_MIXED));
  c->lambda = 3e8 / cfg->freq_hz;
  c->pix_area = (4.0 * VW_PI) / (double)c->npix;
  c->per_ant = cfg->occ_altitude_km < 0.0;
  c->cosphi_glob = c->per_ant ? NAN : vw_cosphi_from_altitude(cfg->occ_radius_km, cfg->occ_altitude_km);
  if (cfg->geometry == VW_GEOM_LUNAR8) {
    /* orbit_online.cuh:238-241 */
    c->orbit_res = (int64_t)std::ceil((2.0 * VW_PI) * (100e3 / c->lambda));
    c->pc = round_away_f(24.0f / 2.3f);
    c->seg_len = c->orbit_res / 3;
    c->T = (int)(c->pc * c->seg_len);
  } else {
    if (cfg->n_epochs <= 0) return c->fail(VW_ERR_ARG, "n_epochs must be > 0");
    if (!(cfg->r_max_km >= cfg->r_min_km) || !(cfg->r_min_km > 0) || !(cfg->gm_km3_s2 > 0))
      return c->fail(VW_ERR_ARG, "circular orbit parameters");
    c->T = cfg->n_epochs;
    /* per-antenna draws: radius, inclination, RAAN, phase (u01 as orbit_online.cuh:324) */
    std::mt19937_64 rng(cfg->orbit_seed);
    auto u01 = [&]() { return (double)(rng() >> 11) * (1.0 / 9007199254740992.0); };
    int A = c->A;
    c->circ.assign(7 * (size_t)A, 0.0);
    for (int a = 0; a < A; ++a) {
      double r = cfg->r_min_km + (cfg->r_max_km - cfg->r_min_km) * u01();
      double inc = 180.0 * u01();
      double raan = 360.0 * u01();
      double ph = 360.0 * u01();
      c->circ[a] = r;
      c->circ[A + a] = ph / 360.0;
      c->circ[2 * A + a] = std::sqrt(cfg->gm_km3_s2 / ((r * r) * r)) * VW_INV_TWOPI;
      vw_sincos_turn(raan / 360.0, &c->circ[3 * A + a], &c->circ[4 * A + a]);
      vw_sincos_turn(inc / 360.0, &c->circ[5 * A + a], &c->circ[6 * A + a]);
    }
  }
  /* schedule: explicit or equal split (main_recon_online.cu:123-139) */
  c->seg_t0.clear();
  c->seg_tlen.clear();
  if (cfg->n_seg_explicit > 0) {
    if (!cfg->seg_tlen) return c->fail(VW_ERR_ARG, "seg_tlen is NULL");
    int64_t cur = 0;
    for (int k = 0; k < cfg->n_seg_explicit; ++k) {
      if (cfg->seg_tlen[k] <= 0) return c->fail(VW_ERR_ARG, "segment length <= 0");
      c->seg_t0.push_back((int32_t)cur);
      c->seg_tlen.push_back(cfg->seg_tlen[k]);
      cur += cfg->seg_tlen[k];
    }
    if (cur > c->T) return c->fail(VW_ERR_ARG, "segments exceed the epochs of the run");
  } else {
    int segs = cfg->segs <= 0 ? 1 : cfg->segs;
    if (segs > c->T) segs = c->T;
    int base = c->T / segs, rem = c->T % segs, cur = 0;
    for (int k = 0; k < segs; ++k) {
      int len = base + (k < rem ? 1 : 0);
      c->seg_t0.push_back(cur);
      c->seg_tlen.push_back(len);
      cur += len;
    }
  }
  c->segs = (int)c->seg_t0.size();
  c->max_tlen = *std::max_element(c->seg_tlen.begin(), c->seg_tlen.end());
  c->pa.resize(c->P);
  c->pb.resize(c->P);
  for (int mm = 0, k = 0; mm < c->A - 1; ++mm)
    for (int nn = mm + 1; nn < c->A; ++nn, ++k) { c->pa[k] = nn; c->pb[k] = mm; }
  /* lm grid point order: 16x16 blocks, row-major blocks */
  c->gidx.clear();
  if (cfg->img_n > 0) {
    int64_t N = cfg->img_n;
    for (int64_t bi = 0; bi < N; bi += 16)
      for (int64_t bj = 0; bj < N; bj += 16)
        for (int64_t ii = bi; ii < bi + 16 && ii < N; ++ii)
          for (int64_t jj = bj; jj < bj + 16 && jj < N; ++jj) {
            int64_t a = 2 * jj + 1 - N, b = 2 * ii + 1 - N;
            if (a * a + b * b < N * N) c->gidx.push_back(ii * N + jj);
          }
  }
  c->ngp = (int64_t)c->gidx.size();
  c->ngt = (c->ngp + c->TP - 1) / c->TP;
  c->max_nh = (int64_t)c->max_tlen * c->P;
  c->nh_pad = (c->max_nh + 63) / 64 * 64;
  c->nq_pad = (2 * c->max_nh + 63) / 64 * 64;
  /* occ_mode 1 and 2 read one member per group: the representative kernel,
   * with the finite-value filter of the member sets except for the uv-mirror
   * occ_mode 2 form (recon_seg_blockage_half_rep); occ_mode 0 averages every
   * member and any other value the four samples s0, s0+L/3, s0+2L/3, s1-1 per
   * output point (the else branch of viss_recon_kernel.cuh:605-612) */
  c->generic = cfg->blockage && cfg->occ_mode != 1 && cfg->occ_mode != 2;
  c->nsky = cfg->nsky > 1 ? cfg->nsky : 1;
  c->batch = c->nsky > 1;
  if (c->batch) {
    if (cfg->nsky > 4096) return c->fail(VW_ERR_ARG, "nsky above 4096");
    if (cfg->precision != VW_PREC_MIXED) return c->fail(VW_ERR_ARG, "batched skies (nsky > 1) need the mixed precision");
    if (c->generic)
      return c->fail(VW_ERR_ARG, "batched skies need a recon rule with one member per group (occ_mode 1 or 2) or blockage 0");
    c->vfact = false;
    const int64_t b16 = (c->nsky + 15) / 16;
    c->nbw = (int)std::min<int64_t>(4, b16);
    c->nsg = (b16 + c->nbw - 1) / c->nbw;
    c->nsky_pad = c->nsg * c->nbw * 16;
  }
  if (c->max_nh > (int64_t)INT32_MAX / 4) return c->fail(VW_ERR_RANGE, "segment too long for 32-bit row indices");
  return VW_OK;
}

/* waves per launch the grid shapes aim for.  256 CUs hold 3,072 - 5,120
 * waves of these kernels at a time and the work per wave varies with the
 * culled fraction, so a launch needs many rounds of waves to keep the last
 * round's idle tail small (measured at Nside 128: 8k -> 16k waves took
 * Viss from 8.25 to 7.9 ms, 8k -> 32k took recon from 16.2 to 13.8 ms) */
#ifndef VW_VISS_TARGET_WAVES
#define VW_VISS_TARGET_WAVES 16384
#endif
#ifndef VW_RECON_TARGET_WAVES
#define VW_RECON_TARGET_WAVES 32768
#endif
#ifndef VW_RECON_MAX_CHUNKS
#define VW_RECON_MAX_CHUNKS 128
#endif
#ifndef VW_RECON_MIN_CHUNKS
#define VW_RECON_MIN_CHUNKS 24
#endif

/* Viss launch shape for nh half rows over a pass of nt tiles */
struct VissShape {
  int rpw;
  int64_t tpc, nchunk, nrb, bpc, nblocks;
};
/* half rows per wave: as many accumulators as fit without spilling
 * (the counting variants carry one more register per row); the plain form
 * also holds four weight coefficients per pixel */
#ifndef VW_VISS_RPW_PLAIN
#define VW_VISS_RPW_PLAIN 8
#endif
#ifndef VW_VISS_RPW_MIXED
#define VW_VISS_RPW_MIXED 16
#endif
/* frame rows up to this many bytes are staged in LDS by the group classes
 * (within the 64 KB of dynamic LDS a launch may ask for by default) */
#define VW_GCLS_LDS_MAX 32768
constexpr int viss_rpw(int /*TP*/, bool mixed, bool count) {
  return (mixed ? VW_VISS_RPW_MIXED : VW_VISS_RPW_PLAIN) / (count ? 2 : 1);
}

/* The launch shapes take lim, the chunk count the context's partial buffers
 * hold (0 while they are being sized): a segment shorter than the schedule's
 * would ask for more chunks than the longest one, and gets lim chunks of more
 * tiles each instead.  lim is at least the chunk count of a full tile pass of
 * the longest segment, so the tiles-per-chunk caps below still hold. */
VissShape viss_shape(const Ctx* c, int64_t nh, int64_t nt, bool count = false, int64_t lim = 0) {
  VissShape s;
  bool mixed = c->cfg.precision == VW_PREC_MIXED;
  s.rpw = viss_rpw(c->TP, mixed, count);
  s.nrb = (nh + s.rpw - 1) / s.rpw;
  const int64_t target_waves = VW_VISS_TARGET_WAVES;
  int64_t nchunk = (target_waves + s.nrb - 1) / s.nrb;
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, nt));
  if (lim > 0) nchunk = std::min(nchunk, lim);
  s.tpc = (nt + nchunk - 1) / nchunk;
  const int64_t tpc_cap = mixed ? 16 : 64; /* bounds the FP32 block length per lane */
  if (s.tpc > tpc_cap) s.tpc = tpc_cap;
  s.nchunk = (nt + s.tpc - 1) / s.tpc;
  s.bpc = (s.nrb + 3) / 4;
  s.nblocks = s.nchunk * s.bpc;
  return s;
}

/* factorized Viss launch shape: a wave owns one epoch and a chunk of tiles,
 * a workgroup four epochs of the same chunk (they share the pixel loads) */
#ifndef VW_VISSF_TARGET_WAVES
#define VW_VISSF_TARGET_WAVES 8192
#endif
struct VissFShape {
  int64_t tpc, nchunk, nquad, nblocks;
};
inline int vissf_groups(int A) { return A <= 16 ? 1 : 2; }
inline int64_t vissf_blocks(int G) { return (int64_t)G * (G + 1) / 2 + (int64_t)G * G; }
VissFShape viss_fshape(const Ctx* c, int64_t tlen, int64_t nt, int64_t lim = 0) {
  VissFShape s;
  int64_t nchunk = (VW_VISSF_TARGET_WAVES + tlen - 1) / tlen;
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, nt));
  if (lim > 0) nchunk = std::min(nchunk, lim);
  s.tpc = (nt + nchunk - 1) / nchunk;
  s.nchunk = (nt + s.tpc - 1) / s.tpc;
  s.nquad = (tlen + 3) / 4;
  s.nblocks = s.nchunk * s.nquad;
  return s;
}

/* recon launch shape: PPL points per lane, group ranges to fill the GPU */
#ifndef VW_RECON_PPL
#define VW_RECON_PPL 4
#endif
constexpr int RECON_PPL = VW_RECON_PPL;
struct ReconShape {
  int64_t pwaves, nchunk;
};
ReconShape recon_shape(int64_t npts) {
  ReconShape r;
  r.pwaves = (npts + 64 * RECON_PPL - 1) / (64 * RECON_PPL);
  const int64_t target_waves = VW_RECON_TARGET_WAVES;
  r.nchunk = std::max<int64_t>(1, std::min<int64_t>(VW_RECON_MAX_CHUNKS, (target_waves + r.pwaves - 1) / r.pwaves));
  /* large point sets get up to VW_RECON_MIN_CHUNKS group ranges while the
   * launch stays within twice the target (Nside 256: 21 ranges, not 11,
   * recon -4 to -5% per launch); the largest keep the target's count */
  const int64_t lo = std::min<int64_t>(VW_RECON_MIN_CHUNKS, 2 * target_waves / r.pwaves);
  r.nchunk = std::max(r.nchunk, lo);
  return r;
}

/* output tiles per recon pass: as many as the class scratch holds */
int64_t recon_pass_tiles(const Ctx* c, int64_t ntile_all, int64_t kind_w) {
  return std::max<int64_t>(1, std::min<int64_t>(ntile_all, c->cls_cap / kind_w));
}

/* batched launch shapes.  Viss: a wave is 16 half rows x 16 nbw skies over a
 * chunk of tiles, a workgroup 4 waves (64 rows) of one chunk; recon: a wave
 * is 16 output points x 16 nbw skies over a range of groups.  The chunk
 * counts follow a target of waves per launch, which the autotuner picks
 * from VW_BATCH_TARGETS (vw_autotune). */
#define VW_BATCH_MAX_CHUNKS 64
constexpr int64_t VW_BATCH_TARGETS[] = {2048, 4096, 8192, 16384, 32768, 65536};
struct VissBShape {
  int64_t nrb, nrbw, tpc, nchunk, nblocks;
};
VissBShape vissb_shape(const Ctx* c, int64_t nh, int64_t nt, int64_t target, int64_t lim = 0) {
  VissBShape s;
  s.nrb = (nh + 15) / 16;
  s.nrbw = (s.nrb + 3) / 4;
  int64_t nchunk = (target + s.nrb * c->nsg - 1) / (s.nrb * c->nsg);
  nchunk = std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(nchunk, nt), VW_BATCH_MAX_CHUNKS));
  if (lim > 0) nchunk = std::min(nchunk, lim);
  s.tpc = (nt + nchunk - 1) / nchunk;
  s.nchunk = (nt + s.tpc - 1) / s.tpc;
  s.nblocks = s.nchunk * s.nrbw * c->nsg;
  return s;
}
struct ReconBShape {
  int64_t pblk, pbw, nchunk, nblocks;
};
ReconBShape reconb_shape(const Ctx* c, int64_t np, int64_t target) {
  ReconBShape s;
  s.pblk = (np + 15) / 16;
  s.pbw = (s.pblk + 3) / 4;
  s.nchunk = std::max<int64_t>(1, std::min<int64_t>(VW_BATCH_MAX_CHUNKS, (target + s.pblk * c->nsg - 1) / (s.pblk * c->nsg)));
  s.nblocks = s.nchunk * s.pbw * c->nsg;
  return s;
}

int upload_day(Ctx* c) {
  /* lunar t1s windows (orbit_online.cuh:317-333) */
  if (c->cfg.geometry == VW_GEOM_LUNAR8) {
    int64_t OR = c->orbit_res;
    c->tw = round_away_f((float)OR * 7.0f * 24.0f / 2.3f);
    int mod14 = ((c->cfg.day - 1) % 14 + 14) % 14;
    c->day_offset = round_away_f((float)mod14 * (float)OR * 24.0f / 2.3f);
    uint64_t day_seed = c->cfg.orbit_seed ^ (uint64_t)(int64_t)c->cfg.day * 0x9E3779B97F4A7C15ULL;
    std::mt19937_64 rng(day_seed);
    c->t1s.assign((size_t)c->pc, 0);
    for (int64_t g = 0; g < c->pc; ++g) {
      int64_t os = g * OR, oe = (g + 1) * OR - 1;
      double u01 = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);
      int64_t t1 = os + round_away_f((float)(u01 * (double)OR * (2.0 / 3.0)));
      int64_t t2 = t1 + c->seg_len - 1;
      if (t2 > oe) { t2 = oe; t1 = t2 - c->seg_len + 1; }
      if (t1 < os) t1 = os;
      if (t1 + c->seg_len - 1 > oe) t1 = oe - c->seg_len + 1;
      c->t1s[g] = t1;
    }
    HC(hipMemcpyAsync(c->d_t1s, c->t1s.data(), sizeof(int64_t) * c->t1s.size(), hipMemcpyHostToDevice, c->st));
  }
  DayArgs a{};
  a.geometry = c->cfg.geometry;
  a.T = c->T;
  a.A = c->A;
  a.orbit_res = c->orbit_res;
  a.seg_len = c->seg_len;
  a.tw = c->tw;
  a.day_offset = c->day_offset;
  a.day = c->cfg.day;
  a.t1s = c->d_t1s;
  a.circ = c->d_circ;
  a.dt = c->cfg.dt_s;
  int64_t TA = (int64_t)c->T * c->A;
  k_day_positions<<<blocks_for(TA), 256, 0, c->st>>>(a, c->d_pos);
  KC();
  /* automatic bl_max: lunar 100 km (main_recon_online.cu:90); circular: the
   * longest 3-D baseline of the run (exact max of device positions) */
  if (c->cfg.bl_max_m > 0.0) {
    c->bl_max = c->cfg.bl_max_m;
  } else if (c->cfg.geometry == VW_GEOM_LUNAR8) {
    c->bl_max = 100e3;
  } else {
    std::vector<double> pos(3 * (size_t)TA);
    HC(hipMemcpyAsync(pos.data(), c->d_pos, sizeof(double) * pos.size(), hipMemcpyDeviceToHost, c->st));
    HC(hipStreamSynchronize(c->st));
    double mx = 0.0;
    for (int t = 0; t < c->T; ++t)
      for (int j = 0; j < c->P; ++j) {
        int64_t ka = (int64_t)t * c->A + c->pa[j], kb = (int64_t)t * c->A + c->pb[j];
        double dx = pos[ka] - pos[kb], dy = pos[TA + ka] - pos[TA + kb], dz = pos[2 * TA + ka] - pos[2 * TA + kb];
        mx = std::fmax(mx, std::sqrt(vw_dot3(dx, dy, dz, dx, dy, dz)) * 1000.0);
      }
    c->bl_max = mx;
  }
  return VW_OK;
}

/* make buffer set b the current segment's */
void select_seg(Ctx* c, int b) {
  c->cur = b;
  const Ctx::SegSet& g = c->seg[b];
  c->d_rows = g.rows;
  c->d_rowt = g.rowt;
  c->d_ant8 = g.ant8;
  c->d_vraw = g.vraw;
  c->d_vh = g.vh;
  c->d_vfull = g.vfull;
}

/* a captured step is rebuilt on its next use */
void drop_graph(Ctx* c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  c->graph_exec = nullptr;
}

/* the main stream waits for every recon enqueued so far */
int join_recon(Ctx* c) {
  for (int b = 0; b < 2; ++b) HC(hipStreamWaitEvent(c->st, c->ev_recon[b], 0));
  return VW_OK;
}

/* device buffers of the batched form, sized for every launch shape the
 * autotuner may choose */
int create_batch(Ctx* c) {
  int rc;
  const int64_t npix_pad = c->ntile * c->TP;
  if ((rc = dalloc(c, &c->d_skyf, (size_t)(c->nsky * c->npix))) ||
      (rc = dalloc(c, &c->d_skyh, (size_t)(npix_pad * c->nsky_pad))) ||
      (rc = dalloc(c, &c->d_skyl, (size_t)(npix_pad * c->nsky_pad))))
    return rc;
  for (int k = 0; k < 4; ++k)
    if ((rc = dalloc(c, &c->d_gpk[k], (size_t)(c->nq_pad * c->nsky_pad)))) return rc;
  int64_t vch = 1;
  const int64_t vlast = c->ntile - (c->ntile - 1) / c->vtpp * c->vtpp;
  for (int k = 0; k < c->segs; ++k)
    for (int64_t nt : {c->vtpp, vlast})
      for (int64_t tg : VW_BATCH_TARGETS) vch = std::max(vch, vissb_shape(c, (int64_t)c->seg_tlen[k] * c->P, nt, tg).nchunk);
  c->bpart_cap = vch * c->nsky_pad * c->nh_pad;
  if ((rc = dalloc(c, &c->d_bpart, (size_t)c->bpart_cap))) return rc;
  const int64_t kind_w = c->nq_pad;
  int64_t rp = 1, wp = 1;
  for (int img = 0; img < 2; ++img) {
    const int64_t npts = img ? c->ngp : c->npix, nta = img ? c->ngt : c->ntile;
    if (npts <= 0) continue;
    const int64_t pp = recon_pass_tiles(c, nta, kind_w) * c->TP;
    const int64_t last = npts - (npts - 1) / pp * pp;
    for (int64_t n : {std::min(npts, pp), last})
      for (int64_t tg : VW_BATCH_TARGETS) {
        const int64_t nch = reconb_shape(c, n, tg).nchunk;
        rp = std::max(rp, nch * c->nsky_pad * n);
        wp = std::max(wp, nch * n);
      }
  }
  c->brpart_cap = rp;
  c->bwpart_cap = wp;
  for (int w = 0; w < 2; ++w) {
    const int64_t n = w == 0 ? c->nh_pad : c->nq_pad;
    if ((rc = dalloc(c, &c->d_okey[w], (size_t)n)) || (rc = dalloc(c, &c->d_oidx[w], (size_t)n)) ||
        (rc = dalloc(c, &c->d_okey2[w], (size_t)n)) || (rc = dalloc(c, &c->d_order[w], (size_t)n)))
      return rc;
  }
  if ((rc = dalloc(c, &c->d_orec, (size_t)c->nq_pad)) || (rc = dalloc(c, &c->d_oant, 10 * (size_t)c->nq_pad)) ||
      (rc = dalloc(c, &c->d_oia, (size_t)c->nq_pad)) || (rc = dalloc(c, &c->d_oib, (size_t)c->nq_pad)))
    return rc;
  {
    const int64_t nmax = std::max<int64_t>(c->nq_pad, 2 * c->max_nh);
    c->rshist_cap = 256 * ((nmax + VW_RS_BLOCK - 1) / VW_RS_BLOCK);
    for (int w = 0; w < 2; ++w)
      if ((rc = dalloc(c, &c->d_rshist[w], (size_t)c->rshist_cap + 256))) return rc; /* + digit totals */
  }
  if ((rc = dalloc(c, &c->d_brpart, (size_t)rp)) || (rc = dalloc(c, &c->d_bwpart, (size_t)wp))) return rc;
  if (c->ngp > 0 && ((rc = dalloc(c, &c->d_brpart2, (size_t)rp)) || (rc = dalloc(c, &c->d_bwpart2, (size_t)wp))))
    return rc;
  return VW_OK;
}

int create_device(Ctx* c) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return c->fail(VW_ERR_DEVICE, "no HIP device");
  HC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  if (c->cfg.serial_stages) {
    c->st2 = c->stB = c->st; /* one stream: the stream waits below are no-ops */
  } else {
    HC(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
    HC(hipStreamCreateWithFlags(&c->stB, hipStreamNonBlocking));
  }
  HC(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  HC(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  HC(hipEventCreateWithFlags(&c->ev_conj, hipEventDisableTiming));
  for (int b = 0; b < 2; ++b) HC(hipEventCreateWithFlags(&c->ev_recon[b], hipEventDisableTiming));
  int rc;
  const size_t np = (size_t)c->npix;
  if ((rc = dalloc(c, &c->d_l, np)) || (rc = dalloc(c, &c->d_m, np)) || (rc = dalloc(c, &c->d_n, np)) ||
      (rc = dalloc(c, &c->d_nm1, np)) || (rc = dalloc(c, &c->d_cone, 5 * (size_t)c->ntile)) ||
      (rc = dalloc(c, &c->d_B, np)) || (rc = dalloc(c, &c->d_Bf, np)) ||
      (rc = dalloc(c, &c->d_Cacc, np * (size_t)c->nsky)) || (rc = dalloc(c, &c->d_Wacc, np)))
    return rc;
  if ((rc = dalloc(c, &c->d_tab, (size_t)VW_TAB_N))) return rc;
  k_tab_build<<<(VW_TAB_N + 255) / 256, 256, 0, c->st>>>(c->d_tab);
  k_pix2lmn<<<blocks_for(c->npix), 256, 0, c->st>>>(c->nside, c->npix, c->d_l, c->d_m, c->d_n, c->d_nm1);
  KC();
  if (c->TP == 256) k_tile_cone<256><<<(unsigned)c->ntile, 256, 0, c->st>>>(c->d_l, c->d_m, c->d_n, c->npix, c->ntile, c->d_cone);
  else k_tile_cone<512><<<(unsigned)c->ntile, 256, 0, c->st>>>(c->d_l, c->d_m, c->d_n, c->npix, c->ntile, c->d_cone);
  KC();
  if (c->ngp > 0) {
    const size_t ng = (size_t)c->ngp;
    if ((rc = dalloc(c, &c->d_gidx, ng)) || (rc = dalloc(c, &c->d_gl, ng)) || (rc = dalloc(c, &c->d_gm, ng)) ||
        (rc = dalloc(c, &c->d_gn, ng)) || (rc = dalloc(c, &c->d_gcone, 5 * (size_t)c->ngt)) ||
        (rc = dalloc(c, &c->d_Gacc, ng * (size_t)c->nsky)) || (rc = dalloc(c, &c->d_GWacc, ng)))
      return rc;
    HC(hipMemcpyAsync(c->d_gidx, c->gidx.data(), sizeof(int64_t) * ng, hipMemcpyHostToDevice, c->st));
    k_grid_points<<<blocks_for(c->ngp), 256, 0, c->st>>>(c->cfg.img_n, c->ngp, c->d_gidx, c->d_gl, c->d_gm, c->d_gn);
    KC();
    if (c->TP == 256) k_tile_cone<256><<<(unsigned)c->ngt, 256, 0, c->st>>>(c->d_gl, c->d_gm, c->d_gn, c->ngp, c->ngt, c->d_gcone);
    else k_tile_cone<512><<<(unsigned)c->ngt, 256, 0, c->st>>>(c->d_gl, c->d_gm, c->d_gn, c->ngp, c->ngt, c->d_gcone);
    KC();
  }
  int64_t TA = (int64_t)c->T * c->A;
  if ((rc = dalloc(c, &c->d_pos, 3 * (size_t)TA)) || (rc = dalloc(c, &c->d_t1s, (size_t)std::max<int64_t>(c->pc, 1))) ||
      (rc = dalloc(c, &c->d_circ, 7 * (size_t)c->A)) || (rc = dalloc(c, &c->d_pa, (size_t)c->P)) ||
      (rc = dalloc(c, &c->d_pb, (size_t)c->P)))
    return rc;
  HC(hipMemcpyAsync(c->d_pa, c->pa.data(), sizeof(int32_t) * c->P, hipMemcpyHostToDevice, c->st));
  HC(hipMemcpyAsync(c->d_pb, c->pb.data(), sizeof(int32_t) * c->P, hipMemcpyHostToDevice, c->st));
  if (!c->circ.empty())
    HC(hipMemcpyAsync(c->d_circ, c->circ.data(), sizeof(double) * c->circ.size(), hipMemcpyHostToDevice, c->st));
  /* segment buffers sized for the longest segment; one class scratch buffer
   * serves every tile pass (rows x tiles for Viss, groups x tiles for recon) */
  const size_t mh = (size_t)c->max_nh;
  const int64_t kind_w = c->generic ? c->nh_pad : c->nq_pad; /* recon class row width */
  const int64_t need = std::max<int64_t>(c->ntile * c->nh_pad, std::max<int64_t>(c->ntile, c->ngt) * kind_w);
  const int64_t budget = c->cfg.class_budget_bytes > 0 ? c->cfg.class_budget_bytes : (int64_t)2 << 30;
  c->cls_cap = std::min<int64_t>(need, budget);
  if (c->cls_cap < std::max(c->nh_pad, kind_w)) c->cls_cap = std::max(c->nh_pad, kind_w);
  c->cls_cap = (c->cls_cap + 3) / 4 * 4;
  c->vtpp = std::max<int64_t>(1, std::min<int64_t>(c->ntile, c->cls_cap / c->nh_pad));
  /* partial buffers fit the largest pass shape: a full pass and the last one */
  int64_t nch = 1;
  const int64_t vlast = c->ntile - (c->ntile - 1) / c->vtpp * c->vtpp;
  for (int k = 0; k < c->segs; ++k)
    for (int64_t nt : {c->vtpp, vlast}) nch = std::max(nch, viss_shape(c, (int64_t)c->seg_tlen[k] * c->P, nt).nchunk);
  c->nchunk_max = nch;
  if (c->vfact) {
    int64_t fp = 1;
    for (int k = 0; k < c->segs; ++k)
      for (int64_t nt : {c->vtpp, vlast})
        fp = std::max(fp, (int64_t)c->seg_tlen[k] * viss_fshape(c, c->seg_tlen[k], nt).nchunk *
                              vissf_blocks(vissf_groups(c->A)) * 256);
    c->fpart_cap = fp;
    if ((rc = dalloc(c, &c->d_fpart, (size_t)fp)) || (rc = dalloc(c, &c->d_aph, (size_t)c->max_tlen * c->A)) ||
        (rc = dalloc(c, &c->d_fesum, (size_t)c->max_tlen * vissf_blocks(vissf_groups(c->A)) * 256)))
      return rc;
  }
  for (int b = 0; b < 2; ++b) {
    Ctx::SegSet& g = c->seg[b];
    const size_t msk = mh * (size_t)c->nsky; /* batched: [nsky][nh] */
    if ((rc = dalloc(c, &g.rows, mh)) || (rc = dalloc(c, &g.rowt, (size_t)c->nh_pad)) ||
        (rc = dalloc(c, &g.ant8, 8 * (size_t)c->max_tlen * c->A)) || (rc = dalloc(c, &g.vraw, msk)) ||
        (rc = dalloc(c, &g.vh, msk)) || (rc = dalloc(c, &g.vfull, 2 * msk)))
      return rc;
  }
  select_seg(c, 1);
  if ((rc = dalloc(c, &c->d_cls, (size_t)c->cls_cap)) || (rc = dalloc(c, &c->d_cls_v, (size_t)c->cls_cap)) ||
      (rc = dalloc(c, &c->d_part, (size_t)nch * c->nh_pad)) ||
      (rc = dalloc(c, &c->d_keys, 2 * mh)) || (rc = dalloc(c, &c->d_keys_sorted, 2 * mh)) ||
      (rc = dalloc(c, &c->d_gkeys, 2 * mh + 1)) || (rc = dalloc(c, &c->d_idx_sorted, 2 * mh)) ||
      (rc = dalloc(c, &c->d_goff, 2 * mh + 2)) || (rc = dalloc(c, &c->d_rank, 2 * mh)) ||
      (rc = dalloc(c, &c->d_info, 1)) ||
      (rc = dalloc(c, &c->d_out, (size_t)c->npix * c->nsky)) || (rc = dalloc(c, &c->d_wout, (size_t)c->npix)) ||
      (rc = dalloc(c, &c->d_iout, (size_t)std::max<int64_t>(1, (int64_t)c->cfg.img_n * c->cfg.img_n) * c->nsky)) ||
      (rc = dalloc(c, &c->d_iwout, (size_t)std::max<int64_t>(1, (int64_t)c->cfg.img_n * c->cfg.img_n))))
    return rc;
  {
    const size_t nq = (size_t)c->nq_pad;
    GroupSoA& g = c->gs;
    if ((rc = dalloc(c, &g.rec, nq)) || (rc = dalloc(c, &g.ant, 10 * nq)) ||
        (rc = dalloc(c, &g.gia, nq)) || (rc = dalloc(c, &g.gib, nq)) || (rc = dalloc(c, &g.grep, nq)))
      return rc;
    int64_t rp = 1;
    for (int img = 0; img < 2; ++img) {
      const int64_t npts = img ? c->ngp : c->npix, nta = img ? c->ngt : c->ntile;
      if (npts <= 0) continue;
      const int64_t pp = recon_pass_tiles(c, nta, kind_w) * c->TP;
      const int64_t last = npts - (npts - 1) / pp * pp;
      for (int64_t n : {std::min(npts, pp), last}) rp = std::max(rp, recon_shape(n).nchunk * n);
    }
    c->rpart_cap = rp;
    if ((rc = dalloc(c, &c->d_rpart, (size_t)rp)) || (rc = dalloc(c, &c->d_rwpart, (size_t)rp))) return rc;
    /* cone bits for the widest tile pass of either stage */
    int64_t maxnt = c->vtpp;
    for (int img = 0; img < 2; ++img) {
      const int64_t nta = img ? c->ngt : c->ntile;
      if (nta > 0)
        for (int64_t w : {c->nh_pad, c->nq_pad}) maxnt = std::max(maxnt, recon_pass_tiles(c, nta, w));
    }
    if (!c->batch && 2 * c->max_nh > VW_RANK_MAX) {
      /* digit tables of the radix sort of long segments' (key, row) pairs
       * (batched contexts size theirs for the work orders too) */
      c->rshist_cap = 256 * ((2 * c->max_nh + VW_RS_BLOCK - 1) / VW_RS_BLOCK);
      if ((rc = dalloc(c, &c->d_rshist[1], (size_t)c->rshist_cap + 256))) return rc; /* + digit totals */
    }
    c->nfr_pad = ((int64_t)c->max_tlen * c->A + 63) / 64 * 64;
    c->bits_cap = maxnt * c->nfr_pad;
    if ((rc = dalloc(c, &c->d_bits, (size_t)c->bits_cap)) || (rc = dalloc(c, &c->d_bits_v, (size_t)c->bits_cap)))
      return rc;
    if (c->ngp > 0 && ((rc = dalloc(c, &c->d_bits2, (size_t)c->bits_cap)) ||
                       (rc = dalloc(c, &c->d_cls2, (size_t)c->cls_cap)) ||
                       (rc = dalloc(c, &c->d_rpart2, (size_t)c->rpart_cap)) ||
                       (rc = dalloc(c, &c->d_rwpart2, (size_t)c->rpart_cap))))
      return rc;
  }
  if (c->batch && (rc = create_batch(c))) return rc;
  HC(hipMemsetAsync(c->d_info, 0, sizeof(DevInfo), c->st));
  HC(hipMemsetAsync(c->d_Cacc, 0, sizeof(double) * np * c->nsky, c->st));
  HC(hipMemsetAsync(c->d_Wacc, 0, sizeof(uint32_t) * np, c->st));
  if (c->ngp) {
    HC(hipMemsetAsync(c->d_Gacc, 0, sizeof(double) * c->ngp * c->nsky, c->st));
    HC(hipMemsetAsync(c->d_GWacc, 0, sizeof(uint32_t) * c->ngp, c->st));
  }
  if ((rc = upload_day(c))) return rc;
  HC(hipStreamSynchronize(c->st));
  return VW_OK;
}

/* ---------------- stages (enqueue only) ---------------- */
int enq_gen(Ctx* c, int t0, int tlen) {
  if (t0 < 0 || tlen <= 0 || tlen > c->max_tlen || (int64_t)t0 + tlen > c->T)
    return c->fail(VW_ERR_ARG, "segment [t0, t0+tlen) out of range");
  c->t0 = t0;
  c->tlen = tlen;
  c->nh = (int64_t)tlen * c->P;
  /* the other buffer set: free once the recon of two segments ago is done */
  select_seg(c, c->cur ^ 1);
  HC(hipStreamWaitEvent(c->st, c->ev_recon[c->cur], 0));
  ev_begin(c, ST_GEN);
  int64_t na = (int64_t)tlen * c->A;
  k_seg_ant<<<blocks_for(na), 256, 0, c->st>>>(c->d_pos, c->T, c->A, t0, tlen, c->per_ant ? c->cfg.occ_radius_km : 0.0,
                                               c->cosphi_glob, c->d_ant8);
  KC();
  k_seg_rows<<<blocks_for(c->nh), 256, 0, c->st>>>(c->d_pos, c->T, c->A, c->P, t0, c->nh, c->d_pa, c->d_pb, c->lambda,
                                                   c->d_rows, c->d_rowt);
  KC();
  if (c->vfact) {
    k_ant_phase<<<blocks_for(na), 256, 0, c->st>>>(c->d_pos, c->T, c->A, t0, tlen, c->lambda, c->d_aph);
    KC();
  }
  ev_end(c, ST_GEN);
  c->have_seg = true;
  c->have_viss = c->have_conj = c->have_groups = false;
  return VW_OK;
}

/* row classes of nt tiles into cls ([nt][nh_pad] bytes) from their cone bits */
void launch_classify(Ctx* c, const uint8_t* bits, int64_t nt, uint8_t* cls, hipStream_t st) {
  const int nhb = (int)((c->nh_pad + VW_GCLS_BLOCK - 1) / VW_GCLS_BLOCK);
  const unsigned nb = (unsigned)(nt * nhb);
  if (c->nfr_pad <= VW_GCLS_LDS_MAX)
    k_classify<true><<<nb, 256, (size_t)c->nfr_pad, st>>>(bits, c->nfr_pad, nhb, c->d_rows, c->nh, c->nh_pad,
                                                           (uint32_t*)cls);
  else
    k_classify<false><<<nb, 256, 0, st>>>(bits, c->nfr_pad, nhb, c->d_rows, c->nh, c->nh_pad, (uint32_t*)cls);
}

/* class table of a tile pass into the class scratch: cone bits per
 * (tile, antenna frame), then per (tile, half row) or (tile, group) */
int enq_classes(Ctx* c, const double* cone, int64_t ntile_all, int64_t tile0, int64_t nt, bool groups,
                uint8_t* bits, uint8_t* cls, hipStream_t st, bool work_order = false) {
  const int64_t nfr = (int64_t)c->tlen * c->A;
  if (c->nfr_pad <= VW_GCLS_LDS_MAX) {
    /* one block per tile makes the tile's cone bits in LDS: no bit table */
    const size_t lds = (size_t)c->nfr_pad;
    if (groups)
      k_group_cls_tile<<<(unsigned)nt, 256, lds, st>>>(cone, ntile_all, tile0, c->d_ant8, nfr,
                                                       work_order ? c->d_oia : c->gs.gia,
                                                       work_order ? c->d_oib : c->gs.gib, c->d_info, c->nq_pad,
                                                       (uint32_t*)cls);
    else
      k_classify_tile<<<(unsigned)nt, 256, lds, st>>>(cone, ntile_all, tile0, c->d_ant8, nfr, c->d_rows, c->nh,
                                                      c->nh_pad, (uint32_t*)cls);
    KC();
    return VW_OK;
  }
  if (nt * c->nfr_pad > c->bits_cap) return c->fail(VW_ERR_RANGE, "internal: cone bit buffer too small");
  k_ant_bits<<<blocks_for(nt * c->nfr_pad), 256, 0, st>>>(cone, ntile_all, tile0, nt, c->d_ant8, nfr, c->nfr_pad, bits);
  KC();
  if (groups) {
    const int nqb = (int)((c->nq_pad + VW_GCLS_BLOCK - 1) / VW_GCLS_BLOCK);
    const unsigned nb = (unsigned)(nt * nqb);
    const int32_t* ia = work_order ? c->d_oia : c->gs.gia;
    const int32_t* ib = work_order ? c->d_oib : c->gs.gib;
    /* frame rows here are longer than the LDS staging limit */
    k_group_cls<false><<<nb, 256, 0, st>>>(bits, c->nfr_pad, nqb, ia, ib, c->d_info, c->nq_pad, (uint32_t*)cls);
  } else
    launch_classify(c, bits, nt, cls, st);
  KC();
  return VW_OK;
}

template <int TP, int RPW, bool MIXED, bool COUNT>
void launch_viss(const VissArgs& a, unsigned nb, hipStream_t st) {
  if constexpr (MIXED)
    k_viss<TP, RPW, COUNT><<<nb, 256, 0, st>>>(a);
  else
    k_viss_plain<TP, RPW, COUNT><<<nb, 256, 0, st>>>(a);
}

/* one tile pass of the antenna-factorized Viss: cone bits per (tile, frame),
 * the product blocks, then their fold into the raw half sums */
int enq_viss_fact(Ctx* c, int64_t tile0, int64_t nt) {
  const bool mixed = c->cfg.precision == VW_PREC_MIXED;
  const int G = vissf_groups(c->A);
  VissFShape s = viss_fshape(c, c->tlen, nt, c->fpart_cap / ((int64_t)c->tlen * vissf_blocks(G) * 256));
  if ((int64_t)c->tlen * s.nchunk * vissf_blocks(G) * 256 > c->fpart_cap)
    return c->fail(VW_ERR_RANGE, "internal: factorized partial buffer too small");
  const int64_t nfr = (int64_t)c->tlen * c->A;
  if (nt * c->nfr_pad > c->bits_cap) return c->fail(VW_ERR_RANGE, "internal: cone bit buffer too small");
  k_ant_bits<<<blocks_for(nt * c->nfr_pad), 256, 0, c->st>>>(c->d_cone, c->ntile, tile0, nt, c->d_ant8, nfr,
                                                              c->nfr_pad, c->d_bits_v);
  KC();
  VissFArgs a{};
  a.tab = c->d_tab;
  a.l = c->d_l; a.m = c->d_m; a.n = c->d_n; a.nm1 = c->d_nm1; a.B = c->d_B; a.Bf = c->d_Bf;
  a.npix = c->npix; a.tile0 = tile0; a.ntile = nt;
  a.aph = c->d_aph; a.ant8 = c->d_ant8; a.bits = c->d_bits_v; a.nfr_pad = c->nfr_pad;
  a.A = c->A; a.tlen = c->tlen;
  a.tpc = s.tpc; a.nchunk = s.nchunk; a.nquad = s.nquad;
  a.check = c->cfg.viss_form == VW_VISS_ANT_CHECK ? 1 : 0;
  a.part = c->d_fpart;
  ev_begin(c, ST_VISSK);
  const unsigned nb = (unsigned)s.nblocks;
#define VW_VF(TP_, G_, M_) k_vissf<TP_, G_, M_><<<nb, 256, 0, c->st>>>(a)
  if (c->TP == 256) {
    if (G == 1) mixed ? VW_VF(256, 1, true) : VW_VF(256, 1, false);
    else mixed ? VW_VF(256, 2, true) : VW_VF(256, 2, false);
  } else {
    if (G == 1) mixed ? VW_VF(512, 1, true) : VW_VF(512, 1, false);
    else mixed ? VW_VF(512, 2, true) : VW_VF(512, 2, false);
  }
#undef VW_VF
  KC();
  ev_end(c, ST_VISSK);
  const int64_t nent = vissf_blocks(G) * 256;
  k_vissf_fold<<<(unsigned)(c->tlen * (nent / 32)), 256, 0, c->st>>>(c->d_fpart, s.nchunk, nent, c->d_fesum);
  KC();
  k_vissf_rows<<<blocks_for(c->nh), 256, 0, c->st>>>(c->d_fesum, c->nh, c->P, G, c->d_pa, c->d_pb, tile0 > 0 ? 1 : 0,
                                                     c->d_vraw);
  KC();
  return VW_OK;
}

/* work order of the batched kernels (which = 0: the segment's half rows on
 * the Viss stream; 1: its groups on the grouping stream): a stable radix
 * sort of the antenna-direction pair keys */
/* stable LSD radix sort of n (key, value) pairs over `passes` 8-bit digits
 * from bit 0 with the memset-free kernels of vw_batch.h; the pairs move
 * between (k0, v0) and (k1, v1) and the sorted pairs end in (k1, v1) */
template <typename K, typename V>
int enq_rsort(Ctx* c, int which, hipStream_t st, K* k0, V* v0, K* k1, V* v1, int64_t n, int passes) {
  const int64_t nblk = (n + VW_RS_BLOCK - 1) / VW_RS_BLOCK;
  if (256 * nblk > c->rshist_cap) return c->fail(VW_ERR_RANGE, "internal: radix-sort table too small");
  int32_t* h = c->d_rshist[which];
  K* ka[2] = {k0, k1};
  V* va[2] = {v0, v1};
  int cur = 0;
  for (int p = 0; p < passes; ++p, cur ^= 1) {
    k_rs_hist<K><<<(unsigned)nblk, 256, 0, st>>>(ka[cur], n, 8 * p, h, nblk);
    KC();
    k_rs_scan<<<256, 256, 0, st>>>(h, nblk, h + c->rshist_cap);
    KC();
    k_rs_scatter<K, V><<<(unsigned)nblk, 256, 0, st>>>(ka[cur], va[cur], n, 8 * p, h, nblk, h + c->rshist_cap,
                                                       ka[cur ^ 1], va[cur ^ 1]);
    KC();
  }
  if (cur == 0) { /* an even number of passes ends in (k0, v0) */
    k_copy<K><<<blocks_for(n), 256, 0, st>>>(k0, k1, n);
    KC();
    k_copy<V><<<blocks_for(n), 256, 0, st>>>(v0, v1, n);
    KC();
  }
  return VW_OK;
}

int enq_order(Ctx* c, int which) {
  hipStream_t st = which == 0 ? c->st : c->stB;
  const int64_t n = which == 0 ? c->nh_pad : c->nq_pad;
  if (which == 0)
    k_row_keys<<<blocks_for(n), 256, 0, st>>>(c->d_rows, c->d_ant8, c->nh, n, c->d_okey[0], c->d_oidx[0]);
  else
    k_group_keys<<<blocks_for(n), 256, 0, st>>>(c->gs.gia, c->gs.gib, c->d_ant8, c->d_info, n, c->d_okey[1],
                                                c->d_oidx[1]);
  KC();
  /* keys below 2^13: two 8-bit digits */
  return enq_rsort<int32_t, int32_t>(c, which, st, c->d_okey[which], c->d_oidx[which], c->d_okey2[which],
                                     c->d_order[which], n, 2);
}

/* one tile pass of the batched Viss: classes, the product kernel, the fold */
int enq_viss_batch(Ctx* c, int64_t tile0, int64_t nt) {
  int rc;
  if ((rc = enq_classes(c, c->d_cone, c->ntile, tile0, nt, false, c->d_bits_v, c->d_cls_v, c->st))) return rc;
  const VissBShape s = vissb_shape(c, c->nh, nt, c->vb_target, c->bpart_cap / (c->nsky_pad * c->nh_pad));
  if (s.nchunk * c->nsky_pad * c->nh_pad > c->bpart_cap) return c->fail(VW_ERR_RANGE, "internal: batch partial buffer too small");
  VissBArgs a{};
  a.l = c->d_l; a.m = c->d_m; a.n = c->d_n; a.nm1 = c->d_nm1;
  a.skyh = (const uint4*)c->d_skyh; a.skyl = (const uint4*)c->d_skyl;
  a.npix = c->npix; a.nsky_pad = c->nsky_pad;
  a.tile0 = tile0; a.ntile = nt; a.TP = c->TP;
  a.rows = c->d_rows; a.ant8 = c->d_ant8; a.cls = c->d_cls_v; a.perm = c->d_order[0];
  a.nh = c->nh; a.nh_pad = c->nh_pad;
  a.tpc = s.tpc; a.nchunk = s.nchunk; a.nrb = s.nrb; a.nrbw = s.nrbw;
  a.part = c->d_bpart;
  ev_begin(c, ST_VISSK);
  const unsigned nb = (unsigned)s.nblocks;
  switch (c->nbw) {
    case 1: k_vissb<1><<<nb, 256, 0, c->st>>>(a); break;
    case 2: k_vissb<2><<<nb, 256, 0, c->st>>>(a); break;
    case 3: k_vissb<3><<<nb, 256, 0, c->st>>>(a); break;
    default: k_vissb<4><<<nb, 256, 0, c->st>>>(a); break;
  }
  KC();
  ev_end(c, ST_VISSK);
  k_vissb_fold<<<blocks_for(c->nsky * c->nh), 256, 0, c->st>>>(c->d_bpart, s.nchunk, c->nsky, c->nsky_pad, c->nh,
                                                               c->nh_pad, tile0 > 0 ? 1 : 0, c->d_vraw);
  KC();
  return VW_OK;
}

int enq_viss(Ctx* c, bool count) {
  if (!c->have_seg || !c->have_sky) return c->fail(VW_ERR_STATE, "vw_viss_half needs vw_set_sky and vw_gen_segment");
  if (count && c->batch) return c->fail(VW_ERR_ARG, "per-row counts come from the one-sky form (nsky = 1)");
  if (count && !c->d_cpart) {
    int rc2;
    if ((rc2 = dalloc(c, &c->d_cpart, (size_t)c->nchunk_max * c->nh_pad)) || (rc2 = dalloc(c, &c->d_counts, (size_t)c->max_nh)))
      return rc2;
  }
  const bool mixed = c->cfg.precision == VW_PREC_MIXED;
  int rc;
  ev_begin(c, ST_VISS);
  for (int64_t tile0 = 0; tile0 < c->ntile; tile0 += c->vtpp) {
    const int64_t nt = std::min(c->vtpp, c->ntile - tile0);
    if (c->batch) {
      if (tile0 == 0 && (rc = enq_order(c, 0))) return rc;
      if ((rc = enq_viss_batch(c, tile0, nt))) return rc;
      continue;
    }
    if (c->vfact && !count) {
      if ((rc = enq_viss_fact(c, tile0, nt))) return rc;
      continue;
    }
    VissShape s = viss_shape(c, c->nh, nt, count, c->nchunk_max);
    if (s.nchunk > c->nchunk_max) return c->fail(VW_ERR_RANGE, "internal: chunk buffer too small");
    if ((rc = enq_classes(c, c->d_cone, c->ntile, tile0, nt, false, c->d_bits_v, c->d_cls_v, c->st))) return rc;
    VissArgs a{};
    a.tab = c->d_tab;
    a.l = c->d_l; a.m = c->d_m; a.n = c->d_n; a.nm1 = c->d_nm1; a.B = c->d_B; a.Bf = c->d_Bf;
    a.npix = c->npix; a.tile0 = tile0; a.ntile = nt;
    a.rows = c->d_rows; a.rowt = c->d_rowt; a.ant8 = c->d_ant8; a.cls = (const uint32_t*)c->d_cls_v;
    a.nh = c->nh; a.nh_pad = c->nh_pad;
    a.tpc = s.tpc; a.nchunk = s.nchunk; a.nrb = s.nrb; a.bpc = s.bpc;
    a.part = c->d_part; a.cpart = c->d_cpart;
    ev_begin(c, ST_VISSK);
    unsigned nb = (unsigned)s.nblocks;
#define VW_VISS(TP_, M_, C_) launch_viss<TP_, viss_rpw(TP_, M_, C_), M_, C_>(a, nb, c->st)
    if (c->TP == 256) {
      if (mixed) count ? VW_VISS(256, true, true) : VW_VISS(256, true, false);
      else count ? VW_VISS(256, false, true) : VW_VISS(256, false, false);
    } else {
      if (mixed) count ? VW_VISS(512, true, true) : VW_VISS(512, true, false);
      else count ? VW_VISS(512, false, true) : VW_VISS(512, false, false);
    }
#undef VW_VISS
    KC();
    ev_end(c, ST_VISSK);
    k_viss_reduce<<<blocks_for(c->nh), 256, 0, c->st>>>(c->d_part, s.nchunk, c->nh, c->nh_pad, tile0 > 0 ? 1 : 0,
                                                        c->d_vraw, count ? c->d_cpart : nullptr,
                                                        count ? c->d_counts : nullptr);
    KC();
  }
  ev_end(c, ST_VISS);
  c->have_viss = true;
  c->have_conj = c->have_groups = false;
  return VW_OK;
}

int enq_conj(Ctx* c) {
  if (!c->have_viss) return c->fail(VW_ERR_STATE, "vw_conj_fill needs vw_viss_half");
  ev_begin(c, ST_CONJ);
  if (c->batch)
    k_conj_fill_b<<<blocks_for(c->nsky * c->nh), 256, 0, c->st>>>(c->d_vraw, c->d_rows, c->nh, c->P, c->nsky, c->d_vh,
                                                                  c->d_vfull);
  else
    k_conj_fill<<<blocks_for(c->nh), 256, 0, c->st>>>(c->d_vraw, c->d_rows, c->nh, c->P, c->d_vh, c->d_vfull);
  KC();
  ev_end(c, ST_CONJ);
  HC(hipEventRecord(c->ev_conj, c->st));
  c->have_conj = true;
  c->have_groups = false;
  return VW_OK;
}

/* w-plane fit, keys, stable order and runs of the current segment's rows
 * (grouping stream); the groups depend on the geometry only */
int enq_group_keys(Ctx* c) {
  bool full = c->cfg.recon_mode == VW_RECON_FULL;
  int64_t nr = full ? 2 * c->nh : c->nh;
  k_segparams<<<1, 256, 0, c->stB>>>(c->d_rows, c->nh, c->P, c->bl_max, c->lambda, c->d_info);
  KC();
  k_keys<<<blocks_for(nr), 256, 0, c->stB>>>(c->d_rows, nr, c->P, full ? 1 : 0, c->d_info, c->d_keys);
  KC();
  if (nr > VW_RANK_MAX) {
    /* long segments: stable LSD radix sort of the (key, row index) pairs
     * over the eight bytes of the non-negative int64 keys */
    k_iota<<<blocks_for(nr), 256, 0, c->stB>>>(nr, c->d_rank);
    KC();
    int rc = enq_rsort<int64_t, int32_t>(c, 1, c->stB, c->d_keys, c->d_rank, c->d_keys_sorted, c->d_idx_sorted, nr, 8);
    if (rc) return rc;
  } else {
    /* stable rank by (key, row): a grid of (256 rows) x (jblk rows) compare
     * tiles sized for about 2,048 workgroups, then a scatter */
    k_fill0<<<blocks_for(nr), 256, 0, c->stB>>>((uint32_t*)c->d_rank, nr);
    KC();
    const int64_t gx = blocks_for(nr);
    const int64_t gy = std::max<int64_t>(1, std::min<int64_t>(gx, (2048 + gx - 1) / gx));
    const int64_t jblk = ((nr + gy - 1) / gy + 255) / 256 * 256;
    dim3 grid((unsigned)gx, (unsigned)((nr + jblk - 1) / jblk));
    k_rank_count<<<grid, 256, 0, c->stB>>>(c->d_keys, nr, jblk, c->d_rank);
    KC();
    k_rank_scatter<<<blocks_for(nr), 256, 0, c->stB>>>(c->d_keys, c->d_rank, nr, c->d_idx_sorted, c->d_keys_sorted);
    KC();
  }
  k_groups<<<1, 1024, 0, c->stB>>>(c->d_keys_sorted, nr, c->cfg.blockage, c->d_gkeys, c->d_goff, c->d_info);
  KC();
  return VW_OK;
}

/* group records (cell centre, representative or mean visibility from vh,
 * antenna frames) of the current groups */
int enq_group_data(Ctx* c, const double2* vh) {
  bool full = c->cfg.recon_mode == VW_RECON_FULL;
  int64_t nr = full ? 2 * c->nh : c->nh;
  k_group_data<<<blocks_for(nr), 256, 0, c->stB>>>(c->d_gkeys, c->d_goff, c->d_idx_sorted, vh, c->d_rows,
                                                  c->d_ant8, c->P, full ? 1 : 0, c->cfg.blockage, c->cfg.occ_mode,
                                                  c->cfg.precision == VW_PREC_MIXED ? 1.0 : VW_TAB_SCALE, c->d_info,
                                                  c->gs);
  KC();
  return VW_OK;
}

int enq_group(Ctx* c) {
  if (!c->have_conj) return c->fail(VW_ERR_STATE, "vw_recon_segment needs vw_conj_fill");
  bool full = c->cfg.recon_mode == VW_RECON_FULL;
  HC(hipStreamWaitEvent(c->stB, c->ev_conj, 0));
  ev_begin(c, ST_GROUP, c->stB);
  int rcg;
  if ((rcg = enq_group_keys(c)) || (rcg = enq_group_data(c, c->d_vh))) return rcg;
  if (c->batch) {
    int rc;
    if ((rc = enq_order(c, 1))) return rc;
    k_group_order_recs<<<blocks_for(c->nq_pad), 256, 0, c->stB>>>(c->d_order[1], c->gs.rec, c->gs.ant, c->gs.gia,
                                                                  c->gs.gib, c->nq_pad, c->d_orec, c->d_oant,
                                                                  c->d_oia, c->d_oib);
    KC();
    k_group_pack<<<blocks_for(c->nq_pad * c->nsky_pad), 256, 0, c->stB>>>(
        c->d_goff, c->d_idx_sorted, c->d_vh, c->nh, c->P, full ? 1 : 0, c->cfg.blockage, c->cfg.occ_mode, c->d_info,
        c->d_order[1], c->nq_pad, c->nsky, c->nsky_pad, c->d_gpk[0], c->d_gpk[1], c->d_gpk[2], c->d_gpk[3]);
    KC();
  }
  ev_end(c, ST_GROUP, c->stB);
  c->have_groups = true;
  return VW_OK;
}

template <bool MIXED, int KIND>
void launch_recon(const ReconArgs& a, hipStream_t st) {
  int64_t waves = a.pwaves * a.nchunk;
  k_recon<RECON_PPL, MIXED, KIND><<<blocks_for(waves, 4), 256, 0, st>>>(a);
}

/* batched recon of one output set: per tile pass the (tile, group) classes,
 * the product kernel and the fold into the per-sky accumulators */
int enq_recon_set_batch(Ctx* c, bool img) {
  hipStream_t st = img ? c->st2 : c->stB;
  uint8_t* cls = img ? c->d_cls2 : c->d_cls;
  uint8_t* bits = img ? c->d_bits2 : c->d_bits;
  const bool rep = c->cfg.blockage != 0;
  const int64_t npts_all = img ? c->ngp : c->npix;
  const int64_t ntile_all = img ? c->ngt : c->ntile;
  const double* cone = img ? c->d_gcone : c->d_cone;
  const int64_t otpp = recon_pass_tiles(c, ntile_all, c->nq_pad);
  for (int64_t tile0 = 0; tile0 < ntile_all; tile0 += otpp) {
    const int64_t nt = std::min(otpp, ntile_all - tile0);
    const int64_t pt0 = tile0 * c->TP;
    const int64_t np = std::min(nt * c->TP, npts_all - pt0);
    if (rep) {
      int rc = enq_classes(c, cone, ntile_all, tile0, nt, true, bits, cls, st, true);
      if (rc) return rc;
    }
    const ReconBShape s = reconb_shape(c, np, c->rb_target);
    if (s.nchunk * c->nsky_pad * np > c->brpart_cap || s.nchunk * np > c->bwpart_cap)
      return c->fail(VW_ERR_RANGE, "internal: batch recon partial buffer too small");
    ReconBArgs a{};
    a.l = img ? c->d_gl : c->d_l;
    a.m = img ? c->d_gm : c->d_m;
    a.n = img ? c->d_gn : c->d_n;
    a.pt0 = pt0; a.npts = np; a.TP = c->TP;
    a.gcls = cls; a.nq_pad = c->nq_pad; a.gperm = c->d_order[1]; a.orec = c->d_orec; a.oant = c->d_oant;
    a.rec = c->gs.rec; a.gant = c->gs.ant; a.info = c->d_info;
    a.vrh = (const uint4*)c->d_gpk[0]; a.vrl = (const uint4*)c->d_gpk[1];
    a.vih = (const uint4*)c->d_gpk[2]; a.vil = (const uint4*)c->d_gpk[3];
    a.nsky_pad = c->nsky_pad;
    a.rep = rep ? 1 : 0;
    a.nchunk = s.nchunk; a.pblk = s.pblk; a.pbw = s.pbw;
    a.part = img ? c->d_brpart2 : c->d_brpart;
    a.wpart = img ? c->d_bwpart2 : c->d_bwpart;
    a.pstride = np;
    ev_begin(c, ST_RECONK, st);
    const unsigned nb = (unsigned)s.nblocks;
    switch (c->nbw) {
      case 1: k_reconb<1><<<nb, 256, 0, st>>>(a); break;
      case 2: k_reconb<2><<<nb, 256, 0, st>>>(a); break;
      case 3: k_reconb<3><<<nb, 256, 0, st>>>(a); break;
      default: k_reconb<4><<<nb, 256, 0, st>>>(a); break;
    }
    KC();
    ev_end(c, ST_RECONK, st);
    k_reconb_fold<<<blocks_for(c->nsky * np), 256, 0, st>>>(a.part, a.wpart, s.nchunk, c->nsky, c->nsky_pad, np, pt0,
                                                             np, npts_all, c->cfg.blockage, img ? c->d_Gacc : c->d_Cacc,
                                                             img ? c->d_GWacc : c->d_Wacc);
    KC();
  }
  return VW_OK;
}

/* one-sky recon of one output set into the accumulators acc / wacc (per
 * tile pass: classes, the recon kernel, the ordered fold of its chunks) */
int enq_recon_set_one(Ctx* c, bool img, double* acc, uint32_t* wacc) {
  /* each output set has its own stream and scratch */
  hipStream_t st = img ? c->st2 : c->stB;
  uint8_t* cls = img ? c->d_cls2 : c->d_cls;
  uint8_t* bits = img ? c->d_bits2 : c->d_bits;
  int kind;
  if (!c->cfg.blockage) kind = VW_KIND_AVG;
  else if (!c->generic) kind = VW_KIND_REP;
  else kind = VW_KIND_GENERIC;
  const bool mixed = c->cfg.precision == VW_PREC_MIXED;
  const int64_t npts_all = img ? c->ngp : c->npix;
  const int64_t ntile_all = img ? c->ngt : c->ntile;
  const double* cone = img ? c->d_gcone : c->d_cone;
  const int64_t kind_w = kind == VW_KIND_GENERIC ? c->nh_pad : c->nq_pad;
  const int64_t otpp = recon_pass_tiles(c, ntile_all, kind_w);
  for (int64_t tile0 = 0; tile0 < ntile_all; tile0 += otpp) {
    const int64_t nt = std::min(otpp, ntile_all - tile0);
    const int64_t pt0 = tile0 * c->TP;
    const int64_t np = std::min(nt * c->TP, npts_all - pt0);
    if (kind != VW_KIND_AVG) {
      int rc = enq_classes(c, cone, ntile_all, tile0, nt, kind == VW_KIND_REP, bits, cls, st);
      if (rc) return rc;
    }
    ReconArgs a{};
    a.tab = c->d_tab;
    a.l = img ? c->d_gl : c->d_l;
    a.m = img ? c->d_gm : c->d_m;
    a.n = img ? c->d_gn : c->d_n;
    a.pt0 = pt0;
    a.npts = np;
    a.TP = c->TP;
    a.gcls = (const uint32_t*)cls;
    a.nq_pad = c->nq_pad;
    a.rcls = (const uint32_t*)cls;
    a.nh_pad = c->nh_pad;
    a.g = c->gs;
    a.goff = c->d_goff;
    a.idx_sorted = c->d_idx_sorted;
    a.vh = c->d_vh;
    a.rows = c->d_rows;
    a.ant8 = c->d_ant8;
    a.info = c->d_info;
    a.P = c->P;
    a.full = c->cfg.recon_mode == VW_RECON_FULL;
    a.occ_mode = c->cfg.occ_mode;
    /* representatives are filtered for finite values except in the uv-mirror
     * recon with occ_mode 2 */
    const bool rep_finite = !(c->cfg.recon_mode == VW_RECON_MIRROR && c->cfg.occ_mode == 2);
    ReconShape rs = recon_shape(np);
    if (rs.nchunk * np > c->rpart_cap) return c->fail(VW_ERR_RANGE, "internal: recon partial buffer too small");
    a.pwaves = rs.pwaves;
    a.nchunk = rs.nchunk;
    a.pstride = np;
    a.part = img ? c->d_rpart2 : c->d_rpart;
    a.wpart = img ? c->d_rwpart2 : c->d_rwpart;
    ev_begin(c, ST_RECONK, st);
    if (mixed) {
      if (kind == VW_KIND_REP && rep_finite) launch_recon<true, VW_KIND_REP_FINITE>(a, st);
      else if (kind == VW_KIND_REP) launch_recon<true, VW_KIND_REP>(a, st);
      else if (kind == VW_KIND_GENERIC) launch_recon<true, VW_KIND_GENERIC>(a, st);
      else launch_recon<true, VW_KIND_AVG>(a, st);
    } else {
      if (kind == VW_KIND_REP && rep_finite) launch_recon<false, VW_KIND_REP_FINITE>(a, st);
      else if (kind == VW_KIND_REP) launch_recon<false, VW_KIND_REP>(a, st);
      else if (kind == VW_KIND_GENERIC) launch_recon<false, VW_KIND_GENERIC>(a, st);
      else launch_recon<false, VW_KIND_AVG>(a, st);
    }
    KC();
    ev_end(c, ST_RECONK, st);
    k_recon_reduce<<<blocks_for(np), 256, 0, st>>>(a.part, a.wpart, a.nchunk, a.pstride, pt0, np, c->cfg.blockage,
                                                      acc, wacc);
    KC();
  }
  return VW_OK;
}

int enq_recon_set(Ctx* c, bool img) {
  if (c->batch) return enq_recon_set_batch(c, img);
  return enq_recon_set_one(c, img, img ? c->d_Gacc : c->d_Cacc, img ? c->d_GWacc : c->d_Wacc);
}

int enq_recon(Ctx* c) {
  int rc;
  if ((rc = enq_group(c))) return rc;
  ev_begin(c, ST_RECON, c->stB);
  if (c->ngp) {
    /* the image set runs on its own stream beside the HEALPix set; both
     * read the groups and write disjoint accumulators */
    HC(hipEventRecord(c->ev_fork, c->stB));
    HC(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
    if ((rc = enq_recon_set(c, true))) return rc;
    HC(hipEventRecord(c->ev_join, c->st2));
  }
  if (c->cfg.healpix_out && (rc = enq_recon_set(c, false))) return rc;
  if (c->ngp) HC(hipStreamWaitEvent(c->stB, c->ev_join, 0));
  ev_end(c, ST_RECON, c->stB);
  /* the segment's buffer set is free for the segment after next */
  HC(hipEventRecord(c->ev_recon[c->cur], c->stB));
  return VW_OK;
}

/* one step over the segment schedule (enqueue only): every segment's
 * geometry, Viss, conjugate fill and recon; with per-sky schedules the
 * Viss of the base segments for every sky into the day buffer, then per
 * distinct (t0, tlen) of the skies' schedules the groups once and, for each
 * sky holding that segment, its group records and recon into its own
 * accumulators and weights */
int enq_step(Ctx* c) {
  int rc;
  if (!c->sky_sched) {
    for (int s = 0; s < c->segs; ++s)
      if ((rc = enq_gen(c, c->seg_t0[s], c->seg_tlen[s])) || (rc = enq_viss(c, false)) || (rc = enq_conj(c)) ||
          (rc = enq_recon(c)))
        return rc;
    return VW_OK;
  }
  const int64_t day_rows = (int64_t)c->T * c->P;
  for (int s = 0; s < c->segs; ++s) {
    if ((rc = enq_gen(c, c->seg_t0[s], c->seg_tlen[s])) || (rc = enq_viss(c, false)) || (rc = enq_conj(c))) return rc;
    k_vday_store<<<blocks_for(c->nsky * c->nh), 256, 0, c->st>>>(c->d_vh, c->nsky, c->nh,
                                                                  (int64_t)c->seg_t0[s] * c->P, day_rows, c->d_vday);
    KC();
  }
  for (const Ctx::SkySeg& u : c->sky_segs) {
    if ((rc = enq_gen(c, u.t0, u.tlen))) return rc;
    /* the segment's rows and the whole day buffer are written on st */
    HC(hipEventRecord(c->ev_conj, c->st));
    HC(hipStreamWaitEvent(c->stB, c->ev_conj, 0));
    ev_begin(c, ST_GROUP, c->stB);
    if ((rc = enq_group_keys(c))) return rc;
    ev_end(c, ST_GROUP, c->stB);
    for (const int32_t sk : u.skies) {
      if ((rc = enq_group_data(c, c->d_vday + sk * day_rows + (int64_t)u.t0 * c->P))) return rc;
      ev_begin(c, ST_RECON, c->stB);
      if (c->ngp) {
        HC(hipEventRecord(c->ev_fork, c->stB));
        HC(hipStreamWaitEvent(c->st2, c->ev_fork, 0));
        if ((rc = enq_recon_set_one(c, true, c->d_Gacc + sk * c->ngp, c->d_GWsky + sk * c->ngp))) return rc;
        HC(hipEventRecord(c->ev_join, c->st2));
      }
      if (c->cfg.healpix_out &&
          (rc = enq_recon_set_one(c, false, c->d_Cacc + sk * c->npix, c->d_Wsky + sk * c->npix)))
        return rc;
      if (c->ngp) HC(hipStreamWaitEvent(c->stB, c->ev_join, 0));
      ev_end(c, ST_RECON, c->stB);
    }
    /* this buffer set is free once the segment's last recon is done */
    HC(hipEventRecord(c->ev_recon[c->cur], c->stB));
  }
  c->have_seg = c->have_viss = c->have_conj = c->have_groups = false;
  return VW_OK;
}

int read_seg_info(Ctx* c, vw_seg_info* o) {
  DevInfo di;
  HC(hipMemcpy(&di, c->d_info, sizeof di, hipMemcpyDeviceToHost));
  if (!o) return VW_OK;
  o->t0 = c->t0;
  o->tlen = c->tlen;
  o->n_half = c->nh;
  o->fa = di.sp.fa;
  o->fb = di.sp.fb;
  o->denom = di.sp.denom;
  o->res = di.sp.res;
  o->half = di.sp.half;
  o->du = di.sp.du;
  o->nuniq = di.nuniq;
  o->nkeys = di.nuniq - di.first_nz;
  return VW_OK;
}

Ctx* as_ctx(void* p) { return (Ctx*)p; }

}  // namespace

/* ================================================================== */
/* C-ABI                                                              */
/* ================================================================== */

VW_API const char* vw_version(void) { return "vissws-gfx950 0.1"; }

VW_API int vw_config_init(vw_config* cfg, int32_t geometry) {
  if (!cfg) return VW_ERR_ARG;
  std::memset(cfg, 0, sizeof *cfg);
  cfg->geometry = geometry;
  cfg->orbit_seed = 42;
  cfg->day = 1;
  cfg->blockage = 1;
  cfg->occ_mode = 2;
  cfg->recon_mode = VW_RECON_MIRROR;
  cfg->tile_pix = 256;
  cfg->precision = VW_PREC_PLAIN;
  cfg->healpix_out = 1;
  cfg->nsky = 1;
  cfg->gm_km3_s2 = 398600.4418;
  if (geometry == VW_GEOM_LUNAR8) {
    cfg->nside = 4096;
    cfg->freq_hz = 1e7;
    cfg->n_ant = 8;
    cfg->segs = 10;
    cfg->occ_radius_km = VW_LUNAR_R_KM;
    cfg->occ_altitude_km = VW_LUNAR_H_KM;
    cfg->r_min_km = cfg->r_max_km = VW_LUNAR_A_KM;
  } else if (geometry == VW_GEOM_CIRCULAR) {
    cfg->nside = 32;
    cfg->freq_hz = 1e7;
    cfg->n_ant = 8;
    cfg->n_epochs = 64;
    cfg->dt_s = 10.0;
    cfg->segs = 4;
    cfg->occ_radius_km = 6371.0;
    cfg->occ_altitude_km = -1.0;
    cfg->r_min_km = 6900.0;
    cfg->r_max_km = 7100.0;
    cfg->img_n = 256;
  } else {
    return VW_ERR_ARG;
  }
  return VW_OK;
}

VW_API int vw_create(const vw_config* cfg, void** out) {
  if (!out || !cfg) {
    g_create_err = "null argument";
    return VW_ERR_ARG;
  }
  *out = nullptr;
  Ctx* c = new Ctx();
  int rc = make_plan(c, cfg);
  if (!rc) rc = create_device(c);
  if (rc) {
    g_create_err = c->err;
    vw_destroy(c);
    return rc;
  }
  *out = c;
  return VW_OK;
}

VW_API void vw_destroy(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  for (void* q : c->allocs) (void)hipFree(q);
  for (int s = 0; s < ST_N; ++s)
    for (auto& e : c->ev[s]) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_conj) (void)hipEventDestroy(c->ev_conj);
  for (int b = 0; b < 2; ++b)
    if (c->ev_recon[b]) (void)hipEventDestroy(c->ev_recon[b]);
  if (c->st2 && c->st2 != c->st) (void)hipStreamDestroy(c->st2);
  if (c->stB && c->stB != c->st) (void)hipStreamDestroy(c->stB);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

VW_API const char* vw_last_error(void* p) {
  if (!p) return g_create_err.c_str();
  return as_ctx(p)->err.c_str();
}

VW_API int vw_get_plan(void* p, vw_plan* o) {
  Ctx* c = as_ctx(p);
  if (!c || !o) return VW_ERR_ARG;
  std::memset(o, 0, sizeof *o);
  o->npix = c->npix;
  o->ntile = c->ntile;
  o->n_img_pts = c->ngp;
  o->n_img_tiles = c->ngt;
  o->n_pairs = c->P;
  o->n_epochs = c->T;
  o->segs = c->segs;
  o->max_tlen = c->max_tlen;
  o->orbit_res = c->orbit_res;
  o->seg_len = c->seg_len;
  o->lambda_m = c->lambda;
  o->bl_max_m = c->bl_max;
  o->cosphi = c->cosphi_glob;
  o->pix_area = c->pix_area;
  o->device_bytes = c->dev_bytes;
  return VW_OK;
}

VW_API int vw_get_schedule(void* p, int32_t* t0, int32_t* tl) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (t0) std::memcpy(t0, c->seg_t0.data(), sizeof(int32_t) * c->segs);
  if (tl) std::memcpy(tl, c->seg_tlen.data(), sizeof(int32_t) * c->segs);
  return VW_OK;
}

VW_API int vw_set_day(void* p, int32_t day) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  c->cfg.day = day;
  drop_graph(c); /* bl_max and the day's scalars are captured by value */
  int rc = upload_day(c);
  if (rc) return rc;
  c->have_seg = c->have_viss = c->have_conj = c->have_groups = false;
  return sync(c);
}

VW_API int vw_set_sky(void* p, const double* B, int64_t npix, int32_t nsky) {
  Ctx* c = as_ctx(p);
  if (!c || !B) return VW_ERR_ARG;
  if (npix != c->npix) return c->fail(VW_ERR_ARG, "sky has " + std::to_string(npix) + " pixels, expected " + std::to_string(c->npix));
  if (nsky != c->nsky) return c->fail(VW_ERR_ARG, "nsky must equal the configured nsky (" + std::to_string(c->nsky) + ")");
  if (c->batch) {
    /* FP32 of B * pix_area (main_recon_online.cu:106-110), then hi / lo BF16 */
    std::vector<float> f((size_t)(npix * nsky));
    for (int64_t i = 0; i < npix * nsky; ++i) {
      if (!std::isfinite(B[i])) return c->fail(VW_ERR_ARG, "batched skies must be finite");
      f[i] = (float)(B[i] * c->pix_area);
    }
    HC(hipMemcpyAsync(c->d_skyf, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, c->st));
    const int64_t npix_pad = c->ntile * c->TP;
    k_sky_pack<<<blocks_for(npix_pad * c->nsky_pad), 256, 0, c->st>>>(c->d_skyf, c->nsky, npix, npix_pad, c->nsky_pad,
                                                                       c->d_skyh, c->d_skyl);
    KC();
    HC(hipStreamSynchronize(c->st));
    c->have_sky = true;
    return VW_OK;
  }
  std::vector<double> b((size_t)npix);
  std::vector<float> bf((size_t)npix);
  for (int64_t i = 0; i < npix; ++i) {
    b[i] = B[i] * c->pix_area; /* main_recon_online.cu:106-110 in FP64 */
    bf[i] = (float)b[i];
  }
  HC(hipMemcpyAsync(c->d_B, b.data(), sizeof(double) * npix, hipMemcpyHostToDevice, c->st));
  HC(hipMemcpyAsync(c->d_Bf, bf.data(), sizeof(float) * npix, hipMemcpyHostToDevice, c->st));
  HC(hipStreamSynchronize(c->st));
  c->have_sky = true;
  return VW_OK;
}

VW_API int vw_gen_segment(void* p, int32_t t0, int32_t tlen, vw_geom_host* o) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  int rc = enq_gen(c, t0, tlen);
  if (rc) return rc;
  if (o) {
    if (!c->d_full && (rc = dalloc(c, &c->d_full, 9 * 2 * (size_t)c->max_nh))) return rc;
    int64_t nf = 2 * c->nh;
    k_seg_full<<<blocks_for(nf), 256, 0, c->st>>>(c->d_pos, c->T, c->A, c->P, t0, c->nh, c->d_pa, c->d_pb, c->d_rows,
                                                  c->d_full);
    KC();
    double* dst[9] = {o->u, o->v, o->w, o->xa, o->ya, o->za, o->xb, o->yb, o->zb};
    for (int k = 0; k < 9; ++k)
      if (dst[k]) HC(hipMemcpyAsync(dst[k], c->d_full + k * nf, sizeof(double) * nf, hipMemcpyDeviceToHost, c->st));
    HC(hipStreamSynchronize(c->st));
    /* half-symmetry guard (orbit_online.cuh:403-418, common.hpp:222-226) */
    if (o->u && o->v && o->w) {
      for (int64_t g = 0; g < c->tlen; ++g)
        for (int j = 0; j < c->P; ++j) {
          int64_t i0 = g * 2 * c->P + j, i1 = i0 + c->P;
          const double* arr[3] = {o->u, o->v, o->w};
          for (int k = 0; k < 3; ++k) {
            double a = arr[k][i0], b = arr[k][i1];
            double tol = 1e-5 + 1e-5 * std::fmax(std::fabs(a), std::fabs(b));
            if (!(std::fabs(a + b) <= tol)) return c->fail(VW_ERR_HALFSYM, "generated segment uvw violates half symmetry");
          }
        }
    }
  }
  return sync(c);
}

VW_API int vw_viss_half(void* p, vw_c128* out) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  int rc = enq_viss(c, false);
  if (rc) return rc;
  if (out) HC(hipMemcpyAsync(out, c->d_vraw, sizeof(vw_c128) * c->nh * c->nsky, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

VW_API int vw_conj_fill(void* p, vw_c128* oh, vw_c128* of) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  int rc = enq_conj(c);
  if (rc) return rc;
  if (oh) HC(hipMemcpyAsync(oh, c->d_vh, sizeof(vw_c128) * c->nh * c->nsky, hipMemcpyDeviceToHost, c->st));
  if (of) HC(hipMemcpyAsync(of, c->d_vfull, sizeof(vw_c128) * 2 * c->nh * c->nsky, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

VW_API int vw_recon_segment(void* p, vw_seg_info* info) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (c->sky_sched) return c->fail(VW_ERR_STATE, "with per-sky schedules the recon runs in vw_run_all");
  int rc = enq_recon(c);
  if (rc) return rc;
  if ((rc = sync(c))) return rc;
  return read_seg_info(c, info);
}

VW_API int vw_run_all(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (c->sky_sched && !c->have_sky) return c->fail(VW_ERR_STATE, "vw_run_all needs vw_set_sky");
  int rc;
  if ((rc = enq_step(c)) || (rc = join_recon(c))) return rc;
  return sync(c);
}

namespace {
/* accumulators and the running key total cleared by kernels (no memset nodes) */
int enq_reset_fill(Ctx* c) {
  const int64_t nC = 2 * c->npix * c->nsky; /* doubles as 32-bit words */
  k_fill0<<<blocks_for(nC), 256, 0, c->st>>>((uint32_t*)c->d_Cacc, nC);
  KC();
  k_fill0<<<blocks_for(c->npix), 256, 0, c->st>>>(c->d_Wacc, c->npix);
  KC();
  if (c->ngp) {
    const int64_t nG = 2 * c->ngp * c->nsky;
    k_fill0<<<blocks_for(nG), 256, 0, c->st>>>((uint32_t*)c->d_Gacc, nG);
    KC();
    k_fill0<<<blocks_for(c->ngp), 256, 0, c->st>>>(c->d_GWacc, c->ngp);
    KC();
  }
  k_fill0<<<1, 256, 0, c->st>>>((uint32_t*)c->d_info, (int64_t)(sizeof(DevInfo) / 4));
  KC();
  if (c->sky_sched) {
    k_fill0<<<blocks_for(c->npix * c->nsky), 256, 0, c->st>>>(c->d_Wsky, c->npix * c->nsky);
    KC();
    if (c->ngp) {
      k_fill0<<<blocks_for(c->ngp * c->nsky), 256, 0, c->st>>>(c->d_GWsky, c->ngp * c->nsky);
      KC();
    }
  }
  return VW_OK;
}
}  // namespace

/* One batched step from a graph: clear the accumulators, then every segment of
 * the schedule (as vw_reset_accumulators + vw_run_all).  The graph is captured
 * on the first call and replayed afterwards; a change of launch targets or of
 * day captures it again. */
VW_API int vw_run_all_graph(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->batch) return c->fail(VW_ERR_ARG, "graph steps are built for batched contexts (nsky > 1)");
  if (!c->have_sky) return c->fail(VW_ERR_STATE, "vw_run_all_graph needs vw_set_sky");
  int rc;
  if (!c->graph_exec) {
    if ((rc = join_recon(c)) || (rc = sync(c))) return rc;
    HC(hipStreamBeginCapture(c->st, hipStreamCaptureModeRelaxed));
    c->capturing = true;
    /* the waits on the recon events of the first two segments refer to records
     * inside the capture */
    for (int b = 0; b < 2; ++b) (void)hipEventRecord(c->ev_recon[b], c->st);
    rc = enq_reset_fill(c);
    if (!rc) rc = enq_step(c);
    if (!rc) rc = join_recon(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &g);
    c->capturing = false;
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return rc ? rc : c->fail(VW_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      c->graph_exec = nullptr;
      return c->fail(VW_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
    }
  }
  HC(hipGraphLaunch(c->graph_exec, c->st));
  return sync(c);
}

VW_API int vw_finalize(void* p, double* C, uint32_t* W, double* img, uint32_t* Wimg) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (int rj = join_recon(c)) return rj;
  ev_begin(c, ST_FIN);
  /* normalise on the device; copy out only what was asked for */
  if (c->cfg.healpix_out) {
    k_normalize<<<blocks_for(c->npix), 256, 0, c->st>>>(c->d_Cacc, c->d_Wacc, c->npix, c->cfg.blockage, c->d_info,
                                                        nullptr, c->d_out, c->d_wout);
    KC();
    if (c->batch) {
      k_normalize_b<<<blocks_for(c->npix * c->nsky), 256, 0, c->st>>>(
          c->d_Cacc, c->sky_sched ? c->d_Wsky : c->d_Wacc, c->npix, c->nsky, c->cfg.blockage, c->d_info, nullptr,
          c->npix, c->d_out, c->sky_sched ? c->npix : 0);
      KC();
    }
  }
  if (c->ngp) {
    int64_t NN = (int64_t)c->cfg.img_n * c->cfg.img_n;
    HC(hipMemsetAsync(c->d_iout, 0, sizeof(double) * NN * c->nsky, c->st));
    HC(hipMemsetAsync(c->d_iwout, 0, sizeof(uint32_t) * NN, c->st));
    k_normalize<<<blocks_for(c->ngp), 256, 0, c->st>>>(c->d_Gacc, c->d_GWacc, c->ngp, c->cfg.blockage, c->d_info,
                                                       c->d_gidx, c->d_iout, c->d_iwout);
    KC();
    if (c->batch) {
      k_normalize_b<<<blocks_for(c->ngp * c->nsky), 256, 0, c->st>>>(
          c->d_Gacc, c->sky_sched ? c->d_GWsky : c->d_GWacc, c->ngp, c->nsky, c->cfg.blockage, c->d_info, c->d_gidx,
          NN, c->d_iout, c->sky_sched ? c->ngp : 0);
      KC();
    }
  }
  ev_end(c, ST_FIN);
  if (C) {
    if (c->cfg.healpix_out)
      HC(hipMemcpyAsync(C, c->d_out, sizeof(double) * c->npix * c->nsky, hipMemcpyDeviceToHost, c->st));
    else std::memset(C, 0, sizeof(double) * c->npix * c->nsky);
  }
  const int64_t nw = c->sky_sched ? c->nsky : 1; /* weight maps */
  if (W) {
    if (c->cfg.healpix_out)
      HC(hipMemcpyAsync(W, c->sky_sched ? c->d_Wsky : c->d_wout, sizeof(uint32_t) * c->npix * nw,
                        hipMemcpyDeviceToHost, c->st));
    else std::memset(W, 0, sizeof(uint32_t) * c->npix * nw);
  }
  if (c->ngp) {
    int64_t NN = (int64_t)c->cfg.img_n * c->cfg.img_n;
    if (img) HC(hipMemcpyAsync(img, c->d_iout, sizeof(double) * NN * c->nsky, hipMemcpyDeviceToHost, c->st));
    if (Wimg && c->sky_sched) {
      k_fill0<<<blocks_for(NN * c->nsky), 256, 0, c->st>>>(c->d_iwout_sky, NN * c->nsky);
      KC();
      k_scatter_w<<<blocks_for(c->ngp * c->nsky), 256, 0, c->st>>>(c->d_GWsky, c->ngp, c->nsky, c->d_gidx, NN,
                                                                    c->d_iwout_sky);
      KC();
      HC(hipMemcpyAsync(Wimg, c->d_iwout_sky, sizeof(uint32_t) * NN * c->nsky, hipMemcpyDeviceToHost, c->st));
    } else if (Wimg) {
      HC(hipMemcpyAsync(Wimg, c->d_iwout, sizeof(uint32_t) * NN, hipMemcpyDeviceToHost, c->st));
    }
  }
  return sync(c);
}

VW_API int vw_reset_accumulators(void* p) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (int rj = join_recon(c)) return rj;
  HC(hipMemsetAsync(c->d_Cacc, 0, sizeof(double) * c->npix * c->nsky, c->st));
  HC(hipMemsetAsync(c->d_Wacc, 0, sizeof(uint32_t) * c->npix, c->st));
  if (c->ngp) {
    HC(hipMemsetAsync(c->d_Gacc, 0, sizeof(double) * c->ngp * c->nsky, c->st));
    HC(hipMemsetAsync(c->d_GWacc, 0, sizeof(uint32_t) * c->ngp, c->st));
  }
  HC(hipMemsetAsync(c->d_info, 0, sizeof(DevInfo), c->st));
  if (c->sky_sched) {
    HC(hipMemsetAsync(c->d_Wsky, 0, sizeof(uint32_t) * c->npix * c->nsky, c->st));
    if (c->ngp) HC(hipMemsetAsync(c->d_GWsky, 0, sizeof(uint32_t) * c->ngp * c->nsky, c->st));
  }
  return sync(c);
}

VW_API int vw_get_accumulators(void* p, double* C, uint32_t* W, double* G, uint32_t* GW, int64_t* wtotal) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (c->sky_sched) return c->fail(VW_ERR_STATE, "accumulators of per-sky schedules are not exported");
  if (int rj = join_recon(c)) return rj;
  if (C) HC(hipMemcpyAsync(C, c->d_Cacc, sizeof(double) * c->npix * c->nsky, hipMemcpyDeviceToHost, c->st));
  if (W) HC(hipMemcpyAsync(W, c->d_Wacc, sizeof(uint32_t) * c->npix, hipMemcpyDeviceToHost, c->st));
  if (c->ngp && G) HC(hipMemcpyAsync(G, c->d_Gacc, sizeof(double) * c->ngp * c->nsky, hipMemcpyDeviceToHost, c->st));
  if (c->ngp && GW) HC(hipMemcpyAsync(GW, c->d_GWacc, sizeof(uint32_t) * c->ngp, hipMemcpyDeviceToHost, c->st));
  if (wtotal)
    HC(hipMemcpyAsync(wtotal, (char*)c->d_info + offsetof(DevInfo, wtotal), sizeof(int64_t), hipMemcpyDeviceToHost,
                      c->st));
  return sync(c);
}

VW_API int vw_set_accumulators(void* p, const double* C, const uint32_t* W, const double* G, const uint32_t* GW,
                               int64_t wtotal) {
  Ctx* c = as_ctx(p);
  if (!c || wtotal < 0) return VW_ERR_ARG;
  if (c->sky_sched) return c->fail(VW_ERR_STATE, "accumulators of per-sky schedules are not imported");
  if (int rj = join_recon(c)) return rj;
  if (C) HC(hipMemcpyAsync(c->d_Cacc, C, sizeof(double) * c->npix * c->nsky, hipMemcpyHostToDevice, c->st));
  else HC(hipMemsetAsync(c->d_Cacc, 0, sizeof(double) * c->npix * c->nsky, c->st));
  if (W) HC(hipMemcpyAsync(c->d_Wacc, W, sizeof(uint32_t) * c->npix, hipMemcpyHostToDevice, c->st));
  else HC(hipMemsetAsync(c->d_Wacc, 0, sizeof(uint32_t) * c->npix, c->st));
  if (c->ngp) {
    if (G) HC(hipMemcpyAsync(c->d_Gacc, G, sizeof(double) * c->ngp * c->nsky, hipMemcpyHostToDevice, c->st));
    else HC(hipMemsetAsync(c->d_Gacc, 0, sizeof(double) * c->ngp * c->nsky, c->st));
    if (GW) HC(hipMemcpyAsync(c->d_GWacc, GW, sizeof(uint32_t) * c->ngp, hipMemcpyHostToDevice, c->st));
    else HC(hipMemsetAsync(c->d_GWacc, 0, sizeof(uint32_t) * c->ngp, c->st));
  }
  HC(hipMemcpyAsync((char*)c->d_info + offsetof(DevInfo, wtotal), &wtotal, sizeof(int64_t), hipMemcpyHostToDevice,
                    c->st));
  return sync(c);
}

/* launch-shape targets of the batched kernels (waves per launch) */
VW_API int vw_set_batch_targets(void* p, int64_t viss_waves, int64_t recon_waves) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->batch) return c->fail(VW_ERR_ARG, "launch targets apply to batched contexts (nsky > 1)");
  auto ok = [](int64_t t) {
    for (int64_t v : VW_BATCH_TARGETS)
      if (v == t) return true;
    return false;
  };
  if (!ok(viss_waves) || !ok(recon_waves))
    return c->fail(VW_ERR_ARG, "target must be one of 2048, 4096, 8192, 16384, 32768, 65536");
  c->vb_target = viss_waves;
  c->rb_target = recon_waves;
  drop_graph(c); /* the captured launches carry the old shapes */
  return VW_OK;
}

VW_API int vw_get_batch_targets(void* p, int64_t* viss_waves, int64_t* recon_waves) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (viss_waves) *viss_waves = c->vb_target;
  if (recon_waves) *recon_waves = c->rb_target;
  return VW_OK;
}

/* Autotuner: on the schedule's longest segment, times the batched Viss
 * kernel and the recon stage for every launch target (reps runs each, HIP
 * events on the context's streams) and keeps the fastest.  The accumulators
 * are cleared afterwards. */
/* waves per adjoint launch that the row chunking aims at */
#define VW_ADJ_TARGET_WAVES 16384

VW_API int vw_viss_adjoint(void* p, const vw_c128* res, double* g) {
  Ctx* c = as_ctx(p);
  if (!c || !res || !g) return VW_ERR_ARG;
  if (c->batch) return c->fail(VW_ERR_ARG, "the adjoint runs on one-sky contexts");
  if (!c->have_seg) return c->fail(VW_ERR_STATE, "vw_viss_adjoint needs vw_gen_segment");
  int rc;
  if ((rc = join_recon(c)) || (rc = sync(c))) return rc;
  const int64_t nsub_max = c->vtpp * c->TP / 256;
  if (!c->d_adj_res) {
    c->adj_part_cap = (VW_ADJ_TARGET_WAVES + 2 * nsub_max) * 256;
    if ((rc = dalloc(c, &c->d_adj_res, (size_t)c->max_nh)) || (rc = dalloc(c, &c->d_adj_out, (size_t)c->npix)) ||
        (rc = dalloc(c, &c->d_adj_part, (size_t)c->adj_part_cap)))
      return rc;
  }
  HC(hipMemcpyAsync(c->d_adj_res, res, sizeof(vw_c128) * c->nh, hipMemcpyHostToDevice, c->st));
  for (int64_t tile0 = 0; tile0 < c->ntile; tile0 += c->vtpp) {
    const int64_t nt = std::min(c->vtpp, c->ntile - tile0);
    if ((rc = enq_classes(c, c->d_cone, c->ntile, tile0, nt, false, c->d_bits_v, c->d_cls_v, c->st))) return rc;
    VissAdjArgs a{};
    a.tab = c->d_tab;
    a.l = c->d_l; a.m = c->d_m; a.n = c->d_n; a.nm1 = c->d_nm1; a.npix = c->npix;
    a.tile0 = tile0; a.ntile = nt;
    a.rows = c->d_rows; a.rowt = c->d_rowt; a.ant8 = c->d_ant8; a.cls = (const uint32_t*)c->d_cls_v;
    a.nh = c->nh; a.nh_pad = c->nh_pad; a.res = c->d_adj_res;
    a.nsub = nt * c->TP / 256;
    const int64_t want = std::max<int64_t>(1, VW_ADJ_TARGET_WAVES / a.nsub);
    a.rpc = std::max<int64_t>(16, ((c->nh + want - 1) / want + 3) / 4 * 4);
    a.nchunk = (c->nh + a.rpc - 1) / a.rpc;
    if (a.nchunk * a.nsub * 256 > c->adj_part_cap) return c->fail(VW_ERR_RANGE, "internal: adjoint partials too small");
    a.part = c->d_adj_part;
    const unsigned nb = (unsigned)((a.nchunk * a.nsub + 3) / 4);
    if (c->TP == 256) k_viss_adj<256><<<nb, 256, 0, c->st>>>(a);
    else k_viss_adj<512><<<nb, 256, 0, c->st>>>(a);
    KC();
    const int64_t p0 = tile0 * c->TP, np = std::min(nt * c->TP, c->npix - p0);
    k_viss_adj_fold<<<blocks_for(np), 256, 0, c->st>>>(c->d_adj_part, a.nchunk, a.nsub * 256, p0, np, c->pix_area,
                                                       c->d_adj_out);
    KC();
  }
  HC(hipMemcpyAsync(g, c->d_adj_out, sizeof(double) * c->npix, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

VW_API int vw_set_sky_schedules(void* p, const int32_t* nseg, const int32_t* seg_tlen) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->batch) return c->fail(VW_ERR_ARG, "per-sky schedules are for batched contexts (nsky > 1)");
  int rc;
  if ((rc = join_recon(c)) || (rc = sync(c))) return rc;
  drop_graph(c);
  if (!nseg) {
    c->sky_sched = false;
    c->sky_segs.clear();
    return vw_reset_accumulators(c);
  }
  if (!seg_tlen) return VW_ERR_ARG;
  if (!c->cfg.blockage) return c->fail(VW_ERR_ARG, "per-sky schedules need blockage = 1 (per-sky weights)");
  /* validate every sky's schedule: lengths in [1, longest base segment],
   * covering the run's epochs exactly */
  std::vector<Ctx::SkySeg> segs;
  int64_t off = 0;
  for (int64_t sk = 0; sk < c->nsky; ++sk) {
    if (nseg[sk] < 1 || nseg[sk] > c->T)
      return c->fail(VW_ERR_ARG, "sky " + std::to_string(sk) + ": segment count out of range");
    int t0 = 0;
    for (int32_t k = 0; k < nseg[sk]; ++k) {
      const int32_t tl = seg_tlen[off + k];
      if (tl < 1 || tl > c->max_tlen)
        return c->fail(VW_ERR_RANGE, "sky " + std::to_string(sk) + ": segment of " + std::to_string(tl) +
                                         " epochs outside [1, " + std::to_string(c->max_tlen) +
                                         "] (the longest segment of the context's schedule)");
      if ((int64_t)t0 + tl > c->T) return c->fail(VW_ERR_ARG, "sky " + std::to_string(sk) + ": schedule beyond the run");
      bool found = false;
      for (Ctx::SkySeg& u : segs)
        if (u.t0 == t0 && u.tlen == tl) {
          u.skies.push_back((int32_t)sk);
          found = true;
          break;
        }
      if (!found) segs.push_back(Ctx::SkySeg{t0, tl, {(int32_t)sk}});
      t0 += tl;
    }
    if (t0 != c->T) return c->fail(VW_ERR_ARG, "sky " + std::to_string(sk) + ": schedule does not cover the run");
    off += nseg[sk];
  }
  std::sort(segs.begin(), segs.end(), [](const Ctx::SkySeg& a, const Ctx::SkySeg& b) {
    return a.t0 != b.t0 ? a.t0 < b.t0 : a.tlen < b.tlen;
  });
  if (!c->d_vday) {
    if ((rc = dalloc(c, &c->d_vday, (size_t)c->nsky * c->T * c->P)) ||
        (rc = dalloc(c, &c->d_Wsky, (size_t)c->nsky * c->npix)))
      return rc;
    if (c->ngp && ((rc = dalloc(c, &c->d_GWsky, (size_t)c->nsky * c->ngp)) ||
                   (rc = dalloc(c, &c->d_iwout_sky, (size_t)c->nsky * c->cfg.img_n * c->cfg.img_n))))
      return rc;
  }
  c->sky_segs = std::move(segs);
  c->sky_sched = true;
  return vw_reset_accumulators(c);
}

VW_API int vw_autotune(void* p, int32_t reps, int64_t* viss_waves, int64_t* recon_waves) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->batch) return c->fail(VW_ERR_ARG, "the autotuner tunes batched contexts (nsky > 1)");
  if (!c->have_sky) return c->fail(VW_ERR_STATE, "vw_autotune needs vw_set_sky");
  if (reps <= 0 || reps > 100) return c->fail(VW_ERR_ARG, "reps must be in [1, 100]");
  int rc;
  if ((rc = join_recon(c)) || (rc = sync(c))) return rc;
  int longest = 0;
  for (int k = 1; k < c->segs; ++k)
    if (c->seg_tlen[k] > c->seg_tlen[longest]) longest = k;
  if ((rc = enq_gen(c, c->seg_t0[longest], c->seg_tlen[longest]))) return rc;
  hipEvent_t e0, e1;
  HC(hipEventCreate(&e0));
  HC(hipEventCreate(&e1));
  auto timed = [&](hipStream_t st, auto&& body, float* ms) -> int {
    int r = body(); /* warm-up */
    if (r) return r;
    HC(hipEventRecord(e0, st));
    for (int i = 0; i < reps && !r; ++i) r = body();
    if (r) return r;
    HC(hipEventRecord(e1, st));
    HC(hipEventSynchronize(e1));
    HC(hipEventElapsedTime(ms, e0, e1));
    return VW_OK;
  };
  float best = 1e30f, ms = 0.f;
  int64_t bv = c->vb_target, br = c->rb_target;
  for (int64_t t : VW_BATCH_TARGETS) {
    c->vb_target = t;
    if ((rc = timed(c->st, [&]() { return enq_viss(c, false); }, &ms))) break;
    if (ms < best) { best = ms; bv = t; }
  }
  c->vb_target = bv;
  if (!rc) rc = enq_viss(c, false);
  if (!rc) rc = enq_conj(c);
  if (!rc) {
    best = 1e30f;
    for (int64_t t : VW_BATCH_TARGETS) {
      c->rb_target = t;
      /* the recon stage ends on stB after joining the image set's stream */
      HC(hipStreamWaitEvent(c->stB, c->ev_conj, 0));
      if ((rc = timed(c->stB, [&]() { return enq_recon(c); }, &ms))) break;
      if (ms < best) { best = ms; br = t; }
    }
    c->rb_target = br;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = vw_reset_accumulators(c))) return rc;
  c->have_seg = c->have_viss = c->have_conj = c->have_groups = false;
  c->tim = vw_timings{};
  if (viss_waves) *viss_waves = c->vb_target;
  if (recon_waves) *recon_waves = c->rb_target;
  drop_graph(c);
  return VW_OK;
}

VW_API int vw_get_timings(void* p, vw_timings* o) {
  Ctx* c = as_ctx(p);
  if (!c || !o) return VW_ERR_ARG;
  *o = c->tim;
  c->tim = vw_timings{};
  return VW_OK;
}

VW_API int vw_get_lmn(void* p, double* l, double* m, double* n, double* nm1) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  size_t b = sizeof(double) * c->npix;
  if (l) HC(hipMemcpy(l, c->d_l, b, hipMemcpyDeviceToHost));
  if (m) HC(hipMemcpy(m, c->d_m, b, hipMemcpyDeviceToHost));
  if (n) HC(hipMemcpy(n, c->d_n, b, hipMemcpyDeviceToHost));
  if (nm1) HC(hipMemcpy(nm1, c->d_nm1, b, hipMemcpyDeviceToHost));
  return VW_OK;
}

VW_API int vw_get_tiles(void* p, int32_t img, double* cone5) {
  Ctx* c = as_ctx(p);
  if (!c || !cone5) return VW_ERR_ARG;
  if (img && !c->ngp) return c->fail(VW_ERR_ARG, "no lm grid");
  HC(hipMemcpy(cone5, img ? c->d_gcone : c->d_cone, sizeof(double) * 5 * (img ? c->ngt : c->ntile), hipMemcpyDeviceToHost));
  return VW_OK;
}

VW_API int vw_get_img_points(void* p, int64_t* idx, double* l, double* m, double* n) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->ngp) return c->fail(VW_ERR_ARG, "no lm grid");
  size_t b = sizeof(double) * c->ngp;
  if (idx) std::memcpy(idx, c->gidx.data(), sizeof(int64_t) * c->ngp);
  if (l) HC(hipMemcpy(l, c->d_gl, b, hipMemcpyDeviceToHost));
  if (m) HC(hipMemcpy(m, c->d_gm, b, hipMemcpyDeviceToHost));
  if (n) HC(hipMemcpy(n, c->d_gn, b, hipMemcpyDeviceToHost));
  return VW_OK;
}

VW_API int vw_get_positions(void* p, double* pos) {
  Ctx* c = as_ctx(p);
  if (!c || !pos) return VW_ERR_ARG;
  HC(hipMemcpy(pos, c->d_pos, sizeof(double) * 3 * c->T * c->A, hipMemcpyDeviceToHost));
  return VW_OK;
}

VW_API int vw_get_rows(void* p, double* u, double* v, double* w) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->have_seg) return c->fail(VW_ERR_STATE, "no segment");
  std::vector<RowRec> r((size_t)c->nh);
  HC(hipMemcpy(r.data(), c->d_rows, sizeof(RowRec) * c->nh, hipMemcpyDeviceToHost));
  for (int64_t h = 0; h < c->nh; ++h) {
    if (u) u[h] = r[h].u;
    if (v) v[h] = r[h].v;
    if (w) w[h] = r[h].w;
  }
  return VW_OK;
}

namespace {
/* full class table of one tile set for the current segment, on demand */
int full_classes(Ctx* c, bool img, std::vector<uint8_t>& t, uint8_t** dev_out) {
  int64_t nt = img ? c->ngt : c->ntile;
  t.assign((size_t)(nt * c->nh_pad), 0);
  uint8_t *d = nullptr, *bits = nullptr;
  HC(hipMalloc(&d, t.size()));
  hipError_t e = hipMalloc(&bits, (size_t)(nt * c->nfr_pad));
  if (e == hipSuccess) {
    k_ant_bits<<<blocks_for(nt * c->nfr_pad), 256, 0, c->st>>>(img ? c->d_gcone : c->d_cone, nt, 0, nt, c->d_ant8,
                                                               (int64_t)c->tlen * c->A, c->nfr_pad, bits);
    launch_classify(c, bits, nt, d, c->st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(t.data(), d, t.size(), hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (bits) (void)hipFree(bits);
  if (e != hipSuccess || !dev_out) (void)hipFree(d);
  if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("class table: ") + hipGetErrorString(e));
  if (dev_out) *dev_out = d;
  return VW_OK;
}

}  // namespace

VW_API int vw_get_tile_classes(void* p, int32_t img, uint8_t* cls) {
  Ctx* c = as_ctx(p);
  if (!c || !cls) return VW_ERR_ARG;
  if (!c->have_seg) return c->fail(VW_ERR_STATE, "no segment");
  if (img && !c->ngp) return c->fail(VW_ERR_ARG, "no lm grid");
  int64_t nt = img ? c->ngt : c->ntile;
  std::vector<uint8_t> t;
  int rc = full_classes(c, img != 0, t, nullptr);
  if (rc) return rc;
  for (int64_t h = 0; h < c->nh; ++h)
    for (int64_t k = 0; k < nt; ++k) cls[h * nt + k] = t[k * c->nh_pad + h];
  return VW_OK;
}

VW_API int vw_get_masks(void* p, int64_t row0, int64_t nrows, uint8_t* mask) {
  Ctx* c = as_ctx(p);
  if (!c || !mask) return VW_ERR_ARG;
  if (!c->have_seg) return c->fail(VW_ERR_STATE, "no segment");
  if (row0 < 0 || nrows < 0 || row0 + nrows > c->nh) return c->fail(VW_ERR_ARG, "row range");
  std::vector<uint8_t> t;
  uint8_t* dcls = nullptr;
  int rc = full_classes(c, false, t, &dcls);
  if (rc) return rc;
  int64_t tot = nrows * c->npix;
  uint8_t* d = nullptr;
  hipError_t e = hipMalloc(&d, tot ? tot : 1);
  if (e == hipSuccess) {
    k_masks<<<blocks_for(tot), 256, 0, c->st>>>(c->d_l, c->d_m, c->d_n, c->npix, c->TP, dcls, c->nh_pad, c->d_rows,
                                                c->d_ant8, row0, nrows, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(mask, d, tot, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(d);
  (void)hipFree(dcls);
  if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("vw_get_masks: ") + hipGetErrorString(e));
  return VW_OK;
}

VW_API int vw_get_viss_counts(void* p, int64_t* counts) {
  Ctx* c = as_ctx(p);
  if (!c || !counts) return VW_ERR_ARG;
  int rc = enq_viss(c, true);
  if (rc) return rc;
  HC(hipMemcpyAsync(counts, c->d_counts, sizeof(int64_t) * c->nh, hipMemcpyDeviceToHost, c->st));
  return sync(c);
}

VW_API int vw_get_groups(void* p, vw_seg_info* info, int64_t* keys, int32_t* offsets, int32_t* idx_sorted,
                         int32_t* rep) {
  Ctx* c = as_ctx(p);
  if (!c) return VW_ERR_ARG;
  if (!c->have_groups) return c->fail(VW_ERR_STATE, "no groups (call vw_recon_segment)");
  vw_seg_info si;
  int rc = read_seg_info(c, &si);
  if (rc) return rc;
  if (info) *info = si;
  int64_t nr = c->cfg.recon_mode == VW_RECON_FULL ? 2 * c->nh : c->nh;
  if (keys) HC(hipMemcpy(keys, c->d_gkeys, sizeof(int64_t) * si.nuniq, hipMemcpyDeviceToHost));
  if (offsets) HC(hipMemcpy(offsets, c->d_goff, sizeof(int32_t) * (si.nuniq + 1), hipMemcpyDeviceToHost));
  if (idx_sorted) HC(hipMemcpy(idx_sorted, c->d_idx_sorted, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
  if (rep) {
    std::vector<int32_t> srt((size_t)nr), off((size_t)si.nuniq + 1);
    HC(hipMemcpy(srt.data(), c->d_idx_sorted, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
    HC(hipMemcpy(off.data(), c->d_goff, sizeof(int32_t) * (si.nuniq + 1), hipMemcpyDeviceToHost));
    for (int64_t q = 0; q < si.nuniq; ++q) rep[q] = srt[off[q] + (off[q + 1] - off[q]) / 2];
  }
  return VW_OK;
}
