/*
 * vw_oracle.c -- CPU FP64 restatement of the reference path (test
 * infrastructure only; see vw_oracle.h).
 *
 * Stage by stage, with the reference file:line each part restates:
 *   plan / schedule          main_recon_online.cu:88-139, orbit_online.cuh:233-336
 *   t1s windows (mt19937_64) orbit_online.cuh:317-333
 *   lunar positions          orbit_online.cuh:39-184 (vw_lunar_sat)
 *   half rows u,v,w          orbit_online.cuh:186-231
 *   pix2lmn                  viss_recon_kernel.cuh:29-112
 *   tile cones / classes     mirror_tile_vissfast/viss_recon_kernel_mirror_tile_vissfast.cuh:21-98
 *   Viss (tile-cone form)    mirror_tile_vissfast/...vissfast.cuh:437-561
 *   phase correct, conj fill common.hpp:255-277
 *   fa/fb, RES, du           viss_recon_kernel.cuh:673-701, main_recon_online.cu:370-383
 *   keys, stable grouping    viss_recon_kernel.cuh:295-467, recon_mirror/viss_recon_kernel_mirror.cuh:462-610
 *   representative (occ 2)   mirror_tile_vissfast/...vissfast.cuh:867-875 (idx_sorted[s0+L/2])
 *   recon kernels            viss_recon_kernel.cuh:470-639, mirror_tile_vissfast/...vissfast.cuh:1014-1287
 *   accumulate / normalise   viss_recon_kernel.cuh:641-659, main_recon_online.cu:497-512
 * Circular-orbit geometry (seeded Earth orbits) has no reference counterpart;
 * it reuses the reference's orbit rotation (orbit_online.cuh:59-77).
 */
#include "vw_oracle.h"
#include "../viss_optimize_ws_amd/csrc/vw_shared.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* mt19937_64 from its published definition                            */
/* ------------------------------------------------------------------ */
void vo_mt64_seed(vo_mt64* g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
  g->idx = 312;
}

uint64_t vo_mt64_next(vo_mt64* g) {
  const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  if (g->idx >= 312) {
    for (int i = 0; i < 312; ++i) {
      uint64_t x = (g->mt[i] & UM) | (g->mt[(i + 1) % 312] & LM);
      uint64_t xa = x >> 1;
      if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
      g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
    }
    g->idx = 0;
  }
  uint64_t y = g->mt[g->idx++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= y >> 43;
  return y;
}

static double vo_u01(vo_mt64* g) { return (double)(vo_mt64_next(g) >> 11) * (1.0 / 9007199254740992.0); }

/* round_away_from_zero_host (orbit_online.cuh:18-20), float arithmetic */
static int vo_round_away_f(float x) { return (x >= 0.0f) ? (int)floorf(x + 0.5f) : (int)ceilf(x - 0.5f); }

/* ------------------------------------------------------------------ */
/* context                                                              */
/* ------------------------------------------------------------------ */
typedef struct {
  vw_config cfg;
  char err[256];
  int64_t nside, npix, ntile, TP;
  int A, P, T, segs, max_tlen;
  int32_t *seg_t0, *seg_tlen;
  double lambda, bl_max, cosphi_glob, pix_area, occ_R;
  int64_t orbit_res, seg_len, tw, day_offset, pc;
  int64_t* t1s;
  double *c_rad, *c_phase, *c_rate, *c_sr, *c_cr, *c_si, *c_ci;
  int *pa, *pb;
  double *l, *m, *n, *nm1, *cone;
  int64_t ngp, ngt;
  int64_t* gidx;
  double *gl, *gm, *gn, *gcone;
  double* pos;
  double* B;
  int have_sky;
  /* segment */
  int t0, tlen;
  int64_t nh;
  double *ru, *rv, *rw, *ant;
  vw_c128 *vraw, *vh, *vfull;
  int have_seg, have_viss, have_conj, have_groups;
  vw_segparams sp;
  int64_t nrows_grp, nuniq;
  int64_t* keys;
  int32_t* idx_sorted;
  int64_t* gkeys;
  int32_t *goff, *grep;
  vw_c128* gavg;
  /* accumulators */
  double *Cacc, *Gacc;
  uint32_t *Wacc, *GWacc;
  int64_t Wtotal;
} vo_ctx;

static char g_create_err[256];

static int fail(vo_ctx* c, int code, const char* msg) {
  if (c) snprintf(c->err, sizeof c->err, "%s", msg);
  else snprintf(g_create_err, sizeof g_create_err, "%s", msg);
  return code;
}

const char* vo_last_error(void* ctx) { return ctx ? ((vo_ctx*)ctx)->err : g_create_err; }

static void* xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* table of (cos, sin)(2 pi k / VW_TAB_N) for the table-driven term sincos */
static double g_tab[2 * VW_TAB_N];
static void build_tab(void) {
  for (int k = 0; k < VW_TAB_N; ++k) vw_sincos_tab_build(k, g_tab);
}

/* pairwise tile sum in the canonical order: for s = TP/2 .. 1: a[i] += a[i+s] */
static double tree_sum(double* a, int64_t TP) {
  for (int64_t s = TP / 2; s > 0; s >>= 1)
    for (int64_t i = 0; i < s; ++i) a[i] = a[i] + a[i + s];
  return a[0];
}

/* build_tile_cone_meta_kernel (vissfast .cuh:21-81) in FP64 */
static void build_cones(const double* l, const double* m, const double* n, int64_t npts,
                        int64_t TP, int64_t ntile, double* cone) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < ntile; ++t) {
    double* bx = (double*)malloc(sizeof(double) * 3 * TP);
    double *by = bx + TP, *bz = by + TP;
    int64_t p0 = t * TP;
    int64_t cnt = npts - p0 < TP ? npts - p0 : TP;
    for (int64_t i = 0; i < TP; ++i) {
      if (i < cnt) { bx[i] = l[p0 + i]; by[i] = m[p0 + i]; bz[i] = n[p0 + i]; }
      else { bx[i] = 0.0; by[i] = 0.0; bz[i] = 0.0; }
    }
    double sx = tree_sum(bx, TP), sy = tree_sum(by, TP), sz = tree_sum(bz, TP);
    double cx, cy, cz;
    vw_cone_centre(sx, sy, sz, &cx, &cy, &cz);
    double mind = 1.0;
    for (int64_t i = 0; i < cnt; ++i) {
      double d = vw_dot3(cx, cy, cz, l[p0 + i], m[p0 + i], n[p0 + i]);
      mind = fmin(mind, d);
    }
    double ca, sa;
    vw_cone_radius(mind, &ca, &sa);
    cone[0 * ntile + t] = cx;
    cone[1 * ntile + t] = cy;
    cone[2 * ntile + t] = cz;
    cone[3 * ntile + t] = ca;
    cone[4 * ntile + t] = sa;
    free(bx);
  }
}

static void day_positions(vo_ctx* c) {
  int A = c->A;
  for (int t = 0; t < c->T; ++t) {
    for (int a = 0; a < A; ++a) {
      double x, y, z;
      if (c->cfg.geometry == VW_GEOM_LUNAR8) {
        int64_t orb = t / c->seg_len, loc = t % c->seg_len;
        int64_t row = c->t1s[orb] + loc;
        vw_lunar_sat(row, a, c->orbit_res, c->cfg.day, c->tw, c->day_offset, &x, &y, &z);
      } else {
        vw_circ_orbit(c->c_rad[a], c->c_phase[a], c->c_rate[a], (double)t * c->cfg.dt_s,
                      c->c_sr[a], c->c_cr[a], c->c_si[a], c->c_ci[a], &x, &y, &z);
      }
      int64_t k = (int64_t)t * A + a;
      c->pos[0 * (int64_t)c->T * A + k] = x;
      c->pos[1 * (int64_t)c->T * A + k] = y;
      c->pos[2 * (int64_t)c->T * A + k] = z;
    }
  }
}

static int setup_day(vo_ctx* c) {
  if (c->cfg.geometry == VW_GEOM_LUNAR8) {
    int64_t OR = c->orbit_res;
    c->tw = vo_round_away_f((float)OR * 7.0f * 24.0f / 2.3f);
    int mod14 = ((c->cfg.day - 1) % 14 + 14) % 14;
    c->day_offset = vo_round_away_f((float)mod14 * (float)OR * 24.0f / 2.3f);
    uint64_t day_seed = c->cfg.orbit_seed ^ (uint64_t)(int64_t)c->cfg.day * 0x9E3779B97F4A7C15ULL;
    vo_mt64 g;
    vo_mt64_seed(&g, day_seed);
    for (int64_t k = 0; k < c->pc; ++k) {
      int64_t os = k * OR, oe = (k + 1) * OR - 1;
      double u01 = (double)(vo_mt64_next(&g) >> 11) * (1.0 / 9007199254740992.0);
      int64_t t1 = os + vo_round_away_f((float)(u01 * (double)OR * (2.0 / 3.0)));
      int64_t t2 = t1 + c->seg_len - 1;
      if (t2 > oe) { t2 = oe; t1 = t2 - c->seg_len + 1; }
      if (t1 < os) t1 = os;
      if (t1 + c->seg_len - 1 > oe) t1 = oe - c->seg_len + 1;
      c->t1s[k] = t1;
    }
  }
  day_positions(c);
  /* automatic bl_max: lunar 100 km (main_recon_online.cu:90); circular: the
   * longest 3-D baseline of the run */
  if (c->cfg.bl_max_m > 0.0) {
    c->bl_max = c->cfg.bl_max_m;
  } else if (c->cfg.geometry == VW_GEOM_LUNAR8) {
    c->bl_max = 100e3;
  } else {
    double mx = 0.0;
    int64_t TA = (int64_t)c->T * c->A;
    for (int t = 0; t < c->T; ++t)
      for (int j = 0; j < c->P; ++j) {
        int64_t ka = (int64_t)t * c->A + c->pa[j], kb = (int64_t)t * c->A + c->pb[j];
        double dx = c->pos[ka] - c->pos[kb];
        double dy = c->pos[TA + ka] - c->pos[TA + kb];
        double dz = c->pos[2 * TA + ka] - c->pos[2 * TA + kb];
        double d = sqrt(vw_dot3(dx, dy, dz, dx, dy, dz)) * 1000.0;
        mx = fmax(mx, d);
      }
    c->bl_max = mx;
  }
  return VW_OK;
}

int vo_create(const vw_config* cfg, void** out) {
  *out = NULL;
  if (!cfg || cfg->nside <= 0 || cfg->freq_hz <= 0.0) return fail(NULL, VW_ERR_ARG, "bad nside/freq");
  if (cfg->tile_pix != 256 && cfg->tile_pix != 512) return fail(NULL, VW_ERR_ARG, "tile_pix must be 256 or 512");
  if (cfg->geometry == VW_GEOM_LUNAR8 && cfg->n_ant != 8) return fail(NULL, VW_ERR_ARG, "lunar geometry has 8 antennas");
  if (cfg->n_ant < 2 || cfg->n_ant > 256) return fail(NULL, VW_ERR_ARG, "n_ant out of range");
  vo_ctx* c = (vo_ctx*)xcalloc(1, sizeof(vo_ctx));
  build_tab();
  c->cfg = *cfg;
  c->cfg.seg_tlen = NULL;
  c->nside = cfg->nside;
  c->npix = 12 * c->nside * c->nside;
  c->TP = cfg->tile_pix;
  c->ntile = (c->npix + c->TP - 1) / c->TP;
  c->A = cfg->n_ant;
  c->P = c->A * (c->A - 1) / 2;
  c->lambda = 3e8 / cfg->freq_hz;
  c->pix_area = (4.0 * VW_PI) / (double)c->npix;
  c->occ_R = cfg->occ_radius_km;
  c->cosphi_glob = (cfg->occ_altitude_km >= 0.0) ? vw_cosphi_from_altitude(cfg->occ_radius_km, cfg->occ_altitude_km) : NAN;
  /* epochs */
  if (cfg->geometry == VW_GEOM_LUNAR8) {
    c->orbit_res = (int64_t)ceil((2.0 * VW_PI) * (100e3 / c->lambda));
    c->pc = vo_round_away_f(24.0f / 2.3f);
    c->seg_len = c->orbit_res / 3;
    c->T = (int)(c->pc * c->seg_len);
    c->t1s = (int64_t*)xcalloc((size_t)c->pc, sizeof(int64_t));
  } else {
    if (cfg->n_epochs <= 0) { free(c); return fail(NULL, VW_ERR_ARG, "n_epochs must be > 0"); }
    c->T = cfg->n_epochs;
    vo_mt64 g;
    vo_mt64_seed(&g, cfg->orbit_seed);
    c->c_rad = (double*)xcalloc(7 * (size_t)c->A, sizeof(double));
    c->c_phase = c->c_rad + c->A; c->c_rate = c->c_phase + c->A;
    c->c_sr = c->c_rate + c->A; c->c_cr = c->c_sr + c->A;
    c->c_si = c->c_cr + c->A; c->c_ci = c->c_si + c->A;
    for (int a = 0; a < c->A; ++a) {
      double r = cfg->r_min_km + (cfg->r_max_km - cfg->r_min_km) * vo_u01(&g);
      double inc = 180.0 * vo_u01(&g);
      double raan = 360.0 * vo_u01(&g);
      double ph = 360.0 * vo_u01(&g);
      c->c_rad[a] = r;
      c->c_rate[a] = sqrt(cfg->gm_km3_s2 / ((r * r) * r)) * VW_INV_TWOPI;
      c->c_phase[a] = ph / 360.0;
      vw_sincos_turn(raan / 360.0, &c->c_sr[a], &c->c_cr[a]);
      vw_sincos_turn(inc / 360.0, &c->c_si[a], &c->c_ci[a]);
    }
  }
  /* schedule (main_recon_online.cu:123-139) */
  if (cfg->n_seg_explicit > 0 && cfg->seg_tlen) {
    c->segs = cfg->n_seg_explicit;
    c->seg_t0 = (int32_t*)xcalloc((size_t)c->segs, sizeof(int32_t));
    c->seg_tlen = (int32_t*)xcalloc((size_t)c->segs, sizeof(int32_t));
    int64_t cur = 0;
    for (int k = 0; k < c->segs; ++k) {
      if (cfg->seg_tlen[k] <= 0) { vo_destroy(c); return fail(NULL, VW_ERR_ARG, "segment length <= 0"); }
      c->seg_t0[k] = (int32_t)cur;
      c->seg_tlen[k] = cfg->seg_tlen[k];
      cur += cfg->seg_tlen[k];
    }
    if (cur > c->T) { vo_destroy(c); return fail(NULL, VW_ERR_ARG, "segments exceed the epochs of the run"); }
  } else {
    int segs = cfg->segs;
    if (segs <= 0) segs = 1;
    if (segs > c->T) segs = c->T;
    c->segs = segs;
    c->seg_t0 = (int32_t*)xcalloc((size_t)segs, sizeof(int32_t));
    c->seg_tlen = (int32_t*)xcalloc((size_t)segs, sizeof(int32_t));
    int base = c->T / segs, rem = c->T % segs, cur = 0;
    for (int k = 0; k < segs; ++k) {
      int len = base + (k < rem ? 1 : 0);
      c->seg_t0[k] = cur;
      c->seg_tlen[k] = len;
      cur += len;
    }
  }
  c->max_tlen = 0;
  for (int k = 0; k < c->segs; ++k) if (c->seg_tlen[k] > c->max_tlen) c->max_tlen = c->seg_tlen[k];
  /* pairs (orbit_online.cuh:25-37): k -> (m<n) lexicographic; endpoint a = n */
  c->pa = (int*)xcalloc((size_t)c->P, sizeof(int));
  c->pb = (int*)xcalloc((size_t)c->P, sizeof(int));
  {
    int k = 0;
    for (int mm = 0; mm < c->A - 1; ++mm)
      for (int nn = mm + 1; nn < c->A; ++nn) { c->pa[k] = nn; c->pb[k] = mm; ++k; }
  }
  /* resident pixel directions */
  c->l = (double*)xcalloc(4 * (size_t)c->npix, sizeof(double));
  c->m = c->l + c->npix; c->n = c->m + c->npix; c->nm1 = c->n + c->npix;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < c->npix; ++p) {
    vw_pix2lmn_nest(c->nside, p, &c->l[p], &c->m[p], &c->n[p]);
    c->nm1[p] = c->n[p] - 1.0;
  }
  c->cone = (double*)xcalloc(5 * (size_t)c->ntile, sizeof(double));
  build_cones(c->l, c->m, c->n, c->npix, c->TP, c->ntile, c->cone);
  /* lm grid (16x16 blocks, row-major blocks) */
  if (cfg->img_n > 0) {
    int64_t N = cfg->img_n, cnt = 0;
    c->gidx = (int64_t*)xcalloc((size_t)(N * N), sizeof(int64_t));
    for (int64_t bi = 0; bi < N; bi += 16)
      for (int64_t bj = 0; bj < N; bj += 16)
        for (int64_t ii = bi; ii < bi + 16 && ii < N; ++ii)
          for (int64_t jj = bj; jj < bj + 16 && jj < N; ++jj) {
            double l, m, n;
            if (vw_grid_point(N, ii, jj, &l, &m, &n)) c->gidx[cnt++] = ii * N + jj;
          }
    c->ngp = cnt;
    c->ngt = (cnt + c->TP - 1) / c->TP;
    c->gl = (double*)xcalloc(3 * (size_t)cnt, sizeof(double));
    c->gm = c->gl + cnt; c->gn = c->gm + cnt;
    for (int64_t k = 0; k < cnt; ++k)
      vw_grid_point(N, c->gidx[k] / N, c->gidx[k] % N, &c->gl[k], &c->gm[k], &c->gn[k]);
    c->gcone = (double*)xcalloc(5 * (size_t)c->ngt, sizeof(double));
    build_cones(c->gl, c->gm, c->gn, c->ngp, c->TP, c->ngt, c->gcone);
  }
  c->pos = (double*)xcalloc(3 * (size_t)c->T * c->A, sizeof(double));
  setup_day(c);
  /* segment buffers */
  int64_t mh = (int64_t)c->max_tlen * c->P;
  c->ru = (double*)xcalloc(3 * (size_t)mh, sizeof(double));
  c->rv = c->ru + mh; c->rw = c->rv + mh;
  c->ant = (double*)xcalloc(8 * (size_t)c->max_tlen * c->A, sizeof(double));
  c->vraw = (vw_c128*)xcalloc((size_t)mh, sizeof(vw_c128));
  c->vh = (vw_c128*)xcalloc((size_t)mh, sizeof(vw_c128));
  c->vfull = (vw_c128*)xcalloc(2 * (size_t)mh, sizeof(vw_c128));
  c->keys = (int64_t*)xcalloc(2 * (size_t)mh, sizeof(int64_t));
  c->idx_sorted = (int32_t*)xcalloc(2 * (size_t)mh, sizeof(int32_t));
  c->gkeys = (int64_t*)xcalloc(2 * (size_t)mh + 1, sizeof(int64_t));
  c->goff = (int32_t*)xcalloc(2 * (size_t)mh + 2, sizeof(int32_t));
  c->grep = (int32_t*)xcalloc(2 * (size_t)mh + 1, sizeof(int32_t));
  c->gavg = (vw_c128*)xcalloc(2 * (size_t)mh + 1, sizeof(vw_c128));
  c->Cacc = (double*)xcalloc((size_t)c->npix, sizeof(double));
  c->Wacc = (uint32_t*)xcalloc((size_t)c->npix, sizeof(uint32_t));
  c->Gacc = (double*)xcalloc((size_t)(c->ngp ? c->ngp : 1), sizeof(double));
  c->GWacc = (uint32_t*)xcalloc((size_t)(c->ngp ? c->ngp : 1), sizeof(uint32_t));
  c->B = (double*)xcalloc((size_t)c->npix, sizeof(double));
  *out = c;
  return VW_OK;
}

void vo_destroy(void* ctx) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c) return;
  free(c->seg_t0); free(c->seg_tlen); free(c->t1s); free(c->c_rad); free(c->pa); free(c->pb);
  free(c->l); free(c->cone); free(c->gidx); free(c->gl); free(c->gcone); free(c->pos); free(c->B);
  free(c->ru); free(c->ant); free(c->vraw); free(c->vh); free(c->vfull); free(c->keys);
  free(c->idx_sorted); free(c->gkeys); free(c->goff); free(c->grep); free(c->gavg);
  free(c->Cacc); free(c->Wacc); free(c->Gacc); free(c->GWacc);
  free(c);
}

int vo_get_plan(void* ctx, vw_plan* o) {
  vo_ctx* c = (vo_ctx*)ctx;
  memset(o, 0, sizeof *o);
  o->npix = c->npix; o->ntile = c->ntile; o->n_img_pts = c->ngp; o->n_img_tiles = c->ngt;
  o->n_pairs = c->P; o->n_epochs = c->T; o->segs = c->segs; o->max_tlen = c->max_tlen;
  o->orbit_res = c->orbit_res; o->seg_len = c->seg_len; o->lambda_m = c->lambda;
  o->bl_max_m = c->bl_max; o->cosphi = c->cosphi_glob; o->pix_area = c->pix_area; o->device_bytes = 0;
  return VW_OK;
}

int vo_get_schedule(void* ctx, int32_t* t0, int32_t* tl) {
  vo_ctx* c = (vo_ctx*)ctx;
  memcpy(t0, c->seg_t0, sizeof(int32_t) * c->segs);
  memcpy(tl, c->seg_tlen, sizeof(int32_t) * c->segs);
  return VW_OK;
}

int vo_set_sky(void* ctx, const double* B, int64_t npix, int32_t nsky) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (npix != c->npix || nsky != 1) return fail(c, VW_ERR_ARG, "sky size mismatch");
  for (int64_t p = 0; p < npix; ++p) c->B[p] = B[p] * c->pix_area;
  c->have_sky = 1;
  return VW_OK;
}

int vo_gen_segment(void* ctx, int32_t t0, int32_t tlen, vw_geom_host* o) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (t0 < 0 || tlen <= 0 || tlen > c->max_tlen || t0 + tlen > c->T) return fail(c, VW_ERR_ARG, "segment out of range");
  c->t0 = t0; c->tlen = tlen; c->nh = (int64_t)tlen * c->P;
  int A = c->A, P = c->P;
  int64_t TA = (int64_t)c->T * A;
  for (int g = 0; g < tlen; ++g)
    for (int a = 0; a < A; ++a) {
      int64_t k = (int64_t)(t0 + g) * A + a;
      double px = c->pos[k], py = c->pos[TA + k], pz = c->pos[2 * TA + k];
      double* q = c->ant + ((int64_t)g * A + a) * 8;
      q[0] = px; q[1] = py; q[2] = pz;
      vw_antenna_frame(px, py, pz, c->cfg.occ_altitude_km >= 0.0 ? 0.0 : c->occ_R, c->cosphi_glob,
                       &q[3], &q[4], &q[5], &q[6], &q[7]);
    }
  for (int64_t h = 0; h < c->nh; ++h) {
    int64_t g = h / P, j = h % P;
    int64_t ka = (int64_t)(t0 + g) * A + c->pa[j], kb = (int64_t)(t0 + g) * A + c->pb[j];
    c->ru[h] = vw_uvw_comp(c->pos[ka], c->pos[kb], c->lambda);
    c->rv[h] = vw_uvw_comp(c->pos[TA + ka], c->pos[TA + kb], c->lambda);
    c->rw[h] = vw_uvw_comp(c->pos[2 * TA + ka], c->pos[2 * TA + kb], c->lambda);
  }
  if (o) {
    for (int64_t i = 0; i < 2 * c->nh; ++i) {
      int64_t g = i / (2 * P), jj = i % (2 * P);
      int neg = jj >= P;
      int64_t j = neg ? jj - P : jj;
      int64_t h = g * P + j;
      int64_t ka = (int64_t)(t0 + g) * A + c->pa[j], kb = (int64_t)(t0 + g) * A + c->pb[j];
      int64_t k1 = neg ? kb : ka, k2 = neg ? ka : kb;
      o->u[i] = neg ? -c->ru[h] : c->ru[h];
      o->v[i] = neg ? -c->rv[h] : c->rv[h];
      o->w[i] = neg ? -c->rw[h] : c->rw[h];
      o->xa[i] = c->pos[k1]; o->ya[i] = c->pos[TA + k1]; o->za[i] = c->pos[2 * TA + k1];
      o->xb[i] = c->pos[k2]; o->yb[i] = c->pos[TA + k2]; o->zb[i] = c->pos[2 * TA + k2];
    }
  }
  c->have_seg = 1;
  c->have_viss = c->have_conj = c->have_groups = 0;
  return VW_OK;
}

/* class of (half row h, tile t) against a cone table */
static int pair_class(const vo_ctx* c, int64_t h, const double* cone, int64_t ntile, int64_t t) {
  int64_t g = h / c->P, j = h % c->P;
  const double* qa = c->ant + (g * c->A + c->pa[j]) * 8;
  const double* qb = c->ant + (g * c->A + c->pb[j]) * 8;
  double cx = cone[t], cy = cone[ntile + t], cz = cone[2 * ntile + t];
  double ca = cone[3 * ntile + t], sa = cone[4 * ntile + t];
  int ba = vw_cone_bits(qa[4], qa[5], qa[6], cx, cy, cz, ca, sa, qa[7]);
  int bb = vw_cone_bits(qb[4], qb[5], qb[6], cx, cy, cz, ca, sa, qb[7]);
  return vw_pair_class(ba, bb);
}

static int predicate(const vo_ctx* c, int64_t h, double l, double m, double n) {
  int64_t g = h / c->P, j = h % c->P;
  const double* qa = c->ant + (g * c->A + c->pa[j]) * 8;
  const double* qb = c->ant + (g * c->A + c->pb[j]) * 8;
  return vw_visible(l, m, n, qa[0], qa[1], qa[2], qa[3], qa[7]) &&
         vw_visible(l, m, n, qb[0], qb[1], qb[2], qb[3], qb[7]);
}

/* effective mask: tile class, per-point predicate on mixed tiles */
static int eff_mask(const vo_ctx* c, int64_t h, int img, int64_t pt) {
  const double* cone = img ? c->gcone : c->cone;
  int64_t ntile = img ? c->ngt : c->ntile;
  int cls = pair_class(c, h, cone, ntile, pt / c->TP);
  if (cls == VW_CLS_VISIBLE) return 1;
  if (cls == VW_CLS_HIDDEN) return 0;
  if (img) return predicate(c, h, c->gl[pt], c->gm[pt], c->gn[pt]);
  return predicate(c, h, c->l[pt], c->m[pt], c->n[pt]);
}

int vo_viss_half(void* ctx, int64_t row0, int64_t nrows, vw_c128* out, int64_t* counts) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg || !c->have_sky) return fail(c, VW_ERR_STATE, "need sky and segment");
  if (nrows < 0) { row0 = 0; nrows = c->nh; }
  if (row0 < 0 || row0 + nrows > c->nh) return fail(c, VW_ERR_ARG, "row range");
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t h = row0; h < row0 + nrows; ++h) {
    double u = c->ru[h], v = c->rv[h], w = c->rw[h];
    double re = 0.0, im = 0.0;
    int64_t cnt = 0;
    for (int64_t t = 0; t < c->ntile; ++t) {
      int cls = pair_class(c, h, c->cone, c->ntile, t);
      if (cls == VW_CLS_HIDDEN) continue;
      int64_t p1 = (t + 1) * c->TP < c->npix ? (t + 1) * c->TP : c->npix;
      for (int64_t p = t * c->TP; p < p1; ++p) {
        if (cls == VW_CLS_MIXED && !predicate(c, h, c->l[p], c->m[p], c->n[p])) continue;
        double s, co;
        vw_sincos_tab(vw_viss_phase(u, v, w, c->l[p], c->m[p], c->nm1[p]) * VW_TAB_SCALE, g_tab, &s, &co);
        re = re + c->B[p] * co;
        im = im - c->B[p] * s;
        ++cnt;
      }
    }
    c->vraw[h].re = re;
    c->vraw[h].im = im;
    if (out) out[h - row0] = c->vraw[h];
    if (counts) counts[h - row0] = cnt;
  }
  if (row0 == 0 && nrows == c->nh) c->have_viss = 1;
  return VW_OK;
}

/* Adjoint of vo_viss_half over all rows of the current segment: per pixel,
 * g_p = pix_area * sum_h M_hp Re(r_h exp(+j 2 pi (u l + v m + w (n-1))))
 * with the same effective mask (tile class, then the predicate on partly
 * visible tiles) and table sincos as the forward sum.  No reference
 * counterpart (a data-consistency operator of this build's deconvolver);
 * pinned by the adjoint identity <V(B), r> = <B, g(r)> against vo_viss_half. */
int vo_viss_adjoint(void* ctx, const vw_c128* res, double* g) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg) return fail(c, VW_ERR_STATE, "need segment");
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t t = 0; t < c->ntile; ++t) {
    int64_t p1 = (t + 1) * c->TP < c->npix ? (t + 1) * c->TP : c->npix;
    for (int64_t p = t * c->TP; p < p1; ++p) g[p] = 0.0;
    for (int64_t h = 0; h < c->nh; ++h) {
      int cls = pair_class(c, h, c->cone, c->ntile, t);
      if (cls == VW_CLS_HIDDEN) continue;
      double u = c->ru[h], v = c->rv[h], w = c->rw[h];
      for (int64_t p = t * c->TP; p < p1; ++p) {
        if (cls == VW_CLS_MIXED && !predicate(c, h, c->l[p], c->m[p], c->n[p])) continue;
        double s, co;
        vw_sincos_tab(vw_viss_phase(u, v, w, c->l[p], c->m[p], c->nm1[p]) * VW_TAB_SCALE, g_tab, &s, &co);
        g[p] = g[p] + (res[h].re * co - res[h].im * s);
      }
    }
    for (int64_t p = t * c->TP; p < p1; ++p) g[p] = c->pix_area * g[p];
  }
  return VW_OK;
}

int vo_set_viss_raw(void* ctx, const vw_c128* raw) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg) return fail(c, VW_ERR_STATE, "need segment");
  memcpy(c->vraw, raw, sizeof(vw_c128) * (size_t)c->nh);
  c->have_viss = 1;
  c->have_conj = c->have_groups = 0;
  return VW_OK;
}

int vo_conj_fill(void* ctx, vw_c128* oh, vw_c128* of) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_viss) return fail(c, VW_ERR_STATE, "need viss");
  int64_t P = c->P;
  for (int64_t h = 0; h < c->nh; ++h) {
    double re, im;
    vw_phase_correct(c->vraw[h].re, c->vraw[h].im, c->rw[h], &re, &im);
    c->vh[h].re = re;
    c->vh[h].im = im;
    int64_t g = h / P, j = h % P;
    c->vfull[2 * P * g + j] = c->vh[h];
    c->vfull[2 * P * g + P + j].re = re;
    c->vfull[2 * P * g + P + j].im = -im;
  }
  if (oh) memcpy(oh, c->vh, sizeof(vw_c128) * (size_t)c->nh);
  if (of) memcpy(of, c->vfull, sizeof(vw_c128) * (size_t)(2 * c->nh));
  c->have_conj = 1;
  return VW_OK;
}

static const int64_t* g_sort_keys;
static int cmp_key_idx(const void* a, const void* b) {
  int32_t ia = *(const int32_t*)a, ib = *(const int32_t*)b;
  int64_t ka = g_sort_keys[ia], kb = g_sort_keys[ib];
  if (ka < kb) return -1;
  if (ka > kb) return 1;
  return (ia > ib) - (ia < ib);
}

/* mode row index -> half row and sign (full mode rows: 2P per epoch) */
static void row_half(const vo_ctx* c, int64_t r, int64_t* h, int* neg) {
  if (c->cfg.recon_mode == VW_RECON_MIRROR) { *h = r; *neg = 0; return; }
  int64_t P = c->P, g = r / (2 * P), jj = r % (2 * P);
  *neg = jj >= P;
  *h = g * P + (jj >= P ? jj - P : jj);
}

int vo_group_segment(void* ctx, vw_seg_info* info) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_conj) return fail(c, VW_ERR_STATE, "need conj fill");
  int64_t P = c->P;
  /* five sums over the full rows, 256 strided partials + stride-halving tree */
  double part[5][256];
  memset(part, 0, sizeof part);
  for (int64_t i = 0; i < 2 * c->nh; ++i) {
    int64_t g = i / (2 * P), jj = i % (2 * P);
    int neg = jj >= P;
    int64_t h = g * P + (neg ? jj - P : jj);
    double sg = neg ? -1.0 : 1.0;
    double u = sg * c->ru[h], v = sg * c->rv[h], w = sg * c->rw[h];
    int t = (int)(i % 256);
    part[0][t] = part[0][t] + u * u;
    part[1][t] = part[1][t] + v * v;
    part[2][t] = part[2][t] + u * v;
    part[3][t] = part[3][t] + u * w;
    part[4][t] = part[4][t] + v * w;
  }
  for (int k = 0; k < 5; ++k) tree_sum(part[k], 256);
  vw_segparams_from_sums(part[0][0], part[1][0], part[2][0], part[3][0], part[4][0], c->bl_max,
                         c->lambda, &c->sp);
  int mirror = c->cfg.recon_mode == VW_RECON_MIRROR;
  int64_t nr = mirror ? c->nh : 2 * c->nh;
  c->nrows_grp = nr;
  for (int64_t r = 0; r < nr; ++r) {
    int64_t h; int neg;
    row_half(c, r, &h, &neg);
    double sg = neg ? -1.0 : 1.0;
    c->keys[r] = vw_grid_key(sg * c->ru[h], sg * c->rv[h], &c->sp, mirror);
    c->idx_sorted[r] = (int32_t)r;
  }
  g_sort_keys = c->keys;
  qsort(c->idx_sorted, (size_t)nr, sizeof(int32_t), cmp_key_idx);
  int64_t nu = 0;
  for (int64_t k = 0; k < nr; ++k) {
    int64_t key = c->keys[c->idx_sorted[k]];
    if (k == 0 || key != c->gkeys[nu - 1]) { c->gkeys[nu] = key; c->goff[nu] = (int32_t)k; ++nu; }
  }
  c->goff[nu] = (int32_t)nr;
  c->nuniq = nu;
  for (int64_t q = 0; q < nu; ++q) {
    int32_t s0 = c->goff[q], s1 = c->goff[q + 1];
    c->grep[q] = c->idx_sorted[s0 + (s1 - s0) / 2];
    double sr = 0.0, si = 0.0;
    for (int32_t k = s0; k < s1; ++k) {
      int64_t h; int neg;
      row_half(c, c->idx_sorted[k], &h, &neg);
      sr = sr + c->vh[h].re;
      si = si + (neg ? -c->vh[h].im : c->vh[h].im);
    }
    c->gavg[q].re = sr / (double)(s1 - s0);
    c->gavg[q].im = si / (double)(s1 - s0);
  }
  c->have_groups = 1;
  if (info) {
    info->t0 = c->t0; info->tlen = c->tlen; info->n_half = c->nh;
    info->fa = c->sp.fa; info->fb = c->sp.fb; info->denom = c->sp.denom;
    info->res = c->sp.res; info->half = c->sp.half; info->du = c->sp.du;
    info->nuniq = nu;
    info->nkeys = nu - ((nu > 0 && c->gkeys[0] == 0) ? 1 : 0);
  }
  return VW_OK;
}

/* effective mask from the row's class for the point's tile (rowcls[h],
 * computed once per output tile), falling back to the per-point predicate
 * on mixed tiles -- the same value as eff_mask */
static int eff_mask_cls(const vo_ctx* c, const uint8_t* rowcls, int64_t h, int img, int64_t pt) {
  int cls = rowcls[h];
  if (cls == VW_CLS_VISIBLE) return 1;
  if (cls == VW_CLS_HIDDEN) return 0;
  if (img) return predicate(c, h, c->gl[pt], c->gm[pt], c->gn[pt]);
  return predicate(c, h, c->l[pt], c->m[pt], c->n[pt]);
}

/* classes of every half row of the segment for one output tile */
static void tile_row_classes(const vo_ctx* c, int img, int64_t tile, uint8_t* rowcls) {
  const double* cone = img ? c->gcone : c->cone;
  int64_t ntile = img ? c->ngt : c->ntile;
  for (int64_t h = 0; h < c->nh; ++h) rowcls[h] = (uint8_t)pair_class(c, h, cone, ntile, tile);
}

/* recon of one output point for the current segment; rowcls holds the row
 * classes of the point's tile; gu/gv the cell centres of the groups */
static void recon_point(const vo_ctx* c, int img, int64_t pt, const uint8_t* rowcls, const double* gu,
                        const double* gv, double* cseg, uint32_t* wseg) {
  double l = img ? c->gl[pt] : c->l[pt];
  double m = img ? c->gm[pt] : c->m[pt];
  double n = img ? c->gn[pt] : c->n[pt];
  double lp = fma(c->sp.fa, n, l), mp = fma(c->sp.fb, n, m);
  double acc = 0.0;
  uint32_t w = 0;
  int mirror = c->cfg.recon_mode == VW_RECON_MIRROR;
  for (int64_t q = 0; q < c->nuniq; ++q) {
    if (c->gkeys[q] == 0) continue;
    const double t = vw_recon_phase(gu[q], gv[q], lp, mp) * VW_TAB_SCALE;
    double sx, cx; /* sin, cos of the phase, evaluated where the term is used */
    if (!c->cfg.blockage) {
      vw_sincos_tab(t, g_tab, &sx, &cx);
      acc = vw_recon_add(acc, c->gavg[q].re, c->gavg[q].im, sx, cx);
      continue;
    }
    int32_t s0 = c->goff[q], s1 = c->goff[q + 1], L = s1 - s0;
    if (mirror && c->cfg.occ_mode == 2) {
      int64_t h = c->grep[q];
      if (eff_mask_cls(c, rowcls, h, img, pt)) {
        vw_sincos_tab(t, g_tab, &sx, &cx);
        acc = vw_recon_add(acc, c->vh[h].re, c->vh[h].im, sx, cx);
        w += 1;
      }
      continue;
    }
    int32_t posv[4];
    int np;
    if (c->cfg.occ_mode == 0) np = -1;
    else if (c->cfg.occ_mode == 1) { posv[0] = s0; np = 1; }
    else if (c->cfg.occ_mode == 2) { posv[0] = s0 + (L >> 1); np = 1; }
    else { posv[0] = s0; posv[1] = s0 + L / 3; posv[2] = s0 + (2 * L) / 3; posv[3] = s1 - 1; np = 4; }
    double sr = 0.0, si = 0.0;
    int cnt = 0;
    int32_t kend = (np < 0) ? s1 : np;
    for (int32_t kk = (np < 0 ? s0 : 0); kk < kend; ++kk) {
      int32_t pos = (np < 0) ? kk : posv[kk];
      int64_t h; int neg;
      row_half(c, c->idx_sorted[pos], &h, &neg);
      if (!eff_mask_cls(c, rowcls, h, img, pt)) continue;
      double vr = c->vh[h].re, vi = neg ? -c->vh[h].im : c->vh[h].im;
      if (isfinite(vr) && isfinite(vi)) { sr = sr + vr; si = si + vi; ++cnt; }
    }
    if (cnt > 0) {
      double invc = 1.0 / (double)cnt;
      vw_sincos_tab(t, g_tab, &sx, &cx);
      acc = vw_recon_add(acc, sr * invc, si * invc, sx, cx);
      w += 1;
    }
  }
  *cseg = acc;
  *wseg = w;
}

/* group cell centres (wavelengths) of the current segment */
static void group_uv(const vo_ctx* c, double* gu, double* gv) {
  for (int64_t q = 0; q < c->nuniq; ++q) {
    gu[q] = gv[q] = 0.0;
    if (c->gkeys[q] != 0) vw_key_uv(c->gkeys[q], &c->sp, &gu[q], &gv[q]);
  }
}

/* recon of the points [first, first + n) of one output set (img), tile by
 * tile: the row classes of a tile are computed once for all its points */
static void recon_range(const vo_ctx* c, int img, int64_t n, const double* gu, const double* gv, double* acc,
                        uint32_t* wacc, int accumulate) {
  int64_t ntile = (n + c->TP - 1) / c->TP;
#pragma omp parallel
  {
    uint8_t* rowcls = (uint8_t*)xcalloc((size_t)c->nh, 1);
#pragma omp for schedule(dynamic, 1)
    for (int64_t t = 0; t < ntile; ++t) {
      tile_row_classes(c, img, t, rowcls);
      int64_t p1 = (t + 1) * c->TP < n ? (t + 1) * c->TP : n;
      for (int64_t p = t * c->TP; p < p1; ++p) {
        double cs; uint32_t ws;
        recon_point(c, img, p, rowcls, gu, gv, &cs, &ws);
        if (accumulate) { acc[p] = acc[p] + cs; wacc[p] += ws; }
        else { acc[p] = cs; wacc[p] = ws; }
      }
    }
    free(rowcls);
  }
}

int vo_recon_points(void* ctx, int32_t img, const int64_t* pts, int64_t npts, double* cseg, uint32_t* wseg) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_groups) return fail(c, VW_ERR_STATE, "need groups");
  int64_t lim = img ? c->ngp : c->npix;
  for (int64_t k = 0; k < npts; ++k) if (pts[k] < 0 || pts[k] >= lim) return fail(c, VW_ERR_ARG, "point index");
  double* gu = (double*)xcalloc((size_t)c->nuniq, sizeof(double));
  double* gv = (double*)xcalloc((size_t)c->nuniq, sizeof(double));
  group_uv(c, gu, gv);
#pragma omp parallel
  {
    uint8_t* rowcls = (uint8_t*)xcalloc((size_t)c->nh, 1);
    int64_t cur = -1;
#pragma omp for schedule(dynamic, 64)
    for (int64_t k = 0; k < npts; ++k) {
      int64_t t = pts[k] / c->TP;
      if (t != cur) { tile_row_classes(c, img, t, rowcls); cur = t; }
      recon_point(c, img, pts[k], rowcls, gu, gv, &cseg[k], &wseg[k]);
    }
    free(rowcls);
  }
  free(gu); free(gv);
  return VW_OK;
}

int vo_recon_segment(void* ctx, vw_seg_info* info) {
  vo_ctx* c = (vo_ctx*)ctx;
  int rc = vo_group_segment(ctx, info);
  if (rc) return rc;
  double* gu = (double*)xcalloc((size_t)c->nuniq, sizeof(double));
  double* gv = (double*)xcalloc((size_t)c->nuniq, sizeof(double));
  group_uv(c, gu, gv);
  if (c->cfg.healpix_out) recon_range(c, 0, c->npix, gu, gv, c->Cacc, c->Wacc, 1);
  if (c->ngp) recon_range(c, 1, c->ngp, gu, gv, c->Gacc, c->GWacc, 1);
  free(gu); free(gv);
  if (!c->cfg.blockage) c->Wtotal += c->nuniq;
  return VW_OK;
}

int vo_finalize(void* ctx, double* C, uint32_t* W, double* img, uint32_t* Wimg) {
  vo_ctx* c = (vo_ctx*)ctx;
  double inv = (c->Wtotal > 0) ? 1.0 / (double)c->Wtotal : 0.0;
  if (C || W) {
    for (int64_t p = 0; p < c->npix; ++p) {
      double v;
      if (c->cfg.blockage) v = c->Wacc[p] > 0 ? c->Cacc[p] / (double)c->Wacc[p] : 0.0;
      else v = c->Cacc[p] * inv;
      if (C) C[p] = v;
      if (W) W[p] = c->Wacc[p];
    }
  }
  if ((img || Wimg) && c->cfg.img_n > 0) {
    int64_t NN = (int64_t)c->cfg.img_n * c->cfg.img_n;
    if (img) memset(img, 0, sizeof(double) * (size_t)NN);
    if (Wimg) memset(Wimg, 0, sizeof(uint32_t) * (size_t)NN);
    for (int64_t k = 0; k < c->ngp; ++k) {
      double v;
      if (c->cfg.blockage) v = c->GWacc[k] > 0 ? c->Gacc[k] / (double)c->GWacc[k] : 0.0;
      else v = c->Gacc[k] * inv;
      if (img) img[c->gidx[k]] = v;
      if (Wimg) Wimg[c->gidx[k]] = c->GWacc[k];
    }
  }
  return VW_OK;
}

int vo_reset_accumulators(void* ctx) {
  vo_ctx* c = (vo_ctx*)ctx;
  memset(c->Cacc, 0, sizeof(double) * (size_t)c->npix);
  memset(c->Wacc, 0, sizeof(uint32_t) * (size_t)c->npix);
  if (c->ngp) {
    memset(c->Gacc, 0, sizeof(double) * (size_t)c->ngp);
    memset(c->GWacc, 0, sizeof(uint32_t) * (size_t)c->ngp);
  }
  c->Wtotal = 0;
  return VW_OK;
}

int vo_get_lmn(void* ctx, double* l, double* m, double* n, double* nm1) {
  vo_ctx* c = (vo_ctx*)ctx;
  size_t b = sizeof(double) * (size_t)c->npix;
  if (l) memcpy(l, c->l, b);
  if (m) memcpy(m, c->m, b);
  if (n) memcpy(n, c->n, b);
  if (nm1) memcpy(nm1, c->nm1, b);
  return VW_OK;
}

int vo_get_tiles(void* ctx, int32_t img, double* cone5) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (img && !c->ngp) return fail(c, VW_ERR_ARG, "no lm grid");
  memcpy(cone5, img ? c->gcone : c->cone, sizeof(double) * 5 * (size_t)(img ? c->ngt : c->ntile));
  return VW_OK;
}

int vo_get_img_points(void* ctx, int64_t* idx, double* l, double* m, double* n) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->ngp) return fail(c, VW_ERR_ARG, "no lm grid");
  size_t b = sizeof(double) * (size_t)c->ngp;
  if (idx) memcpy(idx, c->gidx, sizeof(int64_t) * (size_t)c->ngp);
  if (l) memcpy(l, c->gl, b);
  if (m) memcpy(m, c->gm, b);
  if (n) memcpy(n, c->gn, b);
  return VW_OK;
}

int vo_get_positions(void* ctx, double* pos) {
  vo_ctx* c = (vo_ctx*)ctx;
  memcpy(pos, c->pos, sizeof(double) * 3 * (size_t)c->T * c->A);
  return VW_OK;
}

int vo_get_rows(void* ctx, double* u, double* v, double* w) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg) return fail(c, VW_ERR_STATE, "need segment");
  size_t b = sizeof(double) * (size_t)c->nh;
  if (u) memcpy(u, c->ru, b);
  if (v) memcpy(v, c->rv, b);
  if (w) memcpy(w, c->rw, b);
  return VW_OK;
}

int vo_get_tile_classes(void* ctx, int32_t img, uint8_t* cls) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg) return fail(c, VW_ERR_STATE, "need segment");
  const double* cone = img ? c->gcone : c->cone;
  int64_t nt = img ? c->ngt : c->ntile;
  if (img && !c->ngp) return fail(c, VW_ERR_ARG, "no lm grid");
#pragma omp parallel for schedule(static)
  for (int64_t h = 0; h < c->nh; ++h)
    for (int64_t t = 0; t < nt; ++t) cls[h * nt + t] = (uint8_t)pair_class(c, h, cone, nt, t);
  return VW_OK;
}

int vo_get_masks(void* ctx, int64_t row0, int64_t nrows, uint8_t* mask) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg || row0 < 0 || row0 + nrows > c->nh) return fail(c, VW_ERR_ARG, "row range");
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nrows; ++r)
    for (int64_t p = 0; p < c->npix; ++p) mask[r * c->npix + p] = (uint8_t)eff_mask(c, row0 + r, 0, p);
  return VW_OK;
}

int vo_get_predicate(void* ctx, int64_t row0, int64_t nrows, uint8_t* mask) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_seg || row0 < 0 || row0 + nrows > c->nh) return fail(c, VW_ERR_ARG, "row range");
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nrows; ++r)
    for (int64_t p = 0; p < c->npix; ++p)
      mask[r * c->npix + p] = (uint8_t)predicate(c, row0 + r, c->l[p], c->m[p], c->n[p]);
  return VW_OK;
}

int vo_get_groups(void* ctx, vw_seg_info* info, int64_t* keys, int32_t* offsets, int32_t* idx_sorted,
                  int32_t* rep) {
  vo_ctx* c = (vo_ctx*)ctx;
  if (!c->have_groups) return fail(c, VW_ERR_STATE, "need groups");
  if (info) {
    info->t0 = c->t0; info->tlen = c->tlen; info->n_half = c->nh;
    info->fa = c->sp.fa; info->fb = c->sp.fb; info->denom = c->sp.denom;
    info->res = c->sp.res; info->half = c->sp.half; info->du = c->sp.du;
    info->nuniq = c->nuniq;
    info->nkeys = c->nuniq - ((c->nuniq > 0 && c->gkeys[0] == 0) ? 1 : 0);
  }
  if (keys) memcpy(keys, c->gkeys, sizeof(int64_t) * (size_t)c->nuniq);
  if (offsets) memcpy(offsets, c->goff, sizeof(int32_t) * (size_t)(c->nuniq + 1));
  if (idx_sorted) memcpy(idx_sorted, c->idx_sorted, sizeof(int32_t) * (size_t)c->nrows_grp);
  if (rep) memcpy(rep, c->grep, sizeof(int32_t) * (size_t)c->nuniq);
  return VW_OK;
}

void vo_sincos(const double* x, int64_t n, int32_t form, double* s, double* c) {
  build_tab();
  for (int64_t i = 0; i < n; ++i) {
    if (form == 0) vw_sincos_turn(x[i], &s[i], &c[i]);
    else vw_sincos_tab(x[i] * VW_TAB_SCALE, g_tab, &s[i], &c[i]);
  }
}
