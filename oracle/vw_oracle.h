/*
 * vw_oracle.h -- CPU FP64 restatement of the reference path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or the timed CPU baseline).  The product path never links it.
 *
 * Stage API mirrors include/vissws.h so a test can drive both in lockstep.
 * Formulas that feed comparisons come from the shared header
 * viss_optimize_ws_amd/csrc/vw_shared.h so masks, tile classes and keys are
 * bit-comparable with the device.  Each function cites the reference
 * file:line it restates (see vw_oracle.c).
 */
#ifndef VW_ORACLE_H
#define VW_ORACLE_H
#include <stdint.h>
#include "../include/vissws.h"

#ifdef __cplusplus
extern "C" {
#endif

/* mt19937_64 (Matsumoto & Nishimura 2000; C++ [rand.predef]). */
typedef struct {
  uint64_t mt[312];
  int idx;
} vo_mt64;
void vo_mt64_seed(vo_mt64* g, uint64_t seed);
uint64_t vo_mt64_next(vo_mt64* g);

int vo_create(const vw_config* cfg, void** ctx_out);
void vo_destroy(void* ctx);
const char* vo_last_error(void* ctx);
int vo_get_plan(void* ctx, vw_plan* out);
int vo_get_schedule(void* ctx, int32_t* seg_t0, int32_t* seg_tlen);
int vo_set_sky(void* ctx, const double* B, int64_t npix, int32_t nsky);
int vo_gen_segment(void* ctx, int32_t t0, int32_t tlen, vw_geom_host* opt_out);
/* rows [row0, row0+nrows) of the current segment (nrows < 0: all).  counts
 * (may be NULL) receives the number of pixels summed per row. */
int vo_viss_half(void* ctx, int64_t row0, int64_t nrows, vw_c128* out_raw, int64_t* counts);
/* Adjoint of vo_viss_half over the current segment's rows: g [npix]. */
int vo_viss_adjoint(void* ctx, const vw_c128* res, double* g);
/* Replace the current segment's raw half sums (e.g. with device values). */
int vo_set_viss_raw(void* ctx, const vw_c128* raw);
int vo_conj_fill(void* ctx, vw_c128* opt_half_out, vw_c128* opt_full_out);
/* Segment params + grouping only. */
int vo_group_segment(void* ctx, vw_seg_info* info);
/* Segment contribution for listed points of one output set (img=0 HEALPix,
 * img=1 lm grid) without accumulating.  Requires vo_group_segment. */
int vo_recon_points(void* ctx, int32_t img, const int64_t* pts, int64_t npts, double* cseg,
                    uint32_t* wseg);
/* Group + recon of every point + accumulate. */
int vo_recon_segment(void* ctx, vw_seg_info* info);
int vo_finalize(void* ctx, double* C_healpix, uint32_t* W_healpix, double* img, uint32_t* W_img);
int vo_reset_accumulators(void* ctx);

int vo_get_lmn(void* ctx, double* l, double* m, double* n, double* nm1);
int vo_get_tiles(void* ctx, int32_t img, double* cone5);
int vo_get_img_points(void* ctx, int64_t* idx, double* l, double* m, double* n);
int vo_get_positions(void* ctx, double* pos);
int vo_get_rows(void* ctx, double* u, double* v, double* w);
int vo_get_tile_classes(void* ctx, int32_t img, uint8_t* cls);
int vo_get_masks(void* ctx, int64_t row0, int64_t nrows, uint8_t* mask);
int vo_get_groups(void* ctx, vw_seg_info* info, int64_t* keys, int32_t* offsets,
                  int32_t* idx_sorted, int32_t* rep);
/* Pure per-pixel occlusion predicate (no tile cones), [nrows][npix]. */
int vo_get_predicate(void* ctx, int64_t row0, int64_t nrows, uint8_t* mask);
/* sin and cos of 2 pi x[i] (x in turns): form 0 = exact reduction + Taylor
 * (vw_sincos_turn), form 1 = table-driven (vw_sincos_tab of 1024 x). */
void vo_sincos(const double* x, int64_t n, int32_t form, double* s, double* c);

#ifdef __cplusplus
}
#endif
#endif
