/*
 * vissws.h -- C-ABI of the MI355X (gfx950) visibility / reconstruction path.
 *
 * The reference (SSshuishui/viss_optimize_ws) has no library or FFI; its
 * path is a CUDA program whose stages are the functions cited below.  Each
 * entry point here replaces one of those stages, keeps the reference's row
 * layout, segment schedule and output semantics, and returns an int status
 * (0 = success) instead of calling exit(1) (common.hpp:31-37).  The caller
 * owns host memory; a context owns all device memory.  One context per host
 * thread; calls on a context are serialised and synchronous at return.
 *
 * Arithmetic: "plain" precision evaluates every stage in FP64 (phase, sincos,
 * accumulation).  "mixed" keeps the phase u*l+v*m+w*(n-1) and its reduction
 * mod 1 turn in FP64, evaluates sin/cos with the FP32 hardware units and
 * accumulates in FP32 with compensated (Kahan) summation across tiles.
 */
#ifndef VISSWS_H
#define VISSWS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VW_OK 0
#define VW_ERR_ARG 1     /* invalid argument / configuration            */
#define VW_ERR_DEVICE 2  /* HIP runtime error or no device              */
#define VW_ERR_STATE 3   /* call out of order (e.g. recon before viss)  */
#define VW_ERR_RANGE 4   /* size beyond a compiled or allocated maximum */
#define VW_ERR_IO 5      /* file could not be read or written           */
#define VW_ERR_HALFSYM 6 /* generated uvw violates half symmetry        */

/* geometry generators */
#define VW_GEOM_LUNAR8 0   /* reference formation: orbit_online.cuh:4-13     */
#define VW_GEOM_CIRCULAR 1 /* seeded random circular orbits around a sphere */

#define VW_PREC_PLAIN 0 /* FP64 throughout                                 */
#define VW_PREC_MIXED 1 /* FP64 phase + FP32 sincos + compensated FP32 sums */

#define VW_RECON_MIRROR 0 /* half uv grid  (recon_mirror/, mirror_tile_*) */
#define VW_RECON_FULL 1   /* full uv grid  (main_recon_online.cu)         */

/* Viss summation form (vw_config.viss_form) */
#define VW_VISS_AUTO 0      /* per-row sums in the plain form; antenna-factorized in the mixed form
                             * for n_ant <= 32                                                   */
#define VW_VISS_ROWS 1      /* one phase / sin / cos term per (half row, pixel): the reference's
                             * phase rounding, bit for bit                                       */
#define VW_VISS_ANT 2       /* antenna-factorized on the matrix cores in either form (n_ant <= 32);
                             * in FP64 it differs from the per-row sums by the rounding of the
                             * phase, about 2^-52 |u| turns per term                             */
#define VW_VISS_ANT_CHECK 3 /* as VW_VISS_ANT, with the correction pass run on every tile that has a
                             * partly visible antenna (exercises that path; slower)              */

typedef struct {
  double re, im;
} vw_c128;

/* Run configuration.  Fill with vw_config_init() and override fields. */
typedef struct {
  int32_t nside;       /* HEALPix Nside (main_recon_online.cu:7)                 */
  double freq_hz;      /* lambda = 3e8 / freq (main_recon_online.cu:88-89)       */
  int32_t geometry;    /* VW_GEOM_*                                              */
  int32_t n_ant;       /* antennas; VW_GEOM_LUNAR8 requires 8                    */
  int32_t n_epochs;    /* circular: epochs in the run; lunar: ignored (derived)  */
  double dt_s;         /* circular: seconds between epochs                       */
  uint64_t orbit_seed; /* --orbit_seed (main_recon_online.cu:17)                 */
  int32_t day;         /* --day (lunar RAAN drift, t1s seed)                     */
  int32_t segs;        /* --segs: equal split (main_recon_online.cu:123-139)     */
  const int32_t* seg_tlen; /* optional explicit segment lengths (epochs)       */
  int32_t n_seg_explicit;  /* number of entries in seg_tlen (0 = equal split) */
  int32_t blockage;    /* --blockage {0,1}                                       */
  int32_t occ_mode;    /* --occ_mode {0,1,2,other=4-sample}                      */
  int32_t recon_mode;  /* VW_RECON_*                                             */
  int32_t tile_pix;    /* --viss_tile_pix {256,512}                              */
  int32_t precision;   /* VW_PREC_*                                              */
  double bl_max_m;     /* uv-grid half extent in metres; <= 0: automatic        */
  double occ_radius_km;   /* occulting sphere radius                             */
  double occ_altitude_km; /* lunar global cone altitude h; < 0: per-antenna cone */
  int32_t img_n;       /* lm-grid image side (0 = no grid image)                 */
  int32_t healpix_out; /* 1: reconstruct the HEALPix map C                       */
  double r_min_km, r_max_km; /* circular: orbit radius range                     */
  double gm_km3_s2;    /* circular: gravitational parameter of the central body */
  int64_t class_budget_bytes; /* device bytes for the per-(tile, row) class tables of
                               * one tile pass (0 = 2 GiB); a smaller budget splits the
                               * Viss and recon stages into more tile passes      */
  int32_t serial_stages;       /* 1: run every stage on one stream (no overlap of the
                                * recon of a segment with the next segment's Viss;
                                * for profiling single kernels)                     */
  int32_t viss_form;           /* VW_VISS_*: summation form of the Viss stage (the
                                * half visibilities agree to FP64 rounding)         */
  int32_t nsky;                /* sky realisations that share the geometry and the
                                * segment schedule (0 or 1: one sky).  Above 1 the
                                * context runs the batched form: mixed precision
                                * only; Viss and recon as BF16 hi/lo matrix-core
                                * products of on-the-fly phasor tiles against the
                                * sky / group-visibility matrix (FP32 accumulate,
                                * FP64 across blocks); recon rules with one member
                                * per group (occ_mode 1, 2) or blockage 0.  No
                                * counterpart in the reference, which runs one sky */
} vw_config;

/* Derived plan (filled by vw_get_plan). */
typedef struct {
  int64_t npix;        /* 12 nside^2                                    */
  int64_t ntile;       /* ceil(npix / tile_pix)                         */
  int64_t n_img_pts;   /* lm-grid points inside the unit disk           */
  int64_t n_img_tiles; /* ceil(n_img_pts / tile_pix)                    */
  int32_t n_pairs;     /* n_ant (n_ant - 1) / 2                         */
  int32_t n_epochs;    /* T                                             */
  int32_t segs;        /* number of segments                            */
  int32_t max_tlen;    /* longest segment (epochs)                      */
  int64_t orbit_res;   /* lunar: OrbitRes (orbit_online.cuh:238)        */
  int64_t seg_len;     /* lunar: OrbitRes / 3                           */
  double lambda_m;
  double bl_max_m;
  double cosphi;       /* global threshold (lunar) or NaN (per-antenna) */
  double pix_area;     /* 4 pi / npix                                   */
  int64_t device_bytes; /* device memory held by the context (0 in the oracle) */
} vw_plan;

/* Per-segment scalars (printed per segment by main_recon_online.cu:482-494). */
typedef struct {
  int32_t t0, tlen;
  int64_t n_half;  /* tlen * n_pairs                              */
  double fa, fb, denom;
  int64_t res, half;
  double du;
  int64_t nuniq;   /* groups including the key-0 group (reference quirk) */
  int64_t nkeys;   /* groups with key != 0 (recon terms per output point) */
} vw_seg_info;

/* Host copy of a generated segment in the reference's full-row layout
 * (orbit_online.cuh:186-231): 2*n_pairs rows per epoch, first half sign -1.
 * Arrays are caller-allocated with 2 * tlen * n_pairs entries. */
typedef struct {
  double *u, *v, *w;
  double *xa, *ya, *za; /* xyza */
  double *xb, *yb, *zb; /* xyzb */
} vw_geom_host;

/* Device time of the last completed stage calls (HIP events, ms). */
typedef struct {
  double gen_ms, viss_ms, conj_ms, group_ms, recon_ms, finalize_ms;
  double viss_kernel_ms;  /* the Viss summation kernel alone (sum)     */
  double recon_kernel_ms; /* the recon kernels alone (sum)             */
  int32_t n_viss_launch;  /* Viss kernel launches summed above         */
  int32_t n_recon_launch; /* recon kernel launches summed above        */
} vw_timings;

/* ---------------- configuration / lifetime ---------------- */

/* Defaults for a geometry: lunar = main_recon_online.cu:6-45 (10M band);
 * circular = Earth orbits r in [6900,7100] km, R = 6371 km. */
int vw_config_init(vw_config* cfg, int32_t geometry);

/* Allocates device state, builds resident lmn / nm1 / tile cones / lm grid and
 * the day geometry.  Replaces the per-GPU init (main_recon_online.cu:157-217)
 * and orbit_gen_init (orbit_online.cuh:233-336). */
int vw_create(const vw_config* cfg, void** ctx_out);
void vw_destroy(void* ctx);

/* Last error message of a context (or of the last failed vw_create when
 * ctx is NULL).  The string stays valid until the next call on ctx. */
const char* vw_last_error(void* ctx);

int vw_get_plan(void* ctx, vw_plan* out);
/* Segment schedule: seg_t0 / seg_tlen arrays of length plan.segs. */
int vw_get_schedule(void* ctx, int32_t* seg_t0, int32_t* seg_tlen);

/* Regenerates the day geometry for another day (lunar RAAN drift and t1s);
 * mirror_tile_vissfast/main_recon_online_mirror_tile_vissfast.cu:274-312. */
int vw_set_day(void* ctx, int32_t day);

/* ---------------- stages ---------------- */

/* Sky intensities B[nsky][npix] (NESTED).  Scaled by pix_area = 4 pi / npix
 * (main_recon_online.cu:105-110).  nsky must equal the configured nsky (1
 * unless batched); batched skies must be finite.
 *
 * Batched contexts (nsky > 1) keep every stage call and its order; the
 * per-sky outputs are stacked: vw_viss_half and vw_conj_fill write
 * [nsky][n_half] (full rows [nsky][2 n_half]); vw_finalize writes
 * C_healpix [nsky][npix] and img [nsky][img_n^2] with one weight map
 * W_healpix [npix] / W_img [img_n^2] (the weights do not depend on the sky);
 * vw_get/set_accumulators take C [nsky][npix] and G [nsky][n_img_pts]. */
int vw_set_sky(void* ctx, const double* B, int64_t npix, int32_t nsky);

/* Generates epochs [t0, t0+tlen): antenna positions, per-antenna occlusion
 * frames and half-row u,v,w.  Replaces orbit_gen_make_segment
 * (orbit_online.cuh:338-366) and its half-symmetry guard
 * (orbit_online.cuh:403-418).  opt_out (may be NULL) receives the full-row
 * layout. */
int vw_gen_segment(void* ctx, int32_t t0, int32_t tlen, vw_geom_host* opt_out);

/* Half-conjugate visibility sum with tile-cone culling: Vh[ih] =
 * sum_p vis_a vis_b B'_p exp(-j 2 pi (u l + v m + w (n-1))), one entry per
 * half row ih in [0, tlen*n_pairs).  Replaces viss_partial_all_halfsym_tilecone
 * (mirror_tile_vissfast/viss_recon_kernel_mirror_tile_vissfast.cuh:437-561)
 * plus the cross-GPU host sum (main_recon_online.cu:329-337).  opt_out (may be
 * NULL) receives the raw (not phase-corrected) sums. */
int vw_viss_half(void* ctx, vw_c128* opt_out);

/* Adjoint of the half-visibility sum of the current segment (one-sky
 * contexts): for residual rows r (raw convention of vw_viss_half, n_half
 * entries) writes g[p] = pix_area * sum_h M_hp Re(r_h exp(+j 2 pi (u l + v m
 * + w (n-1)))) for every pixel (npix entries), with the Viss's own mask M
 * and FP64 phase and sincos: the gradient of 1/2 |V(B) - r_target|^2 with
 * respect to the sky B is g(V(B) - r_target).  For data-consistency
 * refinement of sky estimates; no counterpart in the reference. */
int vw_viss_adjoint(void* ctx, const vw_c128* res, double* g);

/* Phase correction Vh *= exp(-j 2 pi w) and conjugate completion
 * V[2P g + j] = Vh, V[2P g + P + j] = conj(Vh).  Replaces
 * phase_correct_viss_halfsym_host + expand_viss_halfsym_to_full
 * (common.hpp:255-277).  Either output may be NULL. */
int vw_conj_fill(void* ctx, vw_c128* opt_half_out, vw_c128* opt_full_out);

/* Segment reconstruction: w-plane fit, uv-grid keys, stable grouping and the
 * per-point recon accumulated into the HEALPix map and the lm image.
 * Replaces compute_fa_fb_host_matlab_no_intercept (viss_recon_kernel.cuh:673-701),
 * build_groups_device0[_half] (viss_recon_kernel.cuh:349-467,
 * recon_mirror/viss_recon_kernel_mirror.cuh:483-610), the recon kernels
 * (viss_recon_kernel.cuh:470-639, mirror_tile_vissfast/...:1014-1287) and
 * add_inplace (viss_recon_kernel.cuh:641-648). */
int vw_recon_segment(void* ctx, vw_seg_info* opt_info);

/* Runs every segment of the schedule (gen, viss, conj fill, recon). */
int vw_run_all(void* ctx);

/* Normalises the accumulators (viss_recon_kernel.cuh:649-659) and copies
 * them out.  Any pointer may be NULL.  C_healpix/W_healpix: npix entries;
 * img/W_img: img_n*img_n entries (row-major, row = m, column = l; 0 outside
 * the unit disk). */
int vw_finalize(void* ctx, double* C_healpix, uint32_t* W_healpix, double* img,
                uint32_t* W_img);

/* Clears the accumulators (reset_daily_accumulators,
 * mirror_tile_vissfast/main_recon_online_mirror_tile_vissfast.cu:257-265). */
int vw_reset_accumulators(void* ctx);

/* Raw accumulators of the current day, before normalisation: per-pixel sums
 * and contributing-key counts of the HEALPix map (npix each) and of the
 * lm-grid points (n_img_pts each, vw_get_img_points order), and the running
 * key total that blockage=0 divides by (scale_inplace).  For segment-level
 * checkpoint / resume: the reference keeps Cacc/Wacc only in device memory
 * between segments (add_inplace, viss_recon_kernel.cuh:641-659).
 * get: NULL pointers are skipped.  set: a NULL array is cleared. */
int vw_get_accumulators(void* ctx, double* C, uint32_t* W, double* G, uint32_t* GW, int64_t* wtotal);
int vw_set_accumulators(void* ctx, const double* C, const uint32_t* W, const double* G, const uint32_t* GW,
                        int64_t wtotal);

int vw_get_timings(void* ctx, vw_timings* out);

/* ---------------- batched launch shapes ---------------- */

/* Waves per launch that the batched Viss and recon kernels size their chunk
 * counts for (one of 2048, 4096, 8192, 16384, 32768, 65536; defaults 8192
 * and 16384).  Batched contexts only; no counterpart in the reference. */
int vw_set_batch_targets(void* ctx, int64_t viss_waves, int64_t recon_waves);
int vw_get_batch_targets(void* ctx, int64_t* viss_waves, int64_t* recon_waves);
/* Per-sky segment schedules for a batched context (blockage = 1): sky s
 * is segmented into nseg[s] consecutive segments whose lengths (epochs) are
 * the next nseg[s] entries of seg_tlen, covering the run's epochs; every
 * length must lie in [1, the longest segment of the context's own
 * schedule].  vw_run_all and vw_run_all_graph then sum the visibilities of
 * the context's schedule once for all skies (the half visibilities of an
 * epoch do not depend on the segmentation), keep them for the whole day, and
 * group and reconstruct each sky over its own segments (skies that share a
 * segment share its groups), with the one-sky mixed recon kernel and one
 * weight map per sky: vw_finalize then writes W_healpix [nsky][npix] and
 * W_img [nsky][img_n^2].  vw_recon_segment and the accumulator get/set calls
 * are refused in this mode.  nseg = NULL returns to the shared schedule.
 * Allocates the day buffer (nsky x epochs x pairs x 16 bytes) on first use.
 * No counterpart in the reference, which runs one sky. */
int vw_set_sky_schedules(void* ctx, const int32_t* nseg, const int32_t* seg_tlen);

/* Autotuner: times every target on the schedule's longest segment (reps
 * timed runs each, HIP events) and keeps the fastest for each kernel; the
 * accumulators are cleared afterwards.  Needs vw_set_sky. */
int vw_autotune(void* ctx, int32_t reps, int64_t* viss_waves, int64_t* recon_waves);

/* One batched step as a graph: clears the accumulators and runs every
 * segment of the schedule (vw_reset_accumulators + vw_run_all), from a HIP
 * graph captured on the first call and replayed afterwards (recaptured after
 * vw_set_batch_targets, vw_autotune or vw_set_day).  Batched contexts only;
 * the step uses no memset and no library sort.  Follow with vw_finalize.
 * No counterpart in the reference. */
int vw_run_all_graph(void* ctx);

/* ---------------- introspection (tests) ---------------- */

/* Resident pixel directions: l, m, n, nm1 (npix each). */
int vw_get_lmn(void* ctx, double* l, double* m, double* n, double* nm1);
/* Tile cones: 5 x ntile (cx, cy, cz, cosA, sinA); img=1 for the lm-grid tiles. */
int vw_get_tiles(void* ctx, int32_t img, double* cone5);
/* lm-grid points: flat image index and direction (n_img_pts each). */
int vw_get_img_points(void* ctx, int64_t* idx, double* l, double* m, double* n);
/* Day positions (km) [3][n_epochs][n_ant]. */
int vw_get_positions(void* ctx, double* pos);
/* Current segment half rows: u, v, w (n_half each). */
int vw_get_rows(void* ctx, double* u, double* v, double* w);
/* Pair classes of the current segment, [n_half][ntile] bytes (0 mixed,
 * 1 visible, 2 hidden).  img=1 for the lm-grid tiles. */
int vw_get_tile_classes(void* ctx, int32_t img, uint8_t* cls);
/* Effective occlusion mask of half rows [row0, row0+nrows) against every
 * pixel, [nrows][npix] bytes. */
int vw_get_masks(void* ctx, int64_t row0, int64_t nrows, uint8_t* mask);
/* Number of pixels the Viss kernel summed per half row (n_half entries),
 * computed by the counting instance of the Viss kernel. */
int vw_get_viss_counts(void* ctx, int64_t* counts);
/* Current segment parameters and groups after vw_recon_segment.  Arrays
 * (may be NULL) hold nuniq keys, nuniq+1 offsets, the sorted row indices
 * (n_half rows in mirror mode, 2*n_half in full mode) and nuniq median
 * representatives idx_sorted[s0 + L/2]. */
int vw_get_groups(void* ctx, vw_seg_info* info, int64_t* keys, int32_t* offsets,
                  int32_t* idx_sorted, int32_t* rep);

/* ---------------- host utilities (no device needed) ---------------- */

/* Seeded synthetic skies: kind 0 = log-normal(0,1); kind 1 = log-normal plus
 * Poisson(100) point sources (capped at 200) with Pareto(1.5) flux; kind 2 =
 * a smooth band sky, (1 + 9 exp(-|lat| / 8 deg)) times a log-normal field
 * exp(0.5 g - 1/8) drawn per Nside-16 pixel (npix = 12 Nside^2, Nside a
 * power of two >= 16). */
int vw_make_sky(uint64_t seed, int64_t npix, int32_t kind, double* out);

/* Reference file formats (common.hpp:89-153, main_recon_online.cu:238-267,
 * 516-538).  load_sky: mode "auto" | "bin" | "txt" over
 * {sky_dir}/B_{btag}.bin, B.bin, B_{btag}.txt, B.txt (float32 / text with an
 * optional npix header line).  The used path is written to used_path
 * (capacity used_cap, may be NULL). */
int vw_io_load_sky(const char* sky_dir, const char* btag, const char* mode, int64_t npix,
                   double* out, char* used_path, int32_t used_cap);
/* C map: binary float32 (as_text=0) or one value per line (as_text=1). */
int vw_io_write_map(const char* path, const double* C, int64_t n, int32_t as_text);
/* Viss text "re im" per full row, 6 significant digits; append=1 appends. */
int vw_io_write_viss(const char* path, const vw_c128* V, int64_t n, int32_t append);
/* Three-column text at 15 significant digits (uvw / xyza / xyzb). */
int vw_io_write_cols3(const char* path, const double* a, const double* b, const double* c,
                      int64_t n, int32_t append);

/* Library version string. */
const char* vw_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VISSWS_H */
