/*
 * vw_kernels.h -- gfx950 kernels of the visibility / reconstruction path.
 *
 * Execution model (wave64, CDNA4):
 *  - Viss: a wave owns one chunk of sky tiles x RPW half rows.  Lanes hold
 *    pixels (4 per lane per 256-pixel sub-tile, coalesced SoA loads), rows
 *    are wave-uniform: row records, tile-cone classes and antenna frames
 *    arrive through scalar loads, so the all-hidden / all-visible / mixed
 *    decision is a scalar branch with no divergence.  Each lane keeps RPW
 *    complex partial sums in registers; one cross-lane tree per row per chunk
 *    writes a partial that a second kernel folds in chunk order
 *    (deterministic).  The plain form evaluates sin/cos from a 1024-entry
 *    LDS table (vw_shared.h).
 *  - Viss, antenna-factorized form (the mixed form's default): a wave owns
 *    one epoch and a chunk of tiles, lanes hold (antenna, pixel) pairs, and
 *    each step of 4 pixels feeds per-antenna exponentials to 16x16x4 matrix
 *    core products whose accumulators are the epoch's pair sums.
 *  - Recon: lanes hold output points (4 per lane, one output tile per wave),
 *    groups are wave-uniform and pass through a per-wave LDS ring; the group
 *    range of a segment is split into chunks whose partials a second kernel
 *    folds in order.
 * Workgroups are remapped so that consecutive logical blocks land on the same
 * XCD (hardware dispatches block b to XCD b % 8), keeping a chunk's pixel
 * data in one XCD's L2.
 */
#ifndef VW_KERNELS_H
#define VW_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "vw_shared.h"
#include "../../include/vissws.h"

struct RowRec {
  double u, v, w;
  int32_t ia, ib; /* antenna-frame indices g*A + a of the two endpoints */
};

/* chunk partials a fold kernel loads before adding them (in chunk order) */
#define VW_FOLD_BATCH 8

struct DevInfo {
  vw_segparams sp;
  int64_t nuniq;
  int64_t first_nz;
  int64_t wtotal;
  int64_t pad;
};

#define VW_FULLMASK 0xffffffffffffffffULL

__device__ __forceinline__ int vw_rfl(int x) { return __builtin_amdgcn_readfirstlane(x); }

/* logical block index with consecutive indices on one XCD */
__device__ __forceinline__ int64_t vw_xcd_block(int64_t b, int64_t nb) {
  int64_t q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ double vw_wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ __forceinline__ int64_t vw_wave_sum_i(int64_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

/* class byte of (tile, half row) from a [ntile][nh_pad] byte table, uniform */
__device__ __forceinline__ int vw_cls_lookup(const uint32_t* __restrict__ words, int64_t tile,
                                             int64_t nh_pad, int64_t h) {
  int64_t byte = tile * nh_pad + h;
  uint32_t w = words[byte >> 2];
  return vw_rfl((int)((w >> (8 * (byte & 3))) & 0xffu));
}

/* ------------------------------------------------------------------ */
/* resident geometry                                                    */
/* ------------------------------------------------------------------ */
__global__ void k_pix2lmn(int64_t nside, int64_t npix, double* __restrict__ l, double* __restrict__ m,
                          double* __restrict__ n, double* __restrict__ nm1) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  double a, b, c;
  vw_pix2lmn_nest(nside, p, &a, &b, &c);
  l[p] = a;
  m[p] = b;
  n[p] = c;
  nm1[p] = c - 1.0;
}

__global__ void k_grid_points(int64_t N, int64_t npts, const int64_t* __restrict__ gidx,
                              double* __restrict__ l, double* __restrict__ m, double* __restrict__ n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npts) return;
  int64_t id = gidx[k];
  double a, b, c;
  vw_grid_point(N, id / N, id % N, &a, &b, &c);
  l[k] = a;
  m[k] = b;
  n[k] = c;
}

/* Tile cone: stride-halving pairwise sum (a[i] += a[i+s], s = TP/2..1), centre,
 * minimum of c.s over the tile.  One workgroup of 256 per tile. */
template <int TP>
__global__ __launch_bounds__(256) void k_tile_cone(const double* __restrict__ l, const double* __restrict__ m,
                                                   const double* __restrict__ n, int64_t npts, int64_t ntile,
                                                   double* __restrict__ cone) {
  __shared__ double sx[TP / 2], sy[TP / 2], sz[TP / 2];
  __shared__ double smin[256];
  int64_t t = blockIdx.x;
  int64_t p0 = t * TP;
  int64_t cnt = npts - p0 < TP ? npts - p0 : TP;
  for (int i = threadIdx.x; i < TP / 2; i += 256) {
    int i2 = i + TP / 2;
    double x0 = i < cnt ? l[p0 + i] : 0.0, x1 = i2 < cnt ? l[p0 + i2] : 0.0;
    double y0 = i < cnt ? m[p0 + i] : 0.0, y1 = i2 < cnt ? m[p0 + i2] : 0.0;
    double z0 = i < cnt ? n[p0 + i] : 0.0, z1 = i2 < cnt ? n[p0 + i2] : 0.0;
    sx[i] = x0 + x1;
    sy[i] = y0 + y1;
    sz[i] = z0 + z1;
  }
  __syncthreads();
  for (int s = TP / 4; s > 0; s >>= 1) {
    for (int i = threadIdx.x; i < s; i += 256) {
      sx[i] = sx[i] + sx[i + s];
      sy[i] = sy[i] + sy[i + s];
      sz[i] = sz[i] + sz[i + s];
    }
    __syncthreads();
  }
  double cx, cy, cz;
  vw_cone_centre(sx[0], sy[0], sz[0], &cx, &cy, &cz);
  double mind = 1.0;
  for (int64_t i = threadIdx.x; i < cnt; i += 256)
    mind = fmin(mind, vw_dot3(cx, cy, cz, l[p0 + i], m[p0 + i], n[p0 + i]));
  smin[threadIdx.x] = mind;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double ca, sa;
    vw_cone_radius(smin[0], &ca, &sa);
    cone[0 * ntile + t] = cx;
    cone[1 * ntile + t] = cy;
    cone[2 * ntile + t] = cz;
    cone[3 * ntile + t] = ca;
    cone[4 * ntile + t] = sa;
  }
}

/* ------------------------------------------------------------------ */
/* day geometry and segment rows                                        */
/* ------------------------------------------------------------------ */
struct DayArgs {
  int geometry;
  int T, A;
  int64_t orbit_res, seg_len, tw, day_offset;
  int day;
  const int64_t* t1s;
  const double* circ; /* 7*A: radius, phase, rate, sr, cr, si, ci */
  double dt;
};

__global__ void k_day_positions(DayArgs a, double* __restrict__ pos) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t TA = (int64_t)a.T * a.A;
  if (k >= TA) return;
  int64_t t = k / a.A;
  int s = (int)(k - t * a.A);
  double x, y, z;
  if (a.geometry == VW_GEOM_LUNAR8) {
    int64_t orb = t / a.seg_len, loc = t - orb * a.seg_len;
    int64_t row = a.t1s[orb] + loc;
    vw_lunar_sat(row, s, a.orbit_res, a.day, a.tw, a.day_offset, &x, &y, &z);
  } else {
    const double* c = a.circ;
    int A = a.A;
    vw_circ_orbit(c[s], c[A + s], c[2 * A + s], (double)t * a.dt, c[3 * A + s], c[4 * A + s],
                  c[5 * A + s], c[6 * A + s], &x, &y, &z);
  }
  pos[k] = x;
  pos[TA + k] = y;
  pos[2 * TA + k] = z;
}

/* per (epoch, antenna) frame: px,py,pz,invn,ux,uy,uz,cosphi */
__global__ void k_seg_ant(const double* __restrict__ pos, int T, int A, int t0, int tlen, double occ_r,
                          double cosphi_glob, double* __restrict__ ant8) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)tlen * A) return;
  int64_t TA = (int64_t)T * A;
  int64_t src = (int64_t)t0 * A + k;
  double px = pos[src], py = pos[TA + src], pz = pos[2 * TA + src];
  double in, ux, uy, uz, cp;
  vw_antenna_frame(px, py, pz, occ_r, cosphi_glob, &in, &ux, &uy, &uz, &cp);
  double* q = ant8 + k * 8;
  q[0] = px; q[1] = py; q[2] = pz; q[3] = in;
  q[4] = ux; q[5] = uy; q[6] = uz; q[7] = cp;
}

__global__ void k_seg_rows(const double* __restrict__ pos, int T, int A, int P, int t0, int64_t nh,
                           const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, double lambda,
                           RowRec* __restrict__ rows, double4* __restrict__ rowt) {
  int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  int64_t TA = (int64_t)T * A;
  int64_t g = h / P;
  int j = (int)(h - g * P);
  int64_t ka = (t0 + g) * A + pa[j], kb = (t0 + g) * A + pb[j];
  RowRec r;
  r.u = vw_uvw_comp(pos[ka], pos[kb], lambda);
  r.v = vw_uvw_comp(pos[TA + ka], pos[TA + kb], lambda);
  r.w = vw_uvw_comp(pos[2 * TA + ka], pos[2 * TA + kb], lambda);
  r.ia = (int32_t)(g * A + pa[j]);
  r.ib = (int32_t)(g * A + pb[j]);
  rows[h] = r;
  /* the plain Viss phase in 1/1024 turns (exact power-of-two scaling) */
  rowt[h] = make_double4(r.u * VW_TAB_SCALE, r.v * VW_TAB_SCALE, r.w * VW_TAB_SCALE, 0.0);
}

/* reference full-row layout (orbit_online.cuh:186-231) into 9 SoA arrays */
__global__ void k_seg_full(const double* __restrict__ pos, int T, int A, int P, int t0, int64_t nh,
                           const int32_t* __restrict__ pa, const int32_t* __restrict__ pb,
                           const RowRec* __restrict__ rows, double* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t nf = 2 * nh;
  if (i >= nf) return;
  int64_t TA = (int64_t)T * A;
  int64_t g = i / (2 * P);
  int jj = (int)(i - g * 2 * P);
  bool neg = jj >= P;
  int j = neg ? jj - P : jj;
  int64_t h = g * P + j;
  int64_t ka = (t0 + g) * A + pa[j], kb = (t0 + g) * A + pb[j];
  int64_t k1 = neg ? kb : ka, k2 = neg ? ka : kb;
  RowRec r = rows[h];
  out[0 * nf + i] = neg ? -r.u : r.u;
  out[1 * nf + i] = neg ? -r.v : r.v;
  out[2 * nf + i] = neg ? -r.w : r.w;
  out[3 * nf + i] = pos[k1];
  out[4 * nf + i] = pos[TA + k1];
  out[5 * nf + i] = pos[2 * TA + k1];
  out[6 * nf + i] = pos[k2];
  out[7 * nf + i] = pos[TA + k2];
  out[8 * nf + i] = pos[2 * TA + k2];
}

/* cone bits of every antenna frame against tiles [tile0, tile0+nt) ->
 * bytes [nt][nfr_pad] (bit0 all visible, bit1 all hidden); cone arrays hold
 * all ntile tiles.  A row's class only combines the bits of its two frames,
 * so the cone tests run once per (tile, frame) instead of twice per
 * (tile, row): tlen*A frames against tlen*A(A-1)/2 rows. */
__global__ void k_ant_bits(const double* __restrict__ cone, int64_t ntile, int64_t tile0, int64_t nt,
                           const double* __restrict__ ant8, int64_t nfr, int64_t nfr_pad, uint8_t* __restrict__ bits) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nt * nfr_pad) return;
  int64_t tr = k / nfr_pad, f = k - tr * nfr_pad;
  if (f >= nfr) { bits[k] = 0; return; }
  int64_t t = tile0 + tr;
  const double* q = ant8 + f * 8;
  bits[k] = (uint8_t)vw_cone_bits(q[4], q[5], q[6], cone[t], cone[ntile + t], cone[2 * ntile + t],
                                  cone[3 * ntile + t], cone[4 * ntile + t], q[7]);
}

/* the cone bits of tile t's nfr frames (as k_ant_bits) into LDS */
__device__ __forceinline__ void vw_tile_bits_lds(uint8_t* sb, const double* __restrict__ cone, int64_t ntile,
                                                 int64_t t, const double* __restrict__ ant8, int64_t nfr) {
  const double cx = cone[t], cy = cone[ntile + t], cz = cone[2 * ntile + t];
  const double cA = cone[3 * ntile + t], sA = cone[4 * ntile + t];
  for (int64_t f = threadIdx.x; f < nfr; f += blockDim.x) {
    const double* q = ant8 + f * 8;
    sb[f] = (uint8_t)vw_cone_bits(q[4], q[5], q[6], cx, cy, cz, cA, sA, q[7]);
  }
}

/* pair class per (tile, half row) for the tiles of a pass -> bytes
 * [tiles][nh_pad] (rows past nh: hidden), from the frame bits.  A block
 * covers VW_GCLS_BLOCK rows of one tile, four a thread written as one class
 * word.  The rows' frame indices pass through LDS (loaded with consecutive
 * lanes on consecutive records, read back four rows a thread), beside the
 * tile's row of cone bits staged in LDS (STAGE; nfr_pad bytes of dynamic
 * LDS) or, for longer rows, gathered from global memory, as k_group_cls. */
#define VW_GCLS_BLOCK 1024
template <bool STAGE>
__global__ __launch_bounds__(256) void k_classify(const uint8_t* __restrict__ bits, int64_t nfr_pad, int nhb,
                                                  const RowRec* __restrict__ rows, int64_t nh, int64_t nh_pad,
                                                  uint32_t* __restrict__ cls) {
  extern __shared__ uint4 sbits[];
  __shared__ int4 sab[VW_GCLS_BLOCK / 2]; /* (ia, ib) of rows 2 k and 2 k + 1 */
  static_assert(VW_GCLS_BLOCK == 4 * 256, "four entries a thread");
  const uint32_t tr = blockIdx.x / (uint32_t)nhb, hc = blockIdx.x - tr * (uint32_t)nhb;
  const int64_t hb = (int64_t)hc * VW_GCLS_BLOCK;
  const uint8_t* row = bits + (int64_t)tr * nfr_pad;
  /* rows past nh are not frames: frame 0 stands in (the row array holds at
   * least one record) */
  int2* sab2 = (int2*)sab;
  int2 ab[4];
  int keep[4]; /* all ones for a row, zero past nh (a mask, not a select, so
                * the four loads stay unconditional and in flight together) */
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t h = hb + threadIdx.x + 256 * k;
    keep[k] = -(int)(h < nh);
    ab[k] = *(const int2*)&rows[h < nh ? h : 0].ia;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) sab2[threadIdx.x + 256 * k] = make_int2(ab[k].x & keep[k], ab[k].y & keep[k]);
  if (STAGE) {
    /* nfr_pad is a multiple of 64 and the row starts 64-byte aligned */
    for (int64_t i = threadIdx.x; i < nfr_pad / 16; i += 256) sbits[i] = ((const uint4*)row)[i];
  }
  __syncthreads();
  const int64_t h0 = hb + 4 * threadIdx.x;
  if (h0 >= nh_pad) return; /* nh_pad is a multiple of 64: whole words only */
  const int4 p0 = sab[2 * threadIdx.x], p1 = sab[2 * threadIdx.x + 1];
  const int av[4] = {p0.x, p0.z, p1.x, p1.z}, bv[4] = {p0.y, p0.w, p1.y, p1.w};
  const uint8_t* b = STAGE ? (const uint8_t*)sbits : row;
  uint32_t w = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t c = (uint32_t)vw_pair_class(b[av[e]], b[bv[e]]);
    w |= (h0 + e < nh ? c : (uint32_t)VW_CLS_HIDDEN) << (8 * e);
  }
  cls[((int64_t)tr * nh_pad + h0) >> 2] = w;
}

/* k_classify with the cone bits made in place: one block per tile of the
 * pass computes the tile's frame bits into LDS (nfr bytes of dynamic LDS),
 * then walks the tile's rows VW_GCLS_BLOCK at a time as k_classify does, so
 * no bit table passes through global memory and each cone test runs once */
__global__ __launch_bounds__(256) void k_classify_tile(const double* __restrict__ cone, int64_t ntile,
                                                       int64_t tile0, const double* __restrict__ ant8, int64_t nfr,
                                                       const RowRec* __restrict__ rows, int64_t nh, int64_t nh_pad,
                                                       uint32_t* __restrict__ cls) {
  extern __shared__ uint4 sbits[];
  __shared__ int4 sab[VW_GCLS_BLOCK / 2]; /* (ia, ib) of rows 2 k and 2 k + 1 */
  const int64_t tr = blockIdx.x;
  uint8_t* sb = (uint8_t*)sbits;
  vw_tile_bits_lds(sb, cone, ntile, tile0 + tr, ant8, nfr);
  int2* sab2 = (int2*)sab;
  for (int64_t hb = 0; hb < nh_pad; hb += VW_GCLS_BLOCK) {
    int2 ab[4];
    int keep[4]; /* a mask, not a select: the four loads stay in flight together */
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int64_t h = hb + threadIdx.x + 256 * k;
      keep[k] = -(int)(h < nh);
      ab[k] = *(const int2*)&rows[h < nh ? h : 0].ia;
    }
    __syncthreads(); /* the bits are in; the last pass's reads of sab are done */
#pragma unroll
    for (int k = 0; k < 4; ++k) sab2[threadIdx.x + 256 * k] = make_int2(ab[k].x & keep[k], ab[k].y & keep[k]);
    __syncthreads();
    const int64_t h0 = hb + 4 * threadIdx.x;
    if (h0 < nh_pad) {
      const int4 p0 = sab[2 * threadIdx.x], p1 = sab[2 * threadIdx.x + 1];
      const int av[4] = {p0.x, p0.z, p1.x, p1.z}, bv[4] = {p0.y, p0.w, p1.y, p1.w};
      uint32_t w = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t c = (uint32_t)vw_pair_class(sb[av[e]], sb[bv[e]]);
        w |= (h0 + e < nh ? c : (uint32_t)VW_CLS_HIDDEN) << (8 * e);
      }
      cls[(tr * nh_pad + h0) >> 2] = w;
    }
  }
}

/* effective mask of rows [row0, row0+nrows) x pixels (introspection) */
__global__ void k_masks(const double* __restrict__ l, const double* __restrict__ m, const double* __restrict__ n,
                        int64_t npix, int64_t TP, const uint8_t* __restrict__ cls, int64_t nh_pad,
                        const RowRec* __restrict__ rows, const double* __restrict__ ant8, int64_t row0,
                        int64_t nrows, uint8_t* __restrict__ mask) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrows * npix) return;
  int64_t r = k / npix, p = k - r * npix;
  int64_t h = row0 + r;
  int c = cls[(p / TP) * nh_pad + h];
  uint8_t v;
  if (c == VW_CLS_VISIBLE) v = 1;
  else if (c == VW_CLS_HIDDEN) v = 0;
  else {
    RowRec rr = rows[h];
    const double* qa = ant8 + (int64_t)rr.ia * 8;
    const double* qb = ant8 + (int64_t)rr.ib * 8;
    v = (uint8_t)(vw_visible(l[p], m[p], n[p], qa[0], qa[1], qa[2], qa[3], qa[7]) &&
                  vw_visible(l[p], m[p], n[p], qb[0], qb[1], qb[2], qb[3], qb[7]));
  }
  mask[k] = v;
}

/* ------------------------------------------------------------------ */
/* Viss                                                                 */
/* ------------------------------------------------------------------ */
struct VissArgs {
  const double2* tab; /* the sine table (plain forms), built once per context */
  const double *l, *m, *n, *nm1, *B;
  const float* Bf;
  int64_t npix;
  int64_t tile0, ntile; /* tiles of this pass: [tile0, tile0 + ntile) */
  const RowRec* rows;
  const double4* rowt; /* (u, v, w) * 1024 per half row (plain form) */
  const double* ant8;
  const uint32_t* cls; /* [ntile][nh_pad] bytes (pass-relative) viewed as words */
  int64_t nh, nh_pad;
  int64_t tpc, nchunk, nrb, bpc;
  double2* part; /* [nchunk][nh_pad] */
  int64_t* cpart; /* [nchunk][nh_pad] (counting instance) */
};

/* the sine table: entry k = (cos, sin)(2 pi k / N), as vw_sincos_tab_build;
 * built once per context in device memory, copied into LDS by every
 * workgroup of the plain kernels (four 16-byte loads a thread) */
__global__ void k_tab_build(double2* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= VW_TAB_N) return;
  double e[2]; /* (cos, sin) */
  vw_sincos_turn((double)i / (double)VW_TAB_N, &e[1], &e[0]);
  tab[i] = make_double2(e[0], e[1]);
}
__device__ __forceinline__ void vw_tab_fill(double2* tab, const double2* __restrict__ src) {
  for (int i = threadIdx.x; i < VW_TAB_N; i += blockDim.x) tab[i] = src[i];
}

__device__ __forceinline__ void vw_hw_sincos_turn(float r, float* s, float* c) {
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}

/* occlusion of the PPL pixels of a lane against both endpoints of a row: the
 * plain kernels evaluate both tests without short circuit (straight-line
 * code, no exec-mask branches); the mixed kernels, at their tighter register
 * budget, keep the short-circuit form */
template <bool BRANCHFREE>
__device__ __forceinline__ bool vw_vis2(bool ok, double l, double m, double n, double ax, double ay, double az,
                                        double ai, double ac, double bx, double by, double bz, double bi, double bc) {
  if (BRANCHFREE)
    return ((int)ok & vw_visible(l, m, n, ax, ay, az, ai, ac) & vw_visible(l, m, n, bx, by, bz, bi, bc)) != 0;
  return ok && vw_visible(l, m, n, ax, ay, az, ai, ac) && vw_visible(l, m, n, bx, by, bz, bi, bc);
}

/* mixed form, one term: accumulate bf * exp(-j 2 pi x) in FP32 (the sign of
 * the imaginary part is applied when the partial is written) */
__device__ __forceinline__ void vw_viss_term_mixed(const RowRec& rec, double l, double m, double nm1, float bf,
                                                   float& fre, float& fim) {
  const double x = vw_viss_phase(rec.u, rec.v, rec.w, l, m, nm1);
  const float f = (float)__builtin_amdgcn_fract(x);
  float s, c;
  vw_hw_sincos_turn(f, &s, &c);
  fre = fmaf(bf, c, fre);
  fim = fmaf(bf, s, fim);
}

/* plain form, the PPL terms of one lane for one half row: phases (operands
 * pre-scaled by 1024), all table indices, the table reads, then polynomials
 * and rotations, so the LDS reads overlap the arithmetic */
/* The pixel weight bw is folded into the polynomial coefficients once per
 * sub-tile: w0 = bw, w1 = bw a, w2 = bw (-a^2/2), w3 = bw (-a^3/6) give
 *   bw cy = w0 + w2 z,  bw sy = f (w1 + w3 z),
 * and the rotation by the table entry (ec, es) is accumulated directly:
 *   dre += ec (bw cy) - es (bw sy),  dim += es (bw cy) + ec (bw sy),
 * four FMAs, two products fewer per term than rotating, then weighting.
 * MASK: a pixel with vis[k] false adds nothing (selects on bw cy, bw sy). */
template <int PPL, bool MASK>
__device__ __forceinline__ void vw_viss_terms_plain(const double4& rt, const double* pl, const double* pm,
                                                    const double* pz, const double* w0, const double* w1,
                                                    const double* w2, const double* w3, const bool* vis,
                                                    const double2* tab, double& dre, double& dim) {
  double f[PPL];
  int ix[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) f[k] = vw_tab_split(vw_viss_phase(rt.x, rt.y, rt.z, pl[k], pm[k], pz[k]), &ix[k]);
  double2 e[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) e[k] = tab[ix[k]];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    const double z = f[k] * f[k];
    double cw = fma(z, w2[k], w0[k]);
    double sw = f[k] * fma(z, w3[k], w1[k]);
    if (MASK) {
      cw = vis[k] ? cw : 0.0;
      sw = vis[k] ? sw : 0.0;
    }
    dre = fma(-e[k].y, sw, fma(e[k].x, cw, dre));
    dim = fma(e[k].x, sw, fma(e[k].y, cw, dim));
  }
}

/* occupancy: the plain form keeps 8 FP64 row sums and 4 weight
 * coefficients per pixel per lane at 3 waves per SIMD; the mixed form (16
 * FP32 row sums) runs 5 (measured best) */
#ifndef VW_VISS_WAVES_PLAIN
#define VW_VISS_WAVES_PLAIN 3
#endif
#ifndef VW_VISS_WAVES_MIXED
#define VW_VISS_WAVES_MIXED 5
#endif
/* per-row Viss, mixed form (the plain form is k_viss_plain) */
template <int TP, int RPW, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VW_VISS_WAVES_MIXED))) void k_viss(VissArgs a) {
  /* a tile is walked in sub-tiles of 256 pixels (4 per lane), so the
   * register footprint does not grow with the tile size */
  constexpr int PPL = 4;
  constexpr int NSUB = TP / 256;
  static_assert(RPW % 4 == 0, "class words hold 4 rows");
  const int64_t vb = vw_xcd_block(blockIdx.x, gridDim.x);
  const int64_t chunk = vb / a.bpc;
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int64_t rb = (vb - chunk * a.bpc) * 4 + wave;
  if (chunk >= a.nchunk || rb >= a.nrb) return;
  const int64_t r0 = rb * RPW; /* r0 + RPW <= nh_pad: nh_pad is a multiple of 64 */
  const int64_t t0 = a.tile0 + chunk * a.tpc;
  const int64_t tend = a.tile0 + a.ntile;
  const int64_t t1 = (t0 + a.tpc < tend) ? t0 + a.tpc : tend;

  float fre[RPW], fim[RPW];
  int cnt[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) { fre[r] = 0.f; fim[r] = 0.f; cnt[r] = 0; }

  for (int64_t st = t0 * NSUB; st < t1 * NSUB; ++st) {
    const int64_t tile = st / NSUB;
    /* the class bytes of this wave's RPW rows for the tile: RPW/4 uniform words */
    uint32_t cw[RPW / 4];
    const uint32_t* cwp = a.cls + (((tile - a.tile0) * a.nh_pad + r0) >> 2);
#pragma unroll
    for (int j = 0; j < RPW / 4; ++j) cw[j] = (uint32_t)vw_rfl((int)cwp[j]);
    double pl[PPL], pm[PPL], pn[PPL], pz[PPL];
    float pbf[PPL];
    bool ok[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      int64_t p = st * 256 + lane + 64 * k;
      ok[k] = p < a.npix;
      int64_t pp = ok[k] ? p : 0;
      pl[k] = a.l[pp];
      pm[k] = a.m[pp];
      pn[k] = a.n[pp];
      pz[k] = a.nm1[pp];
      /* weights are zero outside the sky, so the visible path needs no mask */
      pbf[k] = ok[k] ? a.Bf[pp] : 0.0f;
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const int64_t h = r0 + r;
      /* rows past nh carry the hidden class */
      const int cls = (int)((cw[r >> 2] >> (8 * (r & 3))) & 0xffu);
      if (cls == VW_CLS_VISIBLE) {
        const RowRec rec = a.rows[h];
#pragma unroll
        for (int k = 0; k < PPL; ++k) vw_viss_term_mixed(rec, pl[k], pm[k], pz[k], pbf[k], fre[r], fim[r]);
        if (COUNT) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) cnt[r] += ok[k] ? 1 : 0;
        }
      } else if (cls == VW_CLS_MIXED) {
        const RowRec rec = a.rows[h];
        const double* qa = a.ant8 + (int64_t)rec.ia * 8;
        const double* qb = a.ant8 + (int64_t)rec.ib * 8;
        const double ax = qa[0], ay = qa[1], az = qa[2], ai = qa[3], ac = qa[7];
        const double bx = qb[0], by = qb[1], bz = qb[2], bi = qb[3], bc = qb[7];
        bool vis[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
          vis[k] = vw_vis2<true>(ok[k], pl[k], pm[k], pn[k], ax, ay, az, ai, ac, bx, by, bz, bi, bc);
#pragma unroll
        for (int k = 0; k < PPL; ++k)
          vw_viss_term_mixed(rec, pl[k], pm[k], pz[k], vis[k] ? pbf[k] : 0.f, fre[r], fim[r]);
        if (COUNT) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) cnt[r] += vis[k] ? 1 : 0;
        }
      }
    }
  }
  /* one cross-lane tree per row; FP64 above the per-lane FP32 blocks */
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int64_t h = r0 + r;
    double re = vw_wave_sum((double)fre[r]);
    double im = vw_wave_sum((double)fim[r]);
    int64_t cc = 0;
    if (COUNT) cc = vw_wave_sum_i((int64_t)cnt[r]);
    if (lane == 0 && h < a.nh) {
      a.part[chunk * a.nh_pad + h] = make_double2(re, -im);
      if (COUNT) a.cpart[chunk * a.nh_pad + h] = cc;
    }
  }
}

/* Plain per-row Viss with the pixel data prefetched into LDS.  Each wave
 * keeps its own LDS copy of the next sub-tile's l, m, n and weights, written
 * by direct global-to-LDS loads (16 bytes a lane: two neighbouring pixels)
 * while the current sub-tile is summed, together with the next class words;
 * a sub-tile then starts with LDS reads, and no registers hold the data in
 * flight.  Nothing else in the sub-tile loop reads global memory, so the
 * loop never waits on the prefetch: the wave's rows are read once up front
 * ((u, v, w) into scalar registers, the antenna data of the occlusion
 * predicates into LDS).  Lane i owns pixels 128 j + 2 i + e (j, e in {0, 1})
 * of a 256-pixel sub-tile; pixel counts are even, so a pair is in the sky or
 * out of it whole, and pairs outside it read pixels 0 and 1 under a zero
 * weight.  n - 1 is formed as in k_pix2lmn.  Terms, masks and partials as
 * k_viss. */
template <int TP, int RPW, bool COUNT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VW_VISS_WAVES_PLAIN))) void k_viss_plain(
    VissArgs a) {
  constexpr int PPL = 4;
  constexpr int SUB = 64 * PPL;
  constexpr int NSUB = TP / SUB;
  constexpr int NCW = RPW / 4; /* class words per sub-tile */
  static_assert(RPW % 4 == 0 && NCW <= 64, "class words hold 4 rows");
  static_assert(TP % SUB == 0, "tiles hold whole sub-tiles");
  __shared__ double2 tab[VW_TAB_N];
  __shared__ double2 pre[4][4][2][64]; /* [wave][l, m, n, B][half][lane] */
  __shared__ double rant[4][RPW][10];  /* [wave][row][antenna a (x y z i c), b (...)] */
  vw_tab_fill(tab, a.tab);
  __syncthreads();
  const int64_t vb = vw_xcd_block(blockIdx.x, gridDim.x);
  const int64_t chunk = vb / a.bpc;
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int64_t rb = (vb - chunk * a.bpc) * 4 + wave;
  if (chunk >= a.nchunk || rb >= a.nrb) return;
  const int64_t r0 = rb * RPW; /* r0 + RPW <= nh_pad: nh_pad is a multiple of 64 */
  const int64_t t0 = a.tile0 + chunk * a.tpc;
  const int64_t tend = a.tile0 + a.ntile;
  const int64_t t1 = (t0 + a.tpc < tend) ? t0 + a.tpc : tend;
  const int64_t s0 = t0 * NSUB, s1 = t1 * NSUB;

  /* the wave's rows; rows past nh are hidden and never summed */
  double4 rt[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) rt[r] = (r0 + r < a.nh) ? a.rowt[r0 + r] : make_double4(0.0, 0.0, 0.0, 0.0);
  if (lane < RPW && r0 + lane < a.nh) {
    const RowRec rec = a.rows[r0 + lane];
    const double* qa = a.ant8 + (int64_t)rec.ia * 8;
    const double* qb = a.ant8 + (int64_t)rec.ib * 8;
    double* d = rant[wave][lane];
    d[0] = qa[0]; d[1] = qa[1]; d[2] = qa[2]; d[3] = qa[3]; d[4] = qa[7];
    d[5] = qb[0]; d[6] = qb[1]; d[7] = qb[2]; d[8] = qb[3]; d[9] = qb[7];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  typedef __attribute__((address_space(3))) void* lds_ptr;
  uint32_t cwn = 0; /* lane j < NCW: class word j of the prefetched sub-tile */
  auto prefetch = [&](int64_t s) {
    const double* src[4] = {a.l, a.m, a.n, a.B};
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t p = s * SUB + 128 * j + 2 * lane;
      const int64_t pp = p < a.npix ? p : 0;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        __builtin_amdgcn_global_load_lds((const void*)(src[q] + pp), (lds_ptr)&pre[wave][q][j][0], 16, 0, 0);
    }
    const uint32_t* cwp = a.cls + ((((s / NSUB) - a.tile0) * a.nh_pad + r0) >> 2);
    cwn = cwp[lane < NCW ? lane : 0];
  };

  double dre[RPW], dim[RPW];
  int cnt[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) { dre[r] = 0.0; dim[r] = 0.0; cnt[r] = 0; }

  prefetch(s0);
  for (int64_t st = s0; st < s1; ++st) {
    /* this sub-tile's prefetch has landed */
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    uint32_t cw[NCW];
#pragma unroll
    for (int j = 0; j < NCW; ++j) cw[j] = (uint32_t)__builtin_amdgcn_readlane((int)cwn, j);
    double pl[PPL], pm[PPL], pn[PPL], pz[PPL], w0[PPL], w1[PPL], w2[PPL], w3[PPL];
    bool ok[PPL];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double2 L = pre[wave][0][j][lane], M = pre[wave][1][j][lane];
      const double2 N = pre[wave][2][j][lane], W = pre[wave][3][j][lane];
      pl[2 * j] = L.x;
      pl[2 * j + 1] = L.y;
      pm[2 * j] = M.x;
      pm[2 * j + 1] = M.y;
      pn[2 * j] = N.x;
      pn[2 * j + 1] = N.y;
      const bool okj = st * SUB + 128 * j + 2 * lane < a.npix;
      ok[2 * j] = okj;
      ok[2 * j + 1] = okj;
      w0[2 * j] = okj ? W.x : 0.0;
      w0[2 * j + 1] = okj ? W.y : 0.0;
    }
    /* the reads are done before the buffer is refilled */
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    if (st + 1 < s1) prefetch(st + 1);
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      pz[k] = pn[k] - 1.0;
      w1[k] = w0[k] * VW_SHORT_S1;
      w2[k] = w0[k] * VW_SHORT_C2;
      w3[k] = w0[k] * VW_SHORT_S3;
    }
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      /* rows past nh carry the hidden class */
      const int cls = (int)((cw[r >> 2] >> (8 * (r & 3))) & 0xffu);
      if (cls == VW_CLS_VISIBLE) {
        vw_viss_terms_plain<PPL, false>(rt[r], pl, pm, pz, w0, w1, w2, w3, ok, tab, dre[r], dim[r]);
        if (COUNT) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) cnt[r] += ok[k] ? 1 : 0;
        }
      } else if (cls == VW_CLS_MIXED) {
        const double* d = rant[wave][r];
        const double ax = d[0], ay = d[1], az = d[2], ai = d[3], ac = d[4];
        const double bx = d[5], by = d[6], bz = d[7], bi = d[8], bc = d[9];
        bool vis[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k)
          vis[k] = vw_vis2<true>(ok[k], pl[k], pm[k], pn[k], ax, ay, az, ai, ac, bx, by, bz, bi, bc);
        vw_viss_terms_plain<PPL, true>(rt[r], pl, pm, pz, w0, w1, w2, w3, vis, tab, dre[r], dim[r]);
        if (COUNT) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) cnt[r] += vis[k] ? 1 : 0;
        }
      }
    }
  }
  /* one cross-lane tree per row */
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int64_t h = r0 + r;
    const double re = vw_wave_sum(dre[r]);
    const double im = vw_wave_sum(dim[r]);
    int64_t cc = 0;
    if (COUNT) cc = vw_wave_sum_i((int64_t)cnt[r]);
    if (lane == 0 && h < a.nh) {
      a.part[chunk * a.nh_pad + h] = make_double2(re, -im);
      if (COUNT) a.cpart[chunk * a.nh_pad + h] = cc;
    }
  }
}

/* adjoint of the plain Viss sum (a real-linear map from pixel weights to
 * raw half visibilities): for one tile pass and one chunk of half rows,
 * g_p = sum_h M_hp Re(r_h exp(+j 2 pi (u l + v m + w (n-1)))) per pixel,
 * with the effective mask M of the Viss (tile class, then the per-pixel
 * predicate on partly visible tiles) and its table sincos.  A wave owns one
 * 256-pixel sub-tile (4 pixels per lane) and a chunk of rows; rows are
 * wave-uniform, their class one byte of a uniform class word.  Partials
 * [chunk][pixel of the pass]; the fold applies the pixel area. */
struct VissAdjArgs {
  const double2* tab; /* the sine table (plain forms), built once per context */
  const double *l, *m, *n, *nm1;
  int64_t npix;
  int64_t tile0, ntile; /* tiles of this pass: [tile0, tile0 + ntile) */
  const RowRec* rows;
  const double4* rowt; /* (u, v, w) * 1024 per half row */
  const double* ant8;
  const uint32_t* cls; /* [ntile][nh_pad] bytes (pass-relative) viewed as words */
  int64_t nh, nh_pad;
  const double2* res;  /* [nh] residual half visibilities (raw convention) */
  int64_t rpc, nchunk; /* rows per chunk (a multiple of 4), chunks */
  int64_t nsub;        /* 256-pixel sub-tiles of the pass */
  double* part;        /* [nchunk][nsub * 256] */
};

template <int TP>
__global__ __launch_bounds__(256) void k_viss_adj(VissAdjArgs a) {
  constexpr int PPL = 4;
  __shared__ double2 tab[VW_TAB_N];
  vw_tab_fill(tab, a.tab);
  __syncthreads();
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int64_t wv = (int64_t)blockIdx.x * 4 + wave;
  const int64_t chunk = wv / a.nsub, sub = wv - chunk * a.nsub;
  if (chunk >= a.nchunk) return;
  const int64_t p0 = a.tile0 * TP + sub * 256; /* first pixel of the sub-tile */
  const int64_t tile = p0 / TP;
  const int64_t h0 = chunk * a.rpc;
  const int64_t h1 = h0 + a.rpc < a.nh ? h0 + a.rpc : a.nh;
  double pl[PPL], pm[PPL], pn[PPL], pz[PPL], acc[PPL];
  bool ok[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    const int64_t p = p0 + lane + 64 * k;
    ok[k] = p < a.npix;
    const int64_t pp = ok[k] ? p : 0;
    pl[k] = a.l[pp];
    pm[k] = a.m[pp];
    pn[k] = a.n[pp];
    pz[k] = a.nm1[pp];
    acc[k] = 0.0;
  }
  const uint32_t* cwp = a.cls + (((tile - a.tile0) * a.nh_pad) >> 2);
  for (int64_t h4 = h0; h4 < h1; h4 += 4) {
    const uint32_t cw = (uint32_t)vw_rfl((int)cwp[h4 >> 2]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t h = h4 + j;
      const int cls = (int)((cw >> (8 * j)) & 0xffu);
      if (h >= h1 || cls == VW_CLS_HIDDEN) continue;
      const double2 r = a.res[h];
      bool mk[PPL];
      if (cls == VW_CLS_VISIBLE) {
#pragma unroll
        for (int k = 0; k < PPL; ++k) mk[k] = ok[k];
      } else {
        const RowRec rec = a.rows[h];
        const double* qa = a.ant8 + (int64_t)rec.ia * 8;
        const double* qb = a.ant8 + (int64_t)rec.ib * 8;
#pragma unroll
        for (int k = 0; k < PPL; ++k)
          mk[k] = vw_vis2<true>(ok[k], pl[k], pm[k], pn[k], qa[0], qa[1], qa[2], qa[3], qa[7], qb[0], qb[1], qb[2],
                                qb[3], qb[7]);
      }
      const double4 rt = a.rowt[h];
      double f[PPL];
      int ix[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) f[k] = vw_tab_split(vw_viss_phase(rt.x, rt.y, rt.z, pl[k], pm[k], pz[k]), &ix[k]);
      double2 e[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) e[k] = tab[ix[k]];
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        double sy, cy, sn, cs;
        vw_tab_poly_short(f[k], &sy, &cy);
        vw_tab_rotate(e[k].x, e[k].y, sy, cy, &sn, &cs);
        /* Re(r exp(+j x)) = r.x cos x - r.y sin x */
        const double t = fma(r.x, cs, fma(-r.y, sn, acc[k]));
        acc[k] = mk[k] ? t : acc[k];
      }
    }
  }
  const int64_t stride = a.nsub * 256;
#pragma unroll
  for (int k = 0; k < PPL; ++k) a.part[chunk * stride + sub * 256 + lane + 64 * k] = acc[k];
}

/* adjoint partials of a pass -> g[p] = area * sum over chunks (ascending) */
__global__ void k_viss_adj_fold(const double* __restrict__ part, int64_t nchunk, int64_t stride, int64_t p0,
                                int64_t np, double area, double* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  double s = 0.0;
  for (int64_t k = 0; k < nchunk; ++k) s = s + part[k * stride + i];
  g[p0 + i] = area * s;
}

/* partials -> raw half sums, chunks folded in ascending order; passes after
 * the first add to the running sums */
__global__ void k_viss_reduce(const double2* __restrict__ part, int64_t nchunk, int64_t nh, int64_t nh_pad,
                              int accumulate, double2* __restrict__ vraw, const int64_t* __restrict__ cpart,
                              int64_t* __restrict__ counts) {
  int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  double re = 0.0, im = 0.0;
  int64_t c = 0;
  /* VW_FOLD_BATCH chunks' loads in flight at a time; the sums keep chunk order */
  int64_t k = 0;
  for (; k + VW_FOLD_BATCH <= nchunk; k += VW_FOLD_BATCH) {
    double2 v[VW_FOLD_BATCH];
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) v[e] = part[(k + e) * nh_pad + h];
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) {
      re += v[e].x;
      im += v[e].y;
    }
    if (counts) {
#pragma unroll
      for (int e = 0; e < VW_FOLD_BATCH; ++e) c += cpart[(k + e) * nh_pad + h];
    }
  }
  for (; k < nchunk; ++k) {
    double2 v = part[k * nh_pad + h];
    re += v.x;
    im += v.y;
    if (counts) c += cpart[k * nh_pad + h];
  }
  if (accumulate) {
    double2 o = vraw[h];
    re = o.x + re;
    im = o.y + im;
    if (counts) c += counts[h];
  }
  vraw[h] = make_double2(re, im);
  if (counts) counts[h] = c;
}

/* ------------------------------------------------------------------ */
/* Viss, antenna-factorized form on the matrix cores                    */
/* ------------------------------------------------------------------ */
/* The phase of half row (a, b) is a difference of antenna phases,
 *   u l + v m + w (n-1) = psi_b - psi_a,  psi_x = (x - x_0).(l, m, n-1) 1e3/lambda
 * (x_0: the epoch's antenna 0, which cancels), and its mask is a product of
 * per-antenna masks (below).  With e_x = exp(+2 pi j psi_x) = c_x + j s_x one
 * epoch's half visibilities are entries of
 *   V[a][b] = sum_p I_p m_a m_b e_a conj(e_b):
 *   Re V = Y11[a][b] + Y22[a][b],  Im V = Y21[a][b] - Y21[b][a],
 *   Y11 = sum I c c^T,  Y22 = sum I s s^T,  Y21 = sum I s c^T,
 * three real (16 x 4) x (4 x 16) products per 4 pixels and 16 antennas on
 * the FP64 matrix cores (FP32 in the mixed form), where the row form
 * evaluates 120 phase / sin / cos terms per pixel one by one.  Lane l
 * computes the exponential of antenna 16 gi + l % 16 at pixel l / 16 of the
 * step, which is at once its A operand (row, k) and its B operand (k, column)
 * element.  Accumulator layout (measured on gfx950): FP64 lane l register r
 * holds D[l/16 + 4r][l%16]; FP32 D[4(l/16) + r][l%16].
 *
 * Masks.  The mask of row (a, b) on a pixel is 1 when both endpoint cones
 * hold the tile (class visible), 0 when either cone misses it (hidden), else
 * the product of the two per-pixel predicates.  In a tile where no antenna
 * is partly visible the per-antenna mask m = hidden ? 0 : 1 gives exactly
 * that; otherwise m = hidden ? 0 : predicate does, unless the predicate of an
 * all-visible antenna fails on a pixel of the tile (the cone bound is
 * conservative, so only rounding at the horizon can do that), which would
 * drop terms of pairs of two all-visible antennas.  The summing pass flags
 * that case; the tile then gets a correction pass adding, for those pairs,
 *   Y += sum_p I_p (o_a o_b - o_a p_a o_b p_b) (...)   (o: all-visible),
 * two more products per block on the same accumulators. */
typedef double vw_d4 __attribute__((ext_vector_type(4)));
typedef float vw_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ vw_d4 vw_mfma(double a, double b, vw_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ vw_f4 vw_mfma(float a, float b, vw_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

struct VissFArgs {
  const double2* tab; /* the sine table (plain forms), built once per context */
  const double *l, *m, *n, *nm1, *B;
  const float* Bf;
  int64_t npix;
  int64_t tile0, ntile;   /* tiles of this pass: [tile0, tile0 + ntile) */
  const double4* aph;     /* [tlen*A] (x - x_0) * 1e3 / lambda * 1024 per antenna frame */
  const double* ant8;     /* [tlen*A][8] antenna frames */
  const uint8_t* bits;    /* [ntile][nfr_pad] cone bits, pass-relative tiles */
  int64_t nfr_pad;
  int32_t A, tlen;
  int64_t tpc, nchunk, nquad;
  int32_t check;          /* 1: every tile with a partly visible antenna takes the two-mask sum */
  double* part;           /* [tlen][nchunk][NB][256] */
};

/* blocks of one wave's accumulators for G groups of 16 antennas: Y11 and Y22
 * for gi >= gj (symmetric), Y21 for all (gi, gj) */
template <int G> struct VfShape {
  static constexpr int NS = G * (G + 1) / 2;
  static constexpr int NB = NS + G * G; /* partial blocks: Re = Y11 + Y22 (NS), Y21 (G*G) */
};

/* one tile's pixel data staged in LDS for the four waves of a workgroup
 * (weights zero past the end of the sky) */
template <int TP> struct VfTileLds {
  double l[TP], m[TP], n[TP], nm1[TP], w[TP];
};

template <int G> struct VfLane {
  double kx[G], ky[G], kz[G];                 /* phase coefficients, 1/1024 turns */
  double px[G], py[G], pz[G], pinv[G], pcos[G]; /* predicate frame */
};

/* the product-mask sum of one tile.  PRED false (no partly visible
 * antenna in the tile): m = hidden ? 0 : 1.  PRED: m = hidden ? 0 :
 * predicate, and the return value is nonzero when the predicate of an
 * all-visible antenna failed on some pixel of the tile.  CORR:
 * the correction pass instead, for pairs of two all-visible antennas
 * (o = all-visible): Y += sum w (o_a o_b - o_a p_a o_b p_b) (...). */
template <int TP, int G, bool MIXED, bool PRED, bool CORR>
__device__ __forceinline__ uint64_t vf_tile(const VfTileLds<TP>& px, const VfLane<G>& L, const int (&bt)[G],
                                            const double2* tab,
                                            typename std::conditional<MIXED, vw_f4, vw_d4>::type (&y11)[VfShape<G>::NS],
                                            typename std::conditional<MIXED, vw_f4, vw_d4>::type (&y22)[VfShape<G>::NS],
                                            typename std::conditional<MIXED, vw_f4, vw_d4>::type (&y21)[G * G]) {
  using T = typename std::conditional<MIXED, float, double>::type;
  constexpr bool NEED_N = PRED || CORR;
  const int q = (threadIdx.x & 63) >> 4;
  bool vis[G], allv[G];
  uint64_t allm[G]; /* lanes whose antenna of group gi is all-visible */
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    vis[gi] = (bt[gi] & 2) == 0;
    allv[gi] = (bt[gi] & 1) != 0;
    allm[gi] = __ballot(allv[gi]);
  }
  uint64_t failed = 0;
  /* not unrolled: the mixed form then fits 5 waves per SIMD (measured 4 %
   * faster than the compiler's two-step unroll at 4 waves) */
#pragma unroll 1
  for (int st = 0; st < TP / 4; ++st) {
    const int e = 4 * st + q;
    const double l = px.l[e], m = px.m[e], nm1 = px.nm1[e];
    const double n = NEED_N ? px.n[e] : 0.0;
    const T w = (T)px.w[e];
    T cm[G], sm[G], co[G], so[G];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      bool msk, omk = false;
      if (PRED || CORR) {
        const bool pv = vw_visible(l, m, n, L.px[gi], L.py[gi], L.pz[gi], L.pinv[gi], L.pcos[gi]);
        if (CORR) {
          omk = allv[gi];
          msk = allv[gi] & pv;
        } else {
          msk = vis[gi] & pv;
          failed |= ~__ballot(pv) & allm[gi];
        }
      } else {
        msk = vis[gi];
      }
      T c, s;
      if (MIXED) {
        /* coefficients in turns for this form; the fraction in [0, 1) */
        const double x = vw_viss_phase(L.kx[gi], L.ky[gi], L.kz[gi], l, m, nm1);
        float sf, cf;
        vw_hw_sincos_turn((float)__builtin_amdgcn_fract(x), &sf, &cf);
        c = cf;
        s = sf;
      } else {
        int ix;
        const double f = vw_tab_split(vw_viss_phase(L.kx[gi], L.ky[gi], L.kz[gi], l, m, nm1), &ix);
        const double2 te = tab[ix];
        double sy, cy, sd, cd;
        vw_tab_poly_short(f, &sy, &cy);
        vw_tab_rotate(te.x, te.y, sy, cy, &sd, &cd);
        c = (T)cd;
        s = (T)sd;
      }
      cm[gi] = msk ? c : (T)0;
      sm[gi] = msk ? s : (T)0;
      co[gi] = omk ? c : (T)0;
      so[gi] = omk ? s : (T)0;
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      /* CORR: the second product of each block carries the minus sign */
      const T wm = CORR ? -w : w;
      const T wc = wm * cm[gi], ws = wm * sm[gi];
      const T wco = w * co[gi], wso = w * so[gi];
#pragma unroll
      for (int gj = 0; gj <= gi; ++gj) {
        const int b = gi * (gi + 1) / 2 + gj;
        if (CORR) {
          y11[b] = vw_mfma(wc, cm[gj], vw_mfma(wco, co[gj], y11[b]));
          y22[b] = vw_mfma(ws, sm[gj], vw_mfma(wso, so[gj], y22[b]));
        } else {
          y11[b] = vw_mfma(wc, cm[gj], y11[b]);
          y22[b] = vw_mfma(ws, sm[gj], y22[b]);
        }
      }
#pragma unroll
      for (int gj = 0; gj < G; ++gj) {
        if (CORR) y21[gi * G + gj] = vw_mfma(ws, cm[gj], vw_mfma(wso, co[gj], y21[gi * G + gj]));
        else y21[gi * G + gj] = vw_mfma(ws, cm[gj], y21[gi * G + gj]);
      }
    }
  }
  return failed;
}

#ifndef VW_VISSF_WAVES_G2
#define VW_VISSF_WAVES_G2 2
#endif
template <int TP, int G, bool MIXED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(G == 1 ? 4 : VW_VISSF_WAVES_G2))) void k_vissf(
    VissFArgs a) {
  using Acc = typename std::conditional<MIXED, vw_f4, vw_d4>::type;
  constexpr int NS = VfShape<G>::NS, NB = VfShape<G>::NB;
  constexpr int EPT = TP / 256; /* pixels staged per thread */
  __shared__ double2 tab[MIXED ? 1 : VW_TAB_N];
  __shared__ VfTileLds<TP> px;
  if (!MIXED) vw_tab_fill(tab, a.tab);
  const int64_t vb = vw_xcd_block(blockIdx.x, gridDim.x);
  const int64_t chunk = vb / a.nquad;
  if (chunk >= a.nchunk) return; /* whole workgroup */
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int64_t g = (vb - chunk * a.nquad) * 4 + wave; /* epoch of this wave */
  /* waves past the last epoch stage pixels and meet the barriers, nothing else */
  const bool active = g < a.tlen;
  const int64_t t0 = a.tile0 + chunk * a.tpc;
  const int64_t tend = a.tile0 + a.ntile;
  const int64_t t1 = (t0 + a.tpc < tend) ? t0 + a.tpc : tend;

  VfLane<G> L;
  int64_t fr[G];
  bool have[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const int ant = gi * 16 + (lane & 15);
    have[gi] = active && ant < a.A;
    fr[gi] = have[gi] ? g * a.A + ant : 0;
    const double4 k = a.aph[fr[gi]];
    const double* q8 = a.ant8 + fr[gi] * 8;
    /* 1/1024 turns for the table sincos, turns (exact rescaling) for the hardware one */
    const double sc = MIXED ? 1.0 / VW_TAB_SCALE : 1.0;
    L.kx[gi] = have[gi] ? k.x * sc : 0.0;
    L.ky[gi] = have[gi] ? k.y * sc : 0.0;
    L.kz[gi] = have[gi] ? k.z * sc : 0.0;
    L.px[gi] = q8[0]; L.py[gi] = q8[1]; L.pz[gi] = q8[2]; L.pinv[gi] = q8[3]; L.pcos[gi] = q8[7];
  }
  /* the plain form sums in the FP64 accumulators of the matrix cores; the
   * mixed form sums each tile in FP32 and adds it to FP64 totals */
  Acc y11[NS], y22[NS], y21[G * G];
  vw_d4 s11[NS], s22[NS], s21[G * G];
#pragma unroll
  for (int i = 0; i < NS; ++i) { y11[i] = Acc{0, 0, 0, 0}; y22[i] = Acc{0, 0, 0, 0}; s11[i] = vw_d4{0, 0, 0, 0}; s22[i] = vw_d4{0, 0, 0, 0}; }
#pragma unroll
  for (int i = 0; i < G * G; ++i) { y21[i] = Acc{0, 0, 0, 0}; s21[i] = vw_d4{0, 0, 0, 0}; }

  /* pixel data of the next tile, loaded one tile ahead */
  double nl[EPT], nm[EPT], nn[EPT], nz[EPT], nw[EPT];
  auto fetch = [&](int64_t tile) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int64_t p = tile * TP + threadIdx.x + 256 * k;
      const bool ok = p < a.npix;
      /* past the end of the sky: the tile's first pixel with weight zero, so
       * the predicate test sees only the tile's own directions */
      const int64_t pp = ok ? p : tile * TP;
      nl[k] = a.l[pp]; nm[k] = a.m[pp]; nn[k] = a.n[pp]; nz[k] = a.nm1[pp];
      nw[k] = ok ? (MIXED ? (double)a.Bf[pp] : a.B[pp]) : 0.0;
    }
  };
  fetch(t0);
  for (int64_t tile = t0; tile < t1; ++tile) {
    __syncthreads(); /* the previous tile's readers are done (and the table is built) */
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = threadIdx.x + 256 * k;
      px.l[e] = nl[k]; px.m[e] = nm[k]; px.n[e] = nn[k]; px.nm1[e] = nz[k]; px.w[e] = nw[k];
    }
    __syncthreads();
    if (tile + 1 < t1) fetch(tile + 1);
    if (!active) continue;
    const uint8_t* brow = a.bits + (tile - a.tile0) * a.nfr_pad;
    int bt[G];
    int nvis = 0;
    bool any_mixed = false;
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      bt[gi] = have[gi] ? (int)brow[fr[gi]] : 2;
      nvis += __popcll(__ballot((bt[gi] & 2) == 0) & 0xffffull);
      any_mixed |= (__ballot((bt[gi] & 3) == 0) & 0xffffull) != 0;
    }
    /* fewer than two antennas see any of the tile: every pair is hidden */
    if (nvis < 2) continue;
    if (!any_mixed) {
      vf_tile<TP, G, MIXED, false, false>(px, L, bt, tab, y11, y22, y21);
    } else {
      const uint64_t failed = vf_tile<TP, G, MIXED, true, false>(px, L, bt, tab, y11, y22, y21);
      if (a.check || failed != 0) vf_tile<TP, G, MIXED, true, true>(px, L, bt, tab, y11, y22, y21);
    }
    if (MIXED) {
#pragma unroll
      for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          s11[i][r] += (double)y11[i][r];
          s22[i][r] += (double)y22[i][r];
          y11[i][r] = 0;
          y22[i][r] = 0;
        }
#pragma unroll
      for (int i = 0; i < G * G; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          s21[i][r] += (double)y21[i][r];
          y21[i][r] = 0;
        }
    }
  }
  if (!active) return;
  /* partial blocks: Re = Y11 + Y22 for gi >= gj, then Y21; row-major 16 x 16 */
  double* out = a.part + (g * a.nchunk + chunk) * (int64_t)NB * 256;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = MIXED ? 4 * (lane >> 4) + r : (lane >> 4) + 4 * r, j = lane & 15;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      out[s * 256 + i * 16 + j] = MIXED ? s11[s][r] + s22[s][r] : (double)y11[s][r] + (double)y22[s][r];
#pragma unroll
    for (int s = 0; s < G * G; ++s) out[(NS + s) * 256 + i * 16 + j] = MIXED ? s21[s][r] : (double)y21[s][r];
  }
}

/* partial blocks -> per-epoch block sums.  Workgroup (epoch, slice of 32
 * entries); its eight 32-lane groups take every eighth chunk (256-byte
 * reads), then the eight sums are added in a fixed order (deterministic). */
__global__ __launch_bounds__(256) void k_vissf_fold(const double* __restrict__ part, int64_t nchunk, int64_t nent,
                                                    double* __restrict__ esum) {
  __shared__ double red[8][32];
  const int64_t nslice = nent / 32;
  const int64_t g = blockIdx.x / nslice, sl = blockIdx.x - g * nslice;
  const int t = threadIdx.x & 31, c0 = threadIdx.x >> 5;
  const int64_t e = sl * 32 + t;
  double s = 0.0;
  for (int64_t k = c0; k < nchunk; k += 8) s += part[(g * nchunk + k) * nent + e];
  red[c0][t] = s;
  __syncthreads();
  if (c0 == 0) {
    double r = red[0][t];
#pragma unroll
    for (int i = 1; i < 8; ++i) r += red[i][t];
    esum[g * nent + e] = r;
  }
}

/* per-epoch block sums -> raw half sums of the epoch's pairs (pa[j], pb[j]);
 * passes after the first add to the running sums */
__global__ void k_vissf_rows(const double* __restrict__ esum, int64_t nh, int P, int G,
                             const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, int accumulate,
                             double2* __restrict__ vraw) {
  int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const int64_t g = h / P;
  const int j = (int)(h - g * P);
  const int ea = pa[j], eb = pb[j];
  const int hi = ea > eb ? ea : eb, lo = ea > eb ? eb : ea;
  const int NS = G * (G + 1) / 2, NB = NS + G * G;
  const int64_t ire = (int64_t)((hi >> 4) * ((hi >> 4) + 1) / 2 + (lo >> 4)) * 256 + (hi & 15) * 16 + (lo & 15);
  const int64_t iab = (int64_t)(NS + (ea >> 4) * G + (eb >> 4)) * 256 + (ea & 15) * 16 + (eb & 15);
  const int64_t iba = (int64_t)(NS + (eb >> 4) * G + (ea >> 4)) * 256 + (eb & 15) * 16 + (ea & 15);
  const double* b = esum + g * (int64_t)NB * 256;
  double re = b[ire], im = b[iab] - b[iba];
  if (accumulate) {
    double2 o = vraw[h];
    re = o.x + re;
    im = o.y + im;
  }
  vraw[h] = make_double2(re, im);
}

/* per-frame phase coefficients of the factorized form: (x - x_0) * 1e3 /
 * lambda * 1024 with x_0 the epoch's antenna 0 */
__global__ void k_ant_phase(const double* __restrict__ pos, int T, int A, int t0, int tlen, double lambda,
                            double4* __restrict__ aph) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)tlen * A) return;
  const int64_t TA = (int64_t)T * A;
  const int64_t g = k / A;
  const int64_t src = (int64_t)t0 * A + k, ref = (t0 + g) * A;
  double r[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) r[c] = (((pos[c * TA + src] - pos[c * TA + ref]) * 1e3) / lambda) * VW_TAB_SCALE;
  aph[k] = make_double4(r[0], r[1], r[2], 0.0);
}

/* phase correction + conjugate completion (common.hpp:255-277) */
__global__ void k_conj_fill(const double2* __restrict__ vraw, const RowRec* __restrict__ rows, int64_t nh, int P,
                            double2* __restrict__ vh, double2* __restrict__ vfull) {
  int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  double2 z = vraw[h];
  double re, im;
  vw_phase_correct(z.x, z.y, rows[h].w, &re, &im);
  vh[h] = make_double2(re, im);
  int64_t g = h / P, j = h - g * P;
  vfull[2 * P * g + j] = make_double2(re, im);
  vfull[2 * P * g + P + j] = make_double2(re, -im);
}

/* ------------------------------------------------------------------ */
/* segment parameters, keys, stable grouping                            */
/* ------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void k_segparams(const RowRec* __restrict__ rows, int64_t nh, int P,
                                                   double bl_max, double lambda, DevInfo* __restrict__ info) {
  __shared__ double s[5][256];
  const int t = threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0;
  for (int64_t i = t; i < 2 * nh; i += 256) {
    int64_t g = i / (2 * P);
    int jj = (int)(i - g * 2 * P);
    bool neg = jj >= P;
    int64_t h = g * P + (neg ? jj - P : jj);
    double sg = neg ? -1.0 : 1.0;
    RowRec r = rows[h];
    double u = sg * r.u, v = sg * r.v, w = sg * r.w;
    a0 = a0 + u * u;
    a1 = a1 + v * v;
    a2 = a2 + u * v;
    a3 = a3 + u * w;
    a4 = a4 + v * w;
  }
  s[0][t] = a0; s[1][t] = a1; s[2][t] = a2; s[3][t] = a3; s[4][t] = a4;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st)
      for (int k = 0; k < 5; ++k) s[k][t] = s[k][t] + s[k][t + st];
    __syncthreads();
  }
  if (t == 0) vw_segparams_from_sums(s[0][0], s[1][0], s[2][0], s[3][0], s[4][0], bl_max, lambda, &info->sp);
}

/* mode row -> (half row, conjugate flag) */
__device__ __forceinline__ int64_t vw_row_half(int64_t r, int P, bool full, bool* neg) {
  if (!full) { *neg = false; return r; }
  int64_t g = r / (2 * P);
  int jj = (int)(r - g * 2 * P);
  *neg = jj >= P;
  return g * P + (jj >= P ? jj - P : jj);
}

__global__ void k_keys(const RowRec* __restrict__ rows, int64_t nr, int P, int full, const DevInfo* __restrict__ info,
                       int64_t* __restrict__ keys) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  bool neg;
  int64_t h = vw_row_half(r, P, full != 0, &neg);
  double sg = neg ? -1.0 : 1.0;
  RowRec rec = rows[h];
  vw_segparams sp = info->sp;
  keys[r] = vw_grid_key(sg * rec.u, sg * rec.v, &sp, full ? 0 : 1);
}

/* Stable order by (key, row index): all-pairs rank counted over a 2-D grid
 * of (256-row block, jblk-row block) compare tiles with integer atomics
 * (exact, order-independent), then a scatter.  The grid is sized for about
 * 2,048 workgroups, so a segment of a few thousand rows costs a few
 * microseconds instead of a one-workgroup sort. */
__global__ __launch_bounds__(256) void k_rank_count(const int64_t* __restrict__ keys, int64_t nr, int64_t jblk,
                                                    int32_t* __restrict__ rank) {
  __shared__ int64_t sk[256];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * jblk;
  const int64_t j1 = j0 + jblk < nr ? j0 + jblk : nr;
  const int64_t ki = i < nr ? keys[i] : 0;
  int32_t cnt = 0;
  for (int64_t base = j0; base < j1; base += 256) {
    int64_t j = base + threadIdx.x;
    sk[threadIdx.x] = j < j1 ? keys[j] : INT64_MAX;
    __syncthreads();
    int64_t lim = j1 - base < 256 ? j1 - base : 256;
    for (int64_t q = 0; q < lim; ++q) {
      int64_t kj = sk[q];
      cnt += (kj < ki) || (kj == ki && base + q < i);
    }
    __syncthreads();
  }
  if (i < nr && cnt) atomicAdd(&rank[i], cnt);
}

/* segments longer than this many rows are ordered by a radix sort */
#define VW_RANK_MAX 8192

__global__ void k_iota(int64_t n, int32_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (int32_t)i;
}

__global__ void k_rank_scatter(const int64_t* __restrict__ keys, const int32_t* __restrict__ rank, int64_t nr,
                               int32_t* __restrict__ idx_sorted, int64_t* __restrict__ keys_sorted) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  idx_sorted[rank[i]] = (int32_t)i;
  keys_sorted[rank[i]] = keys[i];
}

/* run boundaries -> keys, offsets, nuniq (single workgroup) */
__global__ __launch_bounds__(1024) void k_groups(const int64_t* __restrict__ ks, int64_t nr, int blockage,
                                                 int64_t* __restrict__ gkeys, int32_t* __restrict__ goff,
                                                 DevInfo* __restrict__ info) {
  __shared__ int64_t sc[1024];
  const int t = threadIdx.x;
  int64_t per = (nr + 1023) / 1024;
  int64_t b = t * per, e = b + per < nr ? b + per : nr;
  int64_t c = 0;
  for (int64_t k = b; k < e; ++k) c += (k == 0 || ks[k] != ks[k - 1]);
  sc[t] = c;
  __syncthreads();
  for (int s = 1; s < 1024; s <<= 1) {
    int64_t v = t >= s ? sc[t - s] : 0;
    __syncthreads();
    sc[t] += v;
    __syncthreads();
  }
  int64_t g = sc[t] - c; /* exclusive */
  for (int64_t k = b; k < e; ++k) {
    if (k == 0 || ks[k] != ks[k - 1]) {
      gkeys[g] = ks[k];
      goff[g] = (int32_t)k;
      ++g;
    }
  }
  if (t == 1023) {
    int64_t nu = sc[1023];
    goff[nu] = (int32_t)nr;
    info->nuniq = nu;
    info->first_nz = (nu > 0 && ks[0] == 0) ? 1 : 0;
    if (!blockage) info->wtotal += nu;
  }
}

/* Per-group records as SoA (uniform loads of four groups per array in one
 * scalar load): cell centre, visibility (representative or group mean), the
 * representative's antenna frames and member range. */
struct GRec {
  double ugu, vgu, vr, vi;
};
struct GroupSoA {
  GRec* rec;    /* cell centre (wavelengths; x1024 in the plain form) and visibility, 32 B */
  double* ant;  /* [q][10]: representative's antenna frames a and b    */
  int32_t *gia, *gib, *grep;
};

__global__ void k_group_data(const int64_t* __restrict__ gkeys, const int32_t* __restrict__ goff,
                             const int32_t* __restrict__ idx_sorted, const double2* __restrict__ vh,
                             const RowRec* __restrict__ rows, const double* __restrict__ ant8, int P, int full,
                             int blockage, int occ_mode, double uv_scale,
                             const DevInfo* __restrict__ info, GroupSoA g) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= info->nuniq) return;
  vw_segparams sp = info->sp;
  int64_t key = gkeys[q];
  int32_t s0 = goff[q], s1 = goff[q + 1];
  double ugu = 0.0, vgu = 0.0;
  if (key != 0) vw_key_uv(key, &sp, &ugu, &vgu);
  bool neg;
  /* the one member a single-representative rule reads: occ_mode 1 the
   * first of the sorted run, occ_mode 2 the median (viss_recon_kernel.cuh:
   * 603-614; mirror_tile_recon_group .cuh:782-792) */
  int64_t hr = vw_row_half(idx_sorted[occ_mode == 1 ? s0 : s0 + (s1 - s0) / 2], P, full != 0, &neg);
  double vr, vi;
  if (blockage) {
    /* representative value (signed for rows of the conjugate half) */
    double2 v = vh[hr];
    vr = v.x;
    vi = neg ? -v.y : v.y;
  } else {
    /* group mean over members in sorted order (compute_avg, viss_recon_kernel.cuh:327-333) */
    double sr = 0.0, si = 0.0;
    for (int32_t k = s0; k < s1; ++k) {
      bool ng;
      int64_t h = vw_row_half(idx_sorted[k], P, full != 0, &ng);
      double2 v = vh[h];
      sr = sr + v.x;
      si = si + (ng ? -v.y : v.y);
    }
    vr = sr / (double)(s1 - s0);
    vi = si / (double)(s1 - s0);
  }
  GRec gr;
  gr.ugu = ugu * uv_scale; /* 1024 for the plain table sincos, 1 for the mixed form (exact) */
  gr.vgu = vgu * uv_scale;
  gr.vr = vr;
  gr.vi = vi;
  g.rec[q] = gr;
  g.grep[q] = (int32_t)hr;
  RowRec r = rows[hr];
  g.gia[q] = r.ia;
  g.gib[q] = r.ib;
  const double* qa = ant8 + (int64_t)r.ia * 8;
  const double* qb = ant8 + (int64_t)r.ib * 8;
  double* d = g.ant + q * 10;
  d[0] = qa[0]; d[1] = qa[1]; d[2] = qa[2]; d[3] = qa[3]; d[4] = qa[7];
  d[5] = qb[0]; d[6] = qb[1]; d[7] = qb[2]; d[8] = qb[3]; d[9] = qb[7];
}

/* class of (output tile, group) from the representative's antennas: entry j
 * of a tile's row pairs frames ia[j] and ib[j] (the groups' own arrays, or
 * for the batched recon the same arrays in its work order, which puts the
 * groups past nuniq last, so in both entries j >= nuniq are hidden).  A
 * block covers VW_GCLS_BLOCK entries of one tile's row, four a thread
 * written as one class word (byte e = entry 4 i + e).  Each thread issues
 * its index loads first, then the block stages that tile's frame row of
 * cone bits in LDS (STAGE; nfr_pad bytes of dynamic LDS; longer rows are
 * gathered from global memory), so the loads overlap.  Blocks that start
 * past the 64-entry block holding entry nuniq - 1 write nothing: the recon
 * kernels read the classes of entries below nuniq only, in words of 4 (one
 * sky) or blocks of 32 (batched) aligned to their size. */
template <bool STAGE>
__global__ __launch_bounds__(256) void k_group_cls(const uint8_t* __restrict__ bits, int64_t nfr_pad, int nqb,
                                                   const int32_t* __restrict__ ia, const int32_t* __restrict__ ib,
                                                   const DevInfo* __restrict__ info, int64_t nq_pad,
                                                   uint32_t* __restrict__ gcls) {
  extern __shared__ uint4 sbits[];
  static_assert(VW_GCLS_BLOCK == 4 * 256, "four entries a thread");
  const uint32_t tr = blockIdx.x / (uint32_t)nqb, qc = blockIdx.x - tr * (uint32_t)nqb;
  const int64_t j00 = (int64_t)qc * VW_GCLS_BLOCK;
  const int64_t nu = info->nuniq;
  if (j00 >= (nu + 63) / 64 * 64) return;
  const uint8_t* row = bits + (int64_t)tr * nfr_pad;
  const int64_t j0 = j00 + 4 * threadIdx.x;
  const bool live = j0 < nq_pad; /* nq_pad is a multiple of 64: whole words only */
  const int64_t jl = live ? j0 : 0;
  /* indices past nuniq are not frames: gather frame 0 for them instead */
  const int4 a4 = *(const int4*)(ia + jl), b4 = *(const int4*)(ib + jl);
  const int av[4] = {a4.x, a4.y, a4.z, a4.w}, bv[4] = {b4.x, b4.y, b4.z, b4.w};
  if (STAGE) {
    /* nfr_pad is a multiple of 64 and the row starts 64-byte aligned */
    for (int64_t i = threadIdx.x; i < nfr_pad / 16; i += 256) sbits[i] = ((const uint4*)row)[i];
    __syncthreads();
  }
  const uint8_t* b = STAGE ? (const uint8_t*)sbits : row;
  uint32_t w = 0;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const bool grp = jl + e < nu;
    const uint32_t cls = (uint32_t)vw_pair_class(b[grp ? av[e] : 0], b[grp ? bv[e] : 0]);
    w |= (grp ? cls : (uint32_t)VW_CLS_HIDDEN) << (8 * e);
  }
  if (live) gcls[((int64_t)tr * nq_pad + j0) >> 2] = w;
}

/* k_group_cls with the cone bits made in place: one block per tile of the
 * pass computes the tile's frame bits into LDS (nfr bytes of dynamic LDS),
 * then writes the class words of entries below nuniq rounded up to 64 (the
 * entries the recon kernels read), VW_GCLS_BLOCK at a time */
__global__ __launch_bounds__(256) void k_group_cls_tile(const double* __restrict__ cone, int64_t ntile,
                                                        int64_t tile0, const double* __restrict__ ant8, int64_t nfr,
                                                        const int32_t* __restrict__ ia,
                                                        const int32_t* __restrict__ ib,
                                                        const DevInfo* __restrict__ info, int64_t nq_pad,
                                                        uint32_t* __restrict__ gcls) {
  extern __shared__ uint4 sbits[];
  const int64_t tr = blockIdx.x;
  const int64_t nu = info->nuniq;
  const int64_t qend = (nu + 63) / 64 * 64 < nq_pad ? (nu + 63) / 64 * 64 : nq_pad;
  if (qend == 0) return;
  uint8_t* sb = (uint8_t*)sbits;
  vw_tile_bits_lds(sb, cone, ntile, tile0 + tr, ant8, nfr);
  __syncthreads();
  for (int64_t j0 = 4 * threadIdx.x; j0 < qend; j0 += VW_GCLS_BLOCK) {
    /* indices past nuniq are not frames: frame 0 stands in for them */
    const int4 a4 = *(const int4*)(ia + j0), b4 = *(const int4*)(ib + j0);
    const int av[4] = {a4.x, a4.y, a4.z, a4.w}, bv[4] = {b4.x, b4.y, b4.z, b4.w};
    uint32_t w = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool grp = j0 + e < nu;
      const uint32_t cls = (uint32_t)vw_pair_class(sb[grp ? av[e] : 0], sb[grp ? bv[e] : 0]);
      w |= (grp ? cls : (uint32_t)VW_CLS_HIDDEN) << (8 * e);
    }
    gcls[(tr * nq_pad + j0) >> 2] = w;
  }
}

/* ------------------------------------------------------------------ */
/* reconstruction                                                       */
/* ------------------------------------------------------------------ */
struct ReconArgs {
  const double2* tab; /* the sine table (plain forms), built once per context */
  const double *l, *m, *n;
  int64_t pt0, npts, TP; /* points [pt0, pt0 + npts) of this pass, pt0 = tile0 * TP */
  const uint32_t* gcls; /* [ntile_out][nq_pad]: per group (representative kind) */
  int64_t nq_pad;
  const uint32_t* rcls; /* [ntile_out][nh_pad]: per half row (generic kind)     */
  int64_t nh_pad;
  GroupSoA g;
  const int32_t* goff;
  const int32_t* idx_sorted;
  const double2* vh;
  const RowRec* rows;
  const double* ant8;
  const DevInfo* info;
  int P, full, occ_mode;
  int64_t nchunk;     /* group ranges per point block                     */
  int64_t pwaves;     /* point blocks (waves) per group range             */
  int64_t pstride;    /* partial row stride                               */
  double* part;       /* [nchunk][pstride]                                */
  uint32_t* wpart;    /* [nchunk][pstride]                                */
};

#define VW_KIND_REP 0
#define VW_KIND_GENERIC 1
#define VW_KIND_AVG 2
#define VW_KIND_REP_FINITE 3 /* representatives, groups with a non-finite value skipped */
#define VW_KIND_IS_REP(k) ((k) == VW_KIND_REP || (k) == VW_KIND_REP_FINITE)

/* visibility of one row for the PPL points of a lane given the class */
template <int PPL, bool BRANCHFREE>
__device__ __forceinline__ void vw_ant_vis(const double* __restrict__ ant8, int ia, int ib, const double* pl,
                                           const double* pm, const double* pn, const bool* ok, bool* vis) {
  const double* qa = ant8 + (int64_t)ia * 8;
  const double* qb = ant8 + (int64_t)ib * 8;
  const double ax = qa[0], ay = qa[1], az = qa[2], ai = qa[3], ac = qa[7];
  const double bx = qb[0], by = qb[1], bz = qb[2], bi = qb[3], bc = qb[7];
#pragma unroll
  for (int k = 0; k < PPL; ++k)
    vis[k] = vw_vis2<BRANCHFREE>(ok[k], pl[k], pm[k], pn[k], ax, ay, az, ai, ac, bx, by, bz,
             bi, bc);
}

/* FP64 phase and turn reduction, hardware FP32 sin/cos, FP32 block sums
 * folded with Kahan compensation (mixed form, one point) */
struct ReconAccF32 {
  float s = 0.f, c = 0.f, b = 0.f;
  __device__ void term(double ugu, double vgu, float vrf, float vif, double lp, double mp) {
    /* the fraction in [0, 1) turns: one v_fract_f64 */
    const double x = vw_recon_phase(ugu, vgu, lp, mp);
    const float f = (float)__builtin_amdgcn_fract(x);
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    b = fmaf(vrf, cs, fmaf(-vif, sn, b));
  }
  /* the same, kept only where use holds: a select, not a branch */
  __device__ void term_if(bool use, double ugu, double vgu, float vrf, float vif, double lp, double mp) {
    const double x = vw_recon_phase(ugu, vgu, lp, mp);
    const float f = (float)__builtin_amdgcn_fract(x);
    const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
    const float nb = fmaf(vrf, cs, fmaf(-vif, sn, b));
    b = use ? nb : b;
  }
  __device__ void fold() {
    float y = b - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
    b = 0.f;
  }
  __device__ double value() const { return ((double)s - (double)c) + (double)b; }
};

/* the PPL points of one lane and one group.  Plain: (ugu, vgu) arrive
 * scaled by 1024; all table indices and reads come before the polynomials
 * so the LDS reads overlap the arithmetic; a point with use[k] false keeps
 * its sum. */
template <int PPL, bool MIXED>
struct ReconLane {
  double d[PPL];
  ReconAccF32 f[PPL];
  __device__ ReconLane() {
#pragma unroll
    for (int k = 0; k < PPL; ++k) d[k] = 0.0;
  }
  __device__ __forceinline__ void add(double ugu, double vgu, const double* vr, const double* vi, const bool* use,
                                      const double* lp, const double* mp, const double2* tab) {
    if (MIXED) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) f[k].term_if(use[k], ugu, vgu, (float)vr[k], (float)vi[k], lp[k], mp[k]);
    } else {
      double fr[PPL];
      int ix[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) fr[k] = vw_tab_split(vw_recon_phase(ugu, vgu, lp[k], mp[k]), &ix[k]);
      double2 e[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) e[k] = tab[ix[k]];
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        double sy, cy, sn, cs;
        vw_tab_poly_short(fr[k], &sy, &cy);
        vw_tab_rotate(e[k].x, e[k].y, sy, cy, &sn, &cs);
        const double t = vw_recon_add(d[k], vr[k], vi[k], sn, cs);
        d[k] = use[k] ? t : d[k];
      }
    }
  }
  __device__ void fold() {
    if (MIXED) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) f[k].fold();
    }
  }
  __device__ double value(int k) const { return MIXED ? f[k].value() : d[k]; }
};

#ifndef VW_RECON_WAVES_PLAIN
#define VW_RECON_WAVES_PLAIN 4
#endif
#ifndef VW_RECON_WAVES_MIXED
#define VW_RECON_WAVES_MIXED 5
#endif
template <int PPL, bool MIXED, int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    MIXED && KIND != VW_KIND_GENERIC ? VW_RECON_WAVES_MIXED : VW_RECON_WAVES_PLAIN))) void k_recon(ReconArgs a) {
  __shared__ double2 tab[MIXED ? 1 : VW_TAB_N];
  __shared__ double gring[4][2 * 64]; /* per wave: two blocks of 16 group records */
  if (!MIXED) {
    vw_tab_fill(tab, a.tab);
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int64_t wid = (int64_t)blockIdx.x * 4 + vw_rfl((int)(threadIdx.x >> 6));
  const int64_t chunk = wid / a.pwaves;
  const int64_t pw = wid - chunk * a.pwaves;
  if (chunk >= a.nchunk) return;
  const int64_t pbase = pw * 64 * PPL; /* pass-relative */
  const int64_t tile = pbase / a.TP;   /* pass-relative tile (class tables are per pass) */
  const vw_segparams sp = a.info->sp;
  /* group counts fit 32 bits; uniform values keep the range tests scalar */
  const int64_t nu = vw_rfl((int)a.info->nuniq);
  const int64_t q0 = vw_rfl((int)a.info->first_nz);
  /* group range of this chunk, aligned to 16 groups */
  const int64_t base = q0 & ~(int64_t)15;
  const int64_t span = ((nu - base) + a.nchunk - 1) / a.nchunk;
  const int64_t spanA = (span + 15) & ~(int64_t)15;
  const int64_t qa_ = vw_rfl((int)(base + chunk * spanA));
  const int64_t qb_ = vw_rfl((int)(qa_ + spanA < nu ? qa_ + spanA : nu));

  double pl[PPL], pm[PPL], pn[PPL], lp[PPL], mp[PPL];
  bool ok[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    int64_t p = pbase + lane + 64 * k;
    ok[k] = p < a.npts;
    int64_t pp = a.pt0 + (ok[k] ? p : 0);
    pl[k] = a.l[pp];
    pm[k] = a.m[pp];
    pn[k] = a.n[pp];
    lp[k] = fma(sp.fa, pn[k], pl[k]);
    mp[k] = fma(sp.fb, pn[k], pm[k]);
  }
  ReconLane<PPL, MIXED> acc;
  uint32_t wl[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) wl[k] = 0;
  uint32_t wu = 0;
  int nblk = 0;

  if (VW_KIND_IS_REP(KIND) || KIND == VW_KIND_AVG) {
    /* records pass through a per-wave LDS ring in blocks of 16 (512 B: one
     * double per lane, loaded one block ahead); each group is then read
     * with a uniform LDS address.  Class words (4 groups each) are scalar
     * loads fetched one word ahead.  Blocks start at multiples of 16 and
     * nq_pad is a multiple of 64, so every block lies inside the records. */
    double* ring = gring[vw_rfl((int)(threadIdx.x >> 6))];
    const double* recd = (const double*)a.g.rec;
    double stage = (qa_ < qb_) ? recd[qa_ * 4 + lane] : 0.0;
    uint32_t ncw = 0x01010101u; /* AVG: every group counts, no class */
    if (VW_KIND_IS_REP(KIND) && qa_ < qb_) ncw = (uint32_t)vw_rfl((int)a.gcls[(tile * a.nq_pad + qa_) >> 2]);
    int buf = 0;
    for (int64_t q16 = qa_; q16 < qb_; q16 += 16, buf ^= 1) {
      ring[buf * 64 + lane] = stage;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (q16 + 16 < qb_) stage = recd[(q16 + 16) * 4 + lane];
      const GRec* blk = (const GRec*)(ring + buf * 64);
      const int64_t qe = q16 + 16 < qb_ ? q16 + 16 : qb_;
      for (int64_t q4 = q16; q4 < qe; q4 += 4) {
        const uint32_t cw = ncw;
        if (VW_KIND_IS_REP(KIND) && q4 + 4 < qb_)
          ncw = (uint32_t)vw_rfl((int)a.gcls[(tile * a.nq_pad + q4 + 4) >> 2]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int64_t q = q4 + j;
          const int cls = (int)((cw >> (8 * j)) & 0xffu);
          if ((int)q < (int)q0 || (int)q >= (int)qb_ || cls == VW_CLS_HIDDEN) continue;
          const GRec g = blk[q - q16];
          if (KIND == VW_KIND_REP_FINITE && !(isfinite(g.vr) && isfinite(g.vi))) continue;
          double vr[PPL], vi[PPL];
#pragma unroll
          for (int k = 0; k < PPL; ++k) { vr[k] = g.vr; vi[k] = g.vi; }
          if (cls == VW_CLS_VISIBLE) {
            bool all[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) all[k] = true;
            acc.add(g.ugu, g.vgu, vr, vi, all, lp, mp, tab);
            wu += 1;
          } else {
            const double* d = a.g.ant + q * 10;
            const double ax = d[0], ay = d[1], az = d[2], ai = d[3], ac = d[4];
            const double bx = d[5], by = d[6], bz = d[7], bi = d[8], bc = d[9];
            bool vis[PPL];
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
              vis[k] = vw_vis2<true>(ok[k], pl[k], pm[k], pn[k], ax, ay, az, ai, ac, bx, by, bz, bi, bc);
              wl[k] += vis[k] ? 1 : 0;
            }
            acc.add(g.ugu, g.vgu, vr, vi, vis, lp, mp, tab);
          }
          if (MIXED && ++nblk == 32) {
            nblk = 0;
            acc.fold();
          }
        }
      }
    }
  } else {
    /* generic representative sets (viss_recon_kernel.cuh:603-614, isfinite filter) */
    for (int64_t q = (qa_ > q0 ? qa_ : q0); q < qb_; ++q) {
      const int32_t s0 = a.goff[q], s1 = a.goff[q + 1], L = s1 - s0;
      int32_t posv[4];
      int np;
      if (a.occ_mode == 0) np = -1;
      else if (a.occ_mode == 1) { posv[0] = s0; np = 1; }
      else if (a.occ_mode == 2) { posv[0] = s0 + (L >> 1); np = 1; }
      else { posv[0] = s0; posv[1] = s0 + L / 3; posv[2] = s0 + (2 * L) / 3; posv[3] = s1 - 1; np = 4; }
      double sr[PPL], si[PPL];
      int cn[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) { sr[k] = 0.0; si[k] = 0.0; cn[k] = 0; }
      const int32_t kend = (np < 0) ? s1 : np;
      for (int32_t kk = (np < 0 ? s0 : 0); kk < kend; ++kk) {
        const int32_t pos = (np < 0) ? kk : posv[kk];
        bool ng;
        const int64_t h = vw_row_half(a.idx_sorted[pos], a.P, a.full != 0, &ng);
        const int cls = vw_cls_lookup(a.rcls, tile, a.nh_pad, h);
        if (cls == VW_CLS_HIDDEN) continue;
        bool mv[PPL];
        if (cls == VW_CLS_VISIBLE) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) mv[k] = ok[k];
        } else {
          const RowRec rec = a.rows[h];
          vw_ant_vis<PPL, !MIXED>(a.ant8, rec.ia, rec.ib, pl, pm, pn, ok, mv);
        }
        const double2 v = a.vh[h];
        const double wr = v.x, wi = ng ? -v.y : v.y;
        const bool fin = isfinite(wr) && isfinite(wi);
#pragma unroll
        for (int k = 0; k < PPL; ++k)
          if (mv[k] && fin) { sr[k] = sr[k] + wr; si[k] = si[k] + wi; cn[k] += 1; }
      }
      const double ugu = a.g.rec[q].ugu, vgu = a.g.rec[q].vgu;
      double ar[PPL], ai[PPL];
      bool use[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        use[k] = cn[k] > 0;
        const double invc = 1.0 / (double)(use[k] ? cn[k] : 1);
        ar[k] = sr[k] * invc;
        ai[k] = si[k] * invc;
        wl[k] += use[k] ? 1 : 0;
      }
      acc.add(ugu, vgu, ar, ai, use, lp, mp, tab);
      if (MIXED && ++nblk == 32) {
        nblk = 0;
        acc.fold();
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    if (!ok[k]) continue;
    const int64_t p = pbase + lane + 64 * k;
    a.part[chunk * a.pstride + p] = acc.value(k);
    a.wpart[chunk * a.pstride + p] = wu + wl[k];
  }
}

/* fold group-range partials in order into the accumulators (add_inplace) for
 * points [pt0, pt0 + npts) */
__global__ void k_recon_reduce(const double* __restrict__ part, const uint32_t* __restrict__ wpart, int64_t nchunk,
                               int64_t pstride, int64_t pt0, int64_t npts, int weights, double* __restrict__ acc,
                               uint32_t* __restrict__ wacc) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  double s = part[p];
  uint32_t w = wpart[p];
  /* VW_FOLD_BATCH chunks' loads in flight at a time; the sum keeps chunk order */
  int64_t k = 1;
  for (; k + VW_FOLD_BATCH <= nchunk; k += VW_FOLD_BATCH) {
    double x[VW_FOLD_BATCH];
    uint32_t y[VW_FOLD_BATCH];
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) {
      x[e] = part[(k + e) * pstride + p];
      y[e] = wpart[(k + e) * pstride + p];
    }
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) {
      s = s + x[e];
      w += y[e];
    }
  }
  for (; k < nchunk; ++k) {
    s = s + part[k * pstride + p];
    w += wpart[k * pstride + p];
  }
  acc[pt0 + p] = acc[pt0 + p] + s;
  if (weights) wacc[pt0 + p] += w;
}

/* final maps: C = acc / W (blockage 1) or acc * (1/W_total) (blockage 0) */
__global__ void k_normalize(const double* __restrict__ acc, const uint32_t* __restrict__ w, int64_t n, int blockage,
                            const DevInfo* __restrict__ info, const int64_t* __restrict__ scatter,
                            double* __restrict__ out, uint32_t* __restrict__ wout) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v;
  if (blockage) v = w[k] > 0 ? acc[k] / (double)w[k] : 0.0;
  else {
    int64_t wt = info->wtotal;
    double inv = wt > 0 ? 1.0 / (double)wt : 0.0;
    v = acc[k] * inv;
  }
  int64_t dst = scatter ? scatter[k] : k;
  out[dst] = v;
  wout[dst] = w[k];
}

#endif
