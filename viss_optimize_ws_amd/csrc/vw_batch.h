/*
 * vw_batch.h -- sky batches: many sky realisations through one geometry.
 *
 * The skies of a batch share antennas, epochs, the segment schedule and so
 * the rows, tile classes, w-plane fit, uv cells and groups of every segment.
 * Both sums then become products of a phasor matrix that depends only on
 * the geometry with a matrix that holds one column per sky:
 *
 *   Viss   V[h][s] = sum_p  M[h][p] exp(-j 2 pi x_hp) I[p][s]
 *          (rows x pixels) x (pixels x skies)
 *   recon  C[q][s] = sum_g  M[q][g] Re(exp(+j 2 pi y_qg) V_g[s])
 *          (points x groups) x (groups x skies)
 *
 * The phasor tiles are made on the fly (FP64 phase and its reduction mod 1
 * turn, hardware FP32 sin/cos, the occlusion mask as zeros) and never stored;
 * the products run on the BF16 matrix cores (v_mfma_f32_16x16x32_bf16) with
 * both operands split x = hi + lo into two BF16 values and three products
 * hi*hi + hi*lo + lo*hi per real multiply-add (about 2^-16 relative per
 * term), FP32 accumulation over one 256-pixel sub-tile or 256-group block,
 * then FP64 sums across blocks, chunks and passes.
 *
 * Operand layout of v_mfma_f32_16x16x32_bf16 (wave64), as used here:
 *   A (16 x 32, rows x k):  lane l holds row l % 16, k = 8 (l / 16) + i, i < 8
 *   B (32 x 16, k x cols):  lane l holds column l % 16, the same k
 *   D (16 x 16, f32):       lane l holds D[4 (l / 16) + i][l % 16], i < 4
 * A and B use the same k assignment, so any permutation of k inside the
 * instruction leaves the product unchanged.
 *
 * Packed operand matrices (sky, group visibilities) are stored as
 * [k / 32][nsky_pad][32] BF16, so a lane's 8 values are one 16-byte load.
 */
#ifndef VW_BATCH_H
#define VW_BATCH_H

#include "vw_kernels.h"

typedef __bf16 vw_bf16x8 __attribute__((ext_vector_type(8)));
typedef float vw_f32x4b __attribute__((ext_vector_type(4)));

typedef float vw_f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 vw_bf16x2 __attribute__((ext_vector_type(2)));

/* x = hi + lo with hi = bf16(x) and lo = bf16(x - hi), both rounded to
 * nearest even: x - hi is exact in FP32 and |x - hi| <= 2^-9 |x|, so
 * |x - hi - lo| <= 2^-18 |x| and the omitted lo*lo product of a split
 * multiply is below 2^-18 of the term.  Per pair of values: one packed
 * conversion, a shift and a mask back to FP32, two subtractions, one packed
 * conversion of the remainders. */
__device__ __forceinline__ void vw_split8(const float (&x)[8], vw_bf16x8& hi, vw_bf16x8& lo) {
  uint32_t h[4], l[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const vw_f32x2 v = {x[2 * j], x[2 * j + 1]};
    const uint32_t hp = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, vw_bf16x2));
    const float h0 = __builtin_bit_cast(float, hp << 16), h1 = __builtin_bit_cast(float, hp & 0xffff0000u);
    const vw_f32x2 rr = {x[2 * j] - h0, x[2 * j + 1] - h1};
    h[j] = hp;
    l[j] = __builtin_bit_cast(uint32_t, __builtin_convertvector(rr, vw_bf16x2));
  }
  hi = __builtin_bit_cast(vw_bf16x8, h);
  lo = __builtin_bit_cast(vw_bf16x8, l);
}

__device__ __forceinline__ vw_bf16x8 vw_ld_bf16x8(const uint4* p) {
  return __builtin_bit_cast(vw_bf16x8, *p);
}

__device__ __forceinline__ vw_f32x4b vw_mfma_bf16(vw_bf16x8 a, vw_bf16x8 b, vw_f32x4b c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* three-product split multiply-accumulate: (ah + al)(bh + bl) without al*bl */
__device__ __forceinline__ vw_f32x4b vw_mac3(vw_bf16x8 ah, vw_bf16x8 al, vw_bf16x8 bh, vw_bf16x8 bl,
                                             vw_f32x4b c) {
  c = vw_mfma_bf16(al, bh, c);
  c = vw_mfma_bf16(ah, bl, c);
  return vw_mfma_bf16(ah, bh, c);
}

/* ------------------------------------------------------------------ */
/* packing                                                              */
/* ------------------------------------------------------------------ */

/* sky batch [nsky][npix] FP32 (already scaled by the pixel area) ->
 * hi / lo [npix_pad / 32][nsky_pad][32]; zero for padding pixels and skies */
__global__ void k_sky_pack(const float* __restrict__ sky, int64_t nsky, int64_t npix, int64_t npix_pad,
                           int64_t nsky_pad, __bf16* __restrict__ hi, __bf16* __restrict__ lo) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npix_pad * nsky_pad) return;
  const int64_t kb = k / (nsky_pad * 32);
  const int64_t r = k - kb * nsky_pad * 32;
  const int64_t s = r / 32, j = r - s * 32;
  const int64_t p = kb * 32 + j;
  const float x = (s < nsky && p < npix) ? sky[s * npix + p] : 0.0f;
  const __bf16 h = (__bf16)x;
  hi[k] = h;
  lo[k] = (__bf16)(x - (float)h);
}

/* ------------------------------------------------------------------ */
/* work order                                                           */
/* ------------------------------------------------------------------ */

/* The batched kernels walk rows (Viss) and groups (recon) in an order that
 * keeps rows with similar occlusion together, so that more 16-row blocks are
 * hidden on a whole tile and skipped: the sort key of a row is the unordered
 * pair of the coarse direction cells of its two antennas (each unit vector
 * quantised to 4 levels per axis and Morton-interleaved: 64 cells).  At
 * Nside 128 / 16 antennas this takes the share of skipped (tile, 16-row)
 * blocks from 4% to 35% (the per-tile bound is 43%).  Entries past the count
 * get key 4096 and sort last; the sort is stable, so ties keep index order. */
#define VW_ORDER_BITS 13
__device__ __forceinline__ int vw_dir_cell(double x, double y, double z) {
  const double v[3] = {x, y, z};
  int c = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int q = (int)((v[d] + 1.0) * 2.0);
    q = q < 0 ? 0 : (q > 3 ? 3 : q);
    c |= ((q >> 1) & 1) << (3 + d);
    c |= (q & 1) << d;
  }
  return c;
}
__device__ __forceinline__ int32_t vw_pair_key(const double* __restrict__ ant8, int ia, int ib) {
  const double* qa = ant8 + (int64_t)ia * 8;
  const double* qb = ant8 + (int64_t)ib * 8;
  const int ca = vw_dir_cell(qa[4], qa[5], qa[6]), cb = vw_dir_cell(qb[4], qb[5], qb[6]);
  return (ca < cb) ? (ca << 6) | cb : (cb << 6) | ca;
}
__global__ void k_row_keys(const RowRec* __restrict__ rows, const double* __restrict__ ant8, int64_t nh, int64_t n,
                           int32_t* __restrict__ keys, int32_t* __restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = i < nh ? vw_pair_key(ant8, rows[i].ia, rows[i].ib) : 4096;
  idx[i] = (int32_t)i;
}
/* cell centres, representative antenna frames and their frame indices in
 * the recon's work order */
__global__ void k_group_order_recs(const int32_t* __restrict__ order, const GRec* __restrict__ rec,
                                   const double* __restrict__ gant, const int32_t* __restrict__ gia,
                                   const int32_t* __restrict__ gib, int64_t n, double2* __restrict__ orec,
                                   double* __restrict__ oant, int32_t* __restrict__ oia, int32_t* __restrict__ oib) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t q = order[j];
  const GRec g = rec[q];
  orec[j] = make_double2(g.ugu, g.vgu);
  oia[j] = gia[q];
  oib[j] = gib[q];
#pragma unroll
  for (int f = 0; f < 10; ++f) oant[j * 10 + f] = gant[q * 10 + f];
}
__global__ void k_group_keys(const int32_t* __restrict__ gia, const int32_t* __restrict__ gib,
                             const double* __restrict__ ant8, const DevInfo* __restrict__ info, int64_t n,
                             int32_t* __restrict__ keys, int32_t* __restrict__ idx) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = i < info->nuniq ? vw_pair_key(ant8, gia[i], gib[i]) : 4096;
  idx[i] = (int32_t)i;
}

/* ------------------------------------------------------------------ */
/* Viss                                                                 */
/* ------------------------------------------------------------------ */
struct VissBArgs {
  const double *l, *m, *n, *nm1;
  const uint4 *skyh, *skyl; /* packed sky, 16-byte units                        */
  int64_t npix, nsky_pad;
  int64_t tile0, ntile, TP;  /* tiles of this pass                              */
  const RowRec* rows;
  const double* ant8;
  const uint8_t* cls;        /* [ntile][nh_pad] pass-relative class bytes         */
  const int32_t* perm;       /* work order: block row j is half row perm[j]       */
  int64_t nh, nh_pad;
  int64_t tpc, nchunk, nrb, nrbw; /* tiles per chunk, chunks, 16-row blocks, blocks of 4 */
  double2* part;             /* [nchunk][nsky_pad][nh_pad]                        */
};

/* LDS staging of one k-step of a packed operand: NARR arrays x (16 NB)
 * rows of 32 BF16, each row padded to 40 BF16 (80 B) so that the 16 rows a
 * 16-lane phase reads fall on disjoint banks.  A workgroup loads step k + 1
 * into registers (16 B per thread and array) while it computes step k, then
 * stores them into the other buffer. */
#define VW_LROW 40
typedef unsigned int vw_u32x4 __attribute__((ext_vector_type(4)));
template <int NARR, int NB>
struct VwStage {
  static constexpr int ROWS = NARR * 16 * NB; /* LDS rows per k-step */
  vw_u32x4 r0, r1, r2, r3;
  /* thread t < 64 NB holds the 16-byte unit t of every array's block kb:
   * row row0 + t / 4 of the sky group, quarter t % 4 of its 32 values */
  __device__ __forceinline__ void load(const uint4* p0, const uint4* p1, const uint4* p2, const uint4* p3, int64_t kb,
                                       int64_t nsky_pad, int64_t row0) {
    const int t = (int)threadIdx.x;
    if (t < 64 * NB) {
      const int64_t off = (kb * nsky_pad + row0 + (t >> 2)) * 4 + (t & 3);
      r0 = *(const vw_u32x4*)(p0 + off);
      if (NARR > 1) r1 = *(const vw_u32x4*)(p1 + off);
      if (NARR > 2) r2 = *(const vw_u32x4*)(p2 + off);
      if (NARR > 3) r3 = *(const vw_u32x4*)(p3 + off);
    }
  }
  __device__ __forceinline__ void store(__bf16* buf) const {
    const int t = (int)threadIdx.x;
    if (t < 64 * NB) {
      __bf16* d = buf + (t >> 2) * VW_LROW + (t & 3) * 8;
      *(vw_u32x4*)d = r0;
      if (NARR > 1) *(vw_u32x4*)(d + 16 * NB * VW_LROW) = r1;
      if (NARR > 2) *(vw_u32x4*)(d + 32 * NB * VW_LROW) = r2;
      if (NARR > 3) *(vw_u32x4*)(d + 48 * NB * VW_LROW) = r3;
    }
  }
};

__device__ __forceinline__ vw_bf16x8 vw_lds_bf16x8(const __bf16* buf, int row, int g) {
  return *(const vw_bf16x8*)(buf + row * VW_LROW + g * 8);
}

/* one wave: 16 half rows x (16 NB) skies over the tiles of one chunk; a
 * workgroup is four waves (64 consecutive rows) on the same pixels.  Per
 * k-step of 32 pixels the workgroup stages the sky operand, the pixels' l,
 * m, n - 1, n and the 64 rows' class bytes on the step's tile through
 * double-buffered LDS.  Every step issues the same global loads, for the
 * next step, unconditionally (indices clamped in range), and waits for them
 * only before their LDS stores at its end: a load waited for right after it
 * is issued would also wait for the loads in flight before it (the vector
 * memory counter is in order), exposing their latency. */
#ifndef VW_VISSB_WAVES
#define VW_VISSB_WAVES 3
#endif
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VW_VISSB_WAVES))) void k_vissb(VissBArgs a) {
  using Stage = VwStage<2, NB>;
  __shared__ __attribute__((aligned(16))) __bf16 bsh[2][Stage::ROWS * VW_LROW];
  __shared__ double px[2][4][32]; /* l, m, n - 1, n of a k-step's pixels */
  __shared__ uint8_t pcls[2][64]; /* class bytes of the 64 rows on a k-step's tile */
  __shared__ double frs[10][64];  /* the 64 rows' antenna frames (mixed-tile masks) */
  const int64_t vb = vw_xcd_block(blockIdx.x, gridDim.x);
  const int64_t rbw = vb % a.nrbw;
  const int64_t rest = vb / a.nrbw;
  const int64_t chunk = rest % a.nchunk;
  const int64_t sg = rest / a.nchunk;
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  const int64_t rb = rbw * 4 + wave;
  const bool wave_on = rb < a.nrb; /* uniform; every wave still joins the barriers */
  const int64_t jr = rb * 16 + r;  /* block row; < nh_pad: nrb * 16 <= nh_pad */
  /* the sorted order puts the padding rows (index >= nh) last */
  const int64_t h = jr < a.nh ? a.perm[jr] : jr;
  const int64_t hc = h < a.nh ? h : a.nh - 1;
  const RowRec rec = a.rows[hc];
  /* the rows' antenna frames (per-pixel masks on mixed tiles), held in LDS
   * rather than in 20 registers per lane: read only on mixed tiles */
  if (g == 0) {
    const double* qa = a.ant8 + (int64_t)rec.ia * 8;
    const double* qb = a.ant8 + (int64_t)rec.ib * 8;
    const int w_ = wave * 16 + r;
    frs[0][w_] = qa[0]; frs[1][w_] = qa[1]; frs[2][w_] = qa[2]; frs[3][w_] = qa[3]; frs[4][w_] = qa[7];
    frs[5][w_] = qb[0]; frs[6][w_] = qb[1]; frs[7][w_] = qb[2]; frs[8][w_] = qb[3]; frs[9][w_] = qb[7];
  }
  const int64_t t0 = a.tile0 + chunk * a.tpc;
  const int64_t tend = a.tile0 + a.ntile;
  const int64_t t1 = (t0 + a.tpc < tend) ? t0 + a.tpc : tend;
  const int ksh = a.TP == 512 ? 4 : 3; /* k-steps per tile: 1 << ksh */
  const int64_t kb0 = t0 << ksh;       /* first k-step block of the chunk */
  const int64_t K = (t1 - t0) << ksh;  /* k-steps of 32 pixels (> 0) */
  const int64_t row0 = sg * 16 * NB;

  Stage stg;
  /* thread t: coordinate (t / 32) % 4 of pixel t % 32 of a step (threads
   * 128-255 repeat 0-127, only the first store); lane (r, g): its row's
   * class byte (only g = 0 stores) */
  const int xc_ = (int)((threadIdx.x >> 5) & 3);
  const double* xsrc = xc_ == 0 ? a.l : (xc_ == 1 ? a.m : (xc_ == 2 ? a.nm1 : a.n));
  double xp;
  int xcl;
#define VW_X_LOAD(k_)                                                                   \
  do {                                                                                  \
    const int64_t p_ = ((kb0 + (k_)) << 5) + (threadIdx.x & 31);                        \
    const double v_ = xsrc[p_ < a.npix ? p_ : a.npix - 1];                              \
    xp = p_ < a.npix ? v_ : 0.0;                                                        \
    /* rows past nh carry the hidden class in the table */                             \
    const int c_ = (int)a.cls[(((kb0 + (k_)) >> ksh) - a.tile0) * a.nh_pad + h];       \
    xcl = wave_on ? c_ : VW_CLS_HIDDEN;                                                 \
  } while (0)
#define VW_X_STORE(b_)                                           \
  do {                                                           \
    if (threadIdx.x < 128) px[b_][xc_][threadIdx.x & 31] = xp;   \
    if (g == 0) pcls[b_][wave * 16 + r] = (uint8_t)xcl;          \
  } while (0)

  vw_f32x4b accR[NB], accI[NB], sumR[NB], sumI[NB];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    accR[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
    accI[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
    sumR[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
    sumI[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
  }

  /* prologue: step 0 */
  VW_X_LOAD(0);
  stg.load(a.skyh, a.skyl, nullptr, nullptr, kb0, a.nsky_pad, row0);
  VW_X_STORE(0);
  stg.store(bsh[0]);
  __syncthreads();

  for (int64_t k = 0; k < K; ++k) {
    const int buf = (int)(k & 1);
    const int64_t kn = k + 1 < K ? k + 1 : k; /* the last step reloads itself */
    stg.load(a.skyh, a.skyl, nullptr, nullptr, kb0 + kn, a.nsky_pad, row0);
    VW_X_LOAD(kn);
    const int cls = (int)pcls[buf][wave * 16 + r];
    const bool act = __ballot(cls != VW_CLS_HIDDEN) != 0;
    const bool any_mixed = __ballot(cls == VW_CLS_MIXED) != 0;
    if (act) {
      const double* pl_ = px[buf][0];
      const double* pm_ = px[buf][1];
      const double* pz_ = px[buf][2];
      const double* pn_ = px[buf][3];
      float cv[8], sv[8];
      /* two forms of the step, chosen per wave: without mixed rows a lane's
       * mask is its row's class alone (no per-term branch); with them every
       * lane evaluates both antennas' predicates, branch-free */
#define VW_PHASOR(MASKED_)                                                                              \
  do {                                                                                                  \
    _Pragma("unroll") for (int i = 0; i < 8; ++i) {                                                     \
      const int kk = g * 8 + i;                                                                         \
      const double pl = pl_[kk], pm = pm_[kk], pz = pz_[kk];                                            \
      const double x = vw_viss_phase(rec.u, rec.v, rec.w, pl, pm, pz);                                  \
      const float f = (float)__builtin_amdgcn_fract(x);                                                 \
      const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);                         \
      bool on = cls == VW_CLS_VISIBLE;                                                                  \
      if (MASKED_) {                                                                                    \
        const double pn = pn_[kk];                                                                      \
        const int va = vw_visible(pl, pm, pn, fa[0], fa[1], fa[2], fa[3], fa[4]);                       \
        const int vb = vw_visible(pl, pm, pn, fb[0], fb[1], fb[2], fb[3], fb[4]);                       \
        on = on | ((cls == VW_CLS_MIXED) & ((va & vb) != 0));                                           \
      }                                                                                                 \
      cv[i] = on ? cs : 0.0f;                                                                           \
      sv[i] = on ? sn : 0.0f;                                                                           \
    }                                                                                                   \
  } while (0)
      if (any_mixed) {
        const int w_ = wave * 16 + r;
        const double fa[5] = {frs[0][w_], frs[1][w_], frs[2][w_], frs[3][w_], frs[4][w_]};
        const double fb[5] = {frs[5][w_], frs[6][w_], frs[7][w_], frs[8][w_], frs[9][w_]};
        VW_PHASOR(true);
      } else {
        const double fa[5] = {}, fb[5] = {};
        VW_PHASOR(false);
      }
#undef VW_PHASOR
      vw_bf16x8 ach, acl, ash, asl;
      vw_split8(cv, ach, acl);
      vw_split8(sv, ash, asl);
      const __bf16* bb = bsh[buf];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        const vw_bf16x8 bh = vw_lds_bf16x8(bb, nb * 16 + r, g);
        const vw_bf16x8 bl = vw_lds_bf16x8(bb, 16 * NB + nb * 16 + r, g);
        accR[nb] = vw_mac3(ach, acl, bh, bl, accR[nb]);
        accI[nb] = vw_mac3(ash, asl, bh, bl, accI[nb]);
      }
      if ((k & 7) == 7) {
        /* blocked sums: the products of 256 pixels accumulate in the matrix
         * core's FP32 accumulators, which are then added to FP32 block sums
         * (folded across chunks in FP64): error ~ (256 + K / 256) ulp
         * rather than K ulp.  A tile's class is the same on all its steps,
         * so act holds on the whole block */
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
          sumR[nb] += accR[nb];
          sumI[nb] += accI[nb];
          accR[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
          accI[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
    /* buffers of step k - 1: every wave passed the barrier after reading
     * them (the last step writes its own reload there, unread) */
    stg.store(bsh[buf ^ 1]);
    VW_X_STORE(buf ^ 1);
    __syncthreads();
  }
#undef VW_X_LOAD
#undef VW_X_STORE
  if (!wave_on) return;
  /* D[4 g + i][r]: block row rb * 16 + 4 g + i, sky column r */
  int64_t hw[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t jj = rb * 16 + 4 * g + i;
    hw[i] = jj < a.nh ? a.perm[jj] : -1;
  }
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int64_t sky = (sg * NB + nb) * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (hw[i] >= 0) a.part[(chunk * a.nsky_pad + sky) * a.nh_pad + hw[i]] = make_double2((double)sumR[nb][i], -(double)sumI[nb][i]);
  }
}

/* chunk partials -> raw half sums [nsky][nh], chunks in ascending order;
 * passes after the first add to the running sums */
__global__ void k_vissb_fold(const double2* __restrict__ part, int64_t nchunk, int64_t nsky, int64_t nsky_pad,
                             int64_t nh, int64_t nh_pad, int accumulate, double2* __restrict__ vraw) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsky * nh) return;
  const int64_t s = k / nh, h = k - s * nh;
  double re = 0.0, im = 0.0;
  for (int64_t c = 0; c < nchunk; ++c) {
    const double2 v = part[(c * nsky_pad + s) * nh_pad + h];
    re += v.x;
    im += v.y;
  }
  if (accumulate) {
    re += vraw[k].x;
    im += vraw[k].y;
  }
  vraw[k] = make_double2(re, im);
}

/* phase correction and conjugate completion of every sky [nsky][nh] ->
 * [nsky][nh] and [nsky][2 nh] (common.hpp:255-277) */
__global__ void k_conj_fill_b(const double2* __restrict__ vraw, const RowRec* __restrict__ rows, int64_t nh, int P,
                              int64_t nsky, double2* __restrict__ vh, double2* __restrict__ vfull) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsky * nh) return;
  const int64_t s = k / nh, h = k - s * nh;
  double re, im;
  vw_phase_correct(vraw[k].x, vraw[k].y, rows[h].w, &re, &im);
  vh[k] = make_double2(re, im);
  const int64_t g = h / P, j = h - g * P;
  double2* vf = vfull + s * 2 * nh;
  vf[2 * P * g + j] = make_double2(re, im);
  vf[2 * P * g + P + j] = make_double2(re, -im);
}

/* ------------------------------------------------------------------ */
/* reconstruction                                                       */
/* ------------------------------------------------------------------ */

/* group visibilities of every sky -> packed real and imaginary parts in the
 * recon's work order, [nq_pad / 32][nsky_pad][32] hi / lo; zero outside
 * [first_nz, nuniq) and for padding skies.  blockage 1: the representative's value (occ_mode 1
 * first member, else the median member, signed for rows of the conjugate
 * half); blockage 0: the group mean over members in sorted order. */
__global__ void k_group_pack(const int32_t* __restrict__ goff, const int32_t* __restrict__ idx_sorted,
                             const double2* __restrict__ vh, int64_t nh, int P, int full, int blockage, int occ_mode,
                             const DevInfo* __restrict__ info, const int32_t* __restrict__ gperm, int64_t nq_pad,
                             int64_t nsky, int64_t nsky_pad, __bf16* __restrict__ rh, __bf16* __restrict__ rl,
                             __bf16* __restrict__ ih, __bf16* __restrict__ il) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq_pad * nsky_pad) return;
  const int64_t qb = k / (nsky_pad * 32);
  const int64_t rr = k - qb * nsky_pad * 32;
  const int64_t s = rr / 32, j = rr - s * 32;
  const int64_t q = gperm[qb * 32 + j]; /* position j of the work order holds group gperm[j] */
  float vr = 0.f, vi = 0.f;
  if (s < nsky && q >= info->first_nz && q < info->nuniq) {
    const int32_t s0 = goff[q], s1 = goff[q + 1];
    const double2* v = vh + s * nh;
    if (blockage) {
      bool neg;
      const int64_t hr = vw_row_half(idx_sorted[occ_mode == 1 ? s0 : s0 + (s1 - s0) / 2], P, full != 0, &neg);
      vr = (float)v[hr].x;
      vi = (float)(neg ? -v[hr].y : v[hr].y);
    } else {
      double sr = 0.0, si = 0.0;
      for (int32_t t = s0; t < s1; ++t) {
        bool ng;
        const int64_t hh = vw_row_half(idx_sorted[t], P, full != 0, &ng);
        sr = sr + v[hh].x;
        si = si + (ng ? -v[hh].y : v[hh].y);
      }
      vr = (float)(sr / (double)(s1 - s0));
      vi = (float)(si / (double)(s1 - s0));
    }
  }
  const __bf16 a = (__bf16)vr, b = (__bf16)vi;
  rh[k] = a;
  rl[k] = (__bf16)(vr - (float)a);
  ih[k] = b;
  il[k] = (__bf16)(vi - (float)b);
}

struct ReconBArgs {
  const double *l, *m, *n;
  int64_t pt0, npts, TP;      /* points [pt0, pt0 + npts) of this pass            */
  const uint8_t* gcls;        /* [tiles of the pass][nq_pad] classes in work order  */
  const int32_t* gperm;       /* work order: position j is group gperm[j]          */
  const double2* orec;        /* (ugu, vgu) in work order                          */
  const double* oant;         /* [position][10] representative frames, work order  */
  int64_t nq_pad;
  const GRec* rec;            /* cell centres (wavelengths)                        */
  const double* gant;         /* [q][10] representative antenna frames             */
  const DevInfo* info;
  const uint4 *vrh, *vrl, *vih, *vil; /* packed group visibilities, 16-byte units */
  int64_t nsky_pad;
  int rep;                    /* 1: representative classes; 0: every group counts  */
  int64_t nchunk, pblk, pbw;  /* group ranges, 16-point blocks, blocks of 4        */
  double* part;               /* [nchunk][nsky_pad][pstride]                       */
  uint32_t* wpart;            /* [nchunk][pstride]                                 */
  int64_t pstride;
};

/* one wave: 16 output points x (16 NB) skies over one range of groups; a
 * workgroup is four waves (64 consecutive points of one output tile).  Per
 * k-step of 32 groups the workgroup stages the packed group visibilities,
 * the groups' cell centres and their class bytes for the tile through
 * double-buffered LDS. */
#ifndef VW_RECONB_WAVES
#define VW_RECONB_WAVES 3
#endif
template <int NB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VW_RECONB_WAVES))) void k_reconb(ReconBArgs a) {
  using Stage = VwStage<4, NB>;
  __shared__ __attribute__((aligned(16))) __bf16 bsh[2][Stage::ROWS * VW_LROW];
  __shared__ double2 grec[2][32];                      /* (ugu, vgu)      */
  __shared__ __attribute__((aligned(16))) int32_t gq[2][32]; /* group index */
  __shared__ double gfr[2][320];                       /* frames, 10 per position */
  __shared__ __attribute__((aligned(16))) uint8_t gcl[2][32]; /* classes */
  const int64_t vb = vw_xcd_block(blockIdx.x, gridDim.x);
  const int64_t pbw = vb % a.pbw;
  const int64_t rest = vb / a.pbw;
  const int64_t chunk = rest % a.nchunk;
  const int64_t sg = rest / a.nchunk;
  const int wave = vw_rfl((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  const int64_t pb = pbw * 4 + wave;
  const bool wave_on = pb < a.pblk;
  const int64_t pbase = pb * 16;          /* pass-relative */
  const int64_t tile = (pbw * 64) / a.TP; /* the workgroup's 64 points lie in one tile */
  const int64_t p = pbase + r;
  const bool okp = wave_on && p < a.npts;
  const int64_t pp = a.pt0 + (okp ? p : 0);
  const vw_segparams sp = a.info->sp;
  const double pl = a.l[pp], pm = a.m[pp], pn = a.n[pp];
  const double lp = fma(sp.fa, pn, pl), mp = fma(sp.fb, pn, pm);
  const int64_t nu = a.info->nuniq, q0 = a.info->first_nz;
  /* this chunk's positions [qa, qe) of the work order (positions below nuniq
   * hold the segment's groups), aligned to 256 */
  const int64_t span = (nu + a.nchunk - 1) / a.nchunk;
  const int64_t spanA = (span + 255) & ~(int64_t)255;
  const int64_t qa = chunk * spanA;
  const int64_t qe = qa + spanA < nu ? qa + spanA : nu;
  const int64_t K = qe > qa ? (qe - qa + 31) >> 5 : 0; /* uniform over the workgroup */
  const int64_t row0 = sg * 16 * NB;

  Stage stg;
  /* threads 0-31: the group at one position of the k-step and its cell
   * centre (ugu, vgu); 32-33: 16 class bytes of the workgroup's tile; the
   * records and the class table are laid out in work order */
  double2 xr = make_double2(0.0, 0.0);
  double xf0 = 0.0, xf1 = 0.0; /* frame doubles t and t + 256 of the k-step (t + 256 < 320) */
  int32_t xq = 0;
  vw_u32x4 xc = {0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u};
#define VW_X_LOAD(k_)                                                                    \
  do {                                                                                   \
    const int t_ = (int)threadIdx.x;                                                     \
    const int64_t j0_ = qa + 32 * (k_);                                                  \
    if (t_ < 32) {                                                                       \
      /* j0 + 32 <= nq_pad: j0 < qe <= nq_pad and both are multiples of 32 */            \
      xq = a.gperm[j0_ + t_];                                                            \
      xr = a.orec[j0_ + t_];                                                             \
    } else if (t_ < 34 && a.rep) {                                                       \
      /* the lane's offset is formed here each step (opaque thread index):     \
       * held across the loop it was spilled at this kernel's register budget */ \
      int tc_ = t_ - 32;                                                                 \
      __asm__ volatile("" : "+v"(tc_));                                                  \
      xc = *(const vw_u32x4*)(a.gcls + tile * a.nq_pad + j0_ + 16 * tc_);                \
    }                                                                                    \
    if (a.rep) {                                                                         \
      xf0 = a.oant[j0_ * 10 + t_];                                                       \
      if (t_ < 64) xf1 = a.oant[j0_ * 10 + 256 + t_];                                    \
    }                                                                                    \
  } while (0)
#define VW_X_STORE(b_)                                        \
  do {                                                        \
    const int t_ = (int)threadIdx.x;                          \
    if (t_ < 32) {                                            \
      grec[b_][t_] = xr;                                      \
      gq[b_][t_] = xq;                                        \
    } else if (t_ < 34) {                                     \
      *(vw_u32x4*)(&gcl[b_][16 * (t_ - 32)]) = xc;            \
    }                                                         \
    if (a.rep) {                                              \
      gfr[b_][t_] = xf0;                                      \
      if (t_ < 64) gfr[b_][256 + t_] = xf1;                   \
    }                                                         \
  } while (0)
  vw_f32x4b acc[NB];
  double sum[NB][4];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    acc[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int i = 0; i < 4; ++i) sum[nb][i] = 0.0;
  }
  uint32_t cnt = 0;

  if (K > 0) {
    stg.load(a.vrh, a.vrl, a.vih, a.vil, qa >> 5, a.nsky_pad, row0);
    VW_X_LOAD(0);
    stg.store(bsh[0]);
    VW_X_STORE(0);
  }
  __syncthreads();
  for (int64_t k = 0; k < K; ++k) {
    const int buf = (int)(k & 1);
    const bool more = k + 1 < K;
    if (more) {
      stg.load(a.vrh, a.vrl, a.vih, a.vil, (qa >> 5) + k + 1, a.nsky_pad, row0);
      VW_X_LOAD(k + 1);
    }
    /* positions of the lane: 8 g .. 8 g + 7 of the k-step */
    const int64_t jk = qa + 32 * k + g * 8;
    const uint64_t cw = *(const uint64_t*)(&gcl[buf][g * 8]);
    const vw_u32x4 q03 = *(const vw_u32x4*)(&gq[buf][g * 8]), q47 = *(const vw_u32x4*)(&gq[buf][g * 8 + 4]);
    /* a k-step whose groups are all hidden from this tile adds nothing; one
     * whose groups are all in range, visible and of nonzero key needs no
     * per-term mask.  Lane l < 32 tests position l of the step, so that the
     * two step flags cost one test per lane and two ballots */
    const int lt = lane & 31;
    const int cl_ = (int)gcl[buf][lt];
    const bool in_ = qa + 32 * k + lt < qe;
    const bool live1 = lane < 32 && in_ && cl_ != VW_CLS_HIDDEN;
    const bool bad1 = lane < 32 && !(in_ && cl_ == VW_CLS_VISIBLE && (int64_t)gq[buf][lt] >= q0);
    if (wave_on && __ballot(live1) != 0) {
      float cv[8], sv[8];
      if (__ballot(bad1) == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const double2 u = grec[buf][g * 8 + i];
          const double x = vw_recon_phase(u.x, u.y, lp, mp);
          const float f = (float)__builtin_amdgcn_fract(x);
          cv[i] = __builtin_amdgcn_cosf(f);
          sv[i] = -__builtin_amdgcn_sinf(f);
        }
        cnt += 8u;
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int64_t q = (int32_t)(i < 4 ? q03[i & 3] : q47[i & 3]);
          const int cls = (int)((cw >> (8 * i)) & 0xffu);
          bool on = jk + i < qe && q >= q0 && cls != VW_CLS_HIDDEN;
          if (on && cls == VW_CLS_MIXED) {
            const double* d = &gfr[buf][(g * 8 + i) * 10];
            on = vw_visible(pl, pm, pn, d[0], d[1], d[2], d[3], d[4]) &&
                 vw_visible(pl, pm, pn, d[5], d[6], d[7], d[8], d[9]);
          }
          const double2 u = grec[buf][g * 8 + i];
          const double x = vw_recon_phase(u.x, u.y, lp, mp);
          const float f = (float)__builtin_amdgcn_fract(x);
          const float sn = __builtin_amdgcn_sinf(f), cs = __builtin_amdgcn_cosf(f);
          cv[i] = on ? cs : 0.0f;
          sv[i] = on ? -sn : 0.0f;
          cnt += on ? 1u : 0u;
        }
      }
      vw_bf16x8 ach, acl, ash, asl;
      vw_split8(cv, ach, acl);
      vw_split8(sv, ash, asl);
      const __bf16* bb = bsh[buf];
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
        acc[nb] = vw_mac3(ach, acl, vw_lds_bf16x8(bb, nb * 16 + r, g),
                          vw_lds_bf16x8(bb, 16 * NB + nb * 16 + r, g), acc[nb]);
        acc[nb] = vw_mac3(ash, asl, vw_lds_bf16x8(bb, 32 * NB + nb * 16 + r, g),
                          vw_lds_bf16x8(bb, 48 * NB + nb * 16 + r, g), acc[nb]);
      }
    }
    if (wave_on && ((k & 7) == 7 || !more)) {
      /* FP64 sums across blocks of 256 groups (also after skipped steps) */
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int i = 0; i < 4; ++i) sum[nb][i] += (double)acc[nb][i];
        acc[nb] = vw_f32x4b{0.f, 0.f, 0.f, 0.f};
      }
    }
    if (more) {
      stg.store(bsh[buf ^ 1]);
      VW_X_STORE(buf ^ 1);
    }
    __syncthreads();
  }
#undef VW_X_LOAD
#undef VW_X_STORE
  if (!wave_on) return;
  /* the four lanes of a point (r, r + 16, r + 32, r + 48) hold its counts */
  cnt += (uint32_t)__shfl_xor((int)cnt, 16, 64);
  cnt += (uint32_t)__shfl_xor((int)cnt, 32, 64);
  if (sg == 0 && g == 0 && p < a.npts) a.wpart[chunk * a.pstride + p] = cnt;
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int64_t sky = (sg * NB + nb) * 16 + r;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t pt = pbase + 4 * g + i;
      if (pt < a.npts) a.part[(chunk * a.nsky_pad + sky) * a.pstride + pt] = sum[nb][i];
    }
  }
}

/* group-range partials -> accumulators [nsky][npts_all] (add_inplace); the
 * sky-0 threads also fold the weights */
__global__ void k_reconb_fold(const double* __restrict__ part, const uint32_t* __restrict__ wpart, int64_t nchunk,
                              int64_t nsky, int64_t nsky_pad, int64_t pstride, int64_t pt0, int64_t npts,
                              int64_t npts_all, int weights, double* __restrict__ acc, uint32_t* __restrict__ wacc) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsky * npts) return;
  const int64_t s = k / npts, p = k - s * npts;
  double v = part[s * pstride + p];
  /* VW_FOLD_BATCH chunks' loads in flight at a time; the sum keeps chunk order */
  int64_t c = 1;
  for (; c + VW_FOLD_BATCH <= nchunk; c += VW_FOLD_BATCH) {
    double x[VW_FOLD_BATCH];
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) x[e] = part[((c + e) * nsky_pad + s) * pstride + p];
#pragma unroll
    for (int e = 0; e < VW_FOLD_BATCH; ++e) v = v + x[e];
  }
  for (; c < nchunk; ++c) v = v + part[(c * nsky_pad + s) * pstride + p];
  acc[s * npts_all + pt0 + p] += v;
  if (weights && s == 0) {
    uint32_t w = wpart[p];
    for (int64_t c2 = 1; c2 < nchunk; ++c2) w += wpart[c2 * pstride + p];
    wacc[pt0 + p] += w;
  }
}

/* final maps of every sky: acc / W (blockage 1) or acc / W_total (blockage
 * 0); the weights of sky s start at w + s * w_stride (0: one map for all) */
__global__ void k_normalize_b(const double* __restrict__ acc, const uint32_t* __restrict__ w, int64_t n, int64_t nsky,
                              int blockage, const DevInfo* __restrict__ info, const int64_t* __restrict__ scatter,
                              int64_t out_stride, double* __restrict__ out, int64_t w_stride) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nsky) return;
  const int64_t s = k / n, i = k - s * n;
  double v;
  const uint32_t wi = w[s * w_stride + i];
  if (blockage) v = wi > 0 ? acc[k] / (double)wi : 0.0;
  else {
    const int64_t wt = info->wtotal;
    v = acc[k] * (wt > 0 ? 1.0 / (double)wt : 0.0);
  }
  out[s * out_stride + (scatter ? scatter[i] : i)] = v;
}

/* per-sky weights of the lm-grid points [nsky][n] onto the image grid
 * [nsky][out_stride] (points outside the disk stay as cleared) */
__global__ void k_scatter_w(const uint32_t* __restrict__ w, int64_t n, int64_t nsky, const int64_t* __restrict__ scatter,
                            int64_t out_stride, uint32_t* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nsky) return;
  const int64_t s = k / n, i = k - s * n;
  out[s * out_stride + scatter[i]] = w[k];
}

/* every sky's phase-corrected half visibilities of the current segment
 * [nsky][nh] into the day buffer [nsky][day_rows] at row offset row0 */
__global__ void k_vday_store(const double2* __restrict__ vh, int64_t nsky, int64_t nh, int64_t row0, int64_t day_rows,
                             double2* __restrict__ vday) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsky * nh) return;
  const int64_t s = k / nh, h = k - s * nh;
  vday[s * day_rows + row0 + h] = vh[k];
}

/* zero fill of 32-bit words (kept out of hipMemsetAsync so that the batched
 * step can be captured in a graph: memset nodes are not replay-safe here) */
__global__ void k_fill0(uint32_t* __restrict__ p, int64_t n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

/* ------------------------------------------------------------------ */
/* stable LSD radix sort without memsets (graph-capturable)             */
/* ------------------------------------------------------------------ */
/* One pass sorts by the 8-bit digit (key >> shift) & 255 of blocks of
 * VW_RS_BLOCK elements: k_rs_hist counts each block's digits (LDS counters),
 * k_rs_scan turns the digit-major table [256][nblk] into exclusive offsets
 * within each digit plus the digit totals, k_rs_scatter places each block's
 * elements at the digit's base (the scan of the totals) plus the block's
 * offset: lane b walks the block's digits in order and moves the elements of
 * digit b to consecutive positions, so equal digits keep their input order
 * (stable). */
#define VW_RS_BLOCK 128 /* short serial walks in the scatter (one lane per digit) */
template <typename K>
__global__ __launch_bounds__(256) void k_rs_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                 int32_t* __restrict__ hist, int64_t nblk) {
  __shared__ int32_t h[256];
  const int64_t b = blockIdx.x;
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i0 = b * VW_RS_BLOCK, i1 = i0 + VW_RS_BLOCK < n ? i0 + VW_RS_BLOCK : n;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) atomicAdd(&h[(int)((keys[i] >> shift) & 255)], 1);
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblk + b] = h[threadIdx.x];
}
/* exclusive scan of each digit's row of the digit-major table [256][nblk]
 * in place, one workgroup per digit (a lane takes a contiguous run of the
 * row, the run totals are scanned in LDS); the digit's total goes to
 * tot[digit] */
__global__ __launch_bounds__(256) void k_rs_scan(int32_t* __restrict__ a, int64_t nblk, int32_t* __restrict__ tot) {
  __shared__ int32_t t[256];
  int32_t* row = a + (int64_t)blockIdx.x * nblk;
  const int64_t per = (nblk + 255) / 256;
  const int64_t i0 = threadIdx.x * per, i1 = i0 + per < nblk ? i0 + per : nblk;
  int32_t s = 0;
  for (int64_t i = i0; i < i1; ++i) s += row[i];
  t[threadIdx.x] = s;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int32_t v = threadIdx.x >= (unsigned)o ? t[threadIdx.x - o] : 0;
    __syncthreads();
    t[threadIdx.x] += v;
    __syncthreads();
  }
  int32_t run = threadIdx.x ? t[threadIdx.x - 1] : 0;
  for (int64_t i = i0; i < i1; ++i) {
    const int32_t c = row[i];
    row[i] = run;
    run += c;
  }
  if (threadIdx.x == 255) tot[blockIdx.x] = t[255];
}
template <typename K, typename V>
__global__ __launch_bounds__(256) void k_rs_scatter(const K* __restrict__ kin, const V* __restrict__ vin, int64_t n,
                                                    int shift, const int32_t* __restrict__ offs, int64_t nblk,
                                                    const int32_t* __restrict__ tot, K* __restrict__ kout,
                                                    V* __restrict__ vout) {
  __shared__ uint8_t d[VW_RS_BLOCK];
  __shared__ int32_t base[256];
  const int64_t b = blockIdx.x;
  const int64_t i0 = b * VW_RS_BLOCK;
  const int cnt = (int)(i0 + VW_RS_BLOCK < n ? VW_RS_BLOCK : n - i0);
  for (int j = threadIdx.x; j < cnt; j += 256) d[j] = (uint8_t)((kin[i0 + j] >> shift) & 255);
  /* digit bases: exclusive scan of the digit totals */
  base[threadIdx.x] = tot[threadIdx.x];
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    const int32_t v = threadIdx.x >= (unsigned)o ? base[threadIdx.x - o] : 0;
    __syncthreads();
    base[threadIdx.x] += v;
    __syncthreads();
  }
  const int bin = threadIdx.x;
  int64_t pos = (bin ? base[bin - 1] : 0) + offs[(int64_t)bin * nblk + b];
  for (int j = 0; j < cnt; ++j)
    if (d[j] == bin) {
      kout[pos] = kin[i0 + j];
      vout[pos] = vin[i0 + j];
      ++pos;
    }
}
template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

#endif
