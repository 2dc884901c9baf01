/*
 * vw_unet.hip -- the learned deconvolver (include/vw_unet.h): U-Net-lite
 * forward, backward and Adam in hand-written gfx950 kernels, orchestrated
 * by a small C++ layer table.  No counterpart in the reference.
 *
 * Tensors are NCHW with square images.  A 3x3 convolution workgroup computes
 * a 16 x 16 output tile for 16 output channels of one image: each lane owns
 * one pixel and 16 FP32 accumulators, 16 input channels at a time are staged
 * with their 1-pixel halo in LDS (18 x 18 each), and the weights, uniform
 * across the workgroup, come through scalar loads.  The same kernel computes
 * the data gradient with the transposed, flipped weights.  Weight gradients
 * are register-blocked over 16 x 16 channel pairs per workgroup and summed
 * over pixel chunks whose partials are folded in a fixed order
 * (deterministic).
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/vissws.h"
#include "../../include/vw_unet.h"

#define VW_API extern "C" __attribute__((visibility("default")))

/* Staging of vw_unet_predict_noalloc: FP32 images in and out of the device
 * in a buffer of the library's own, so that entry point allocates nothing.
 * 16 MiB (4,194,304 values: 64 images of 256 x 256), held by every device the
 * library runs on until the process ends.  In host code this name is not a
 * device address: vw_unet_create resolves the address on its device with
 * hipGetSymbolAddress, and only that address is used. */
#define VW_UNET_SCRATCH_FLOATS 4194304
static __device__ __align__(256) float scratch[VW_UNET_SCRATCH_FLOATS];

namespace {

thread_local std::string g_unet_err;

/* ------------------------------------------------------------------ */
/* kernels                                                              */
/* ------------------------------------------------------------------ */

/* 3x3 convolution, zero padding.  TR = false: out[o] = B[o] + sum_i W[o][i] * in[i]
 * (layer weights [Co][Ci][3][3], cin = Ci, cout = Co), then ReLU if RELU.
 * TR = true: the data gradient of such a layer, out[i] = sum_o W[o][i] (flipped) * in[o]
 * (cin = layer Co, cout = layer Ci). */
template <typename T, bool TR, bool RELU>
__global__ __launch_bounds__(256) void k_conv3(const T* __restrict__ in, const float* __restrict__ W,
                                               const float* __restrict__ B, T* __restrict__ out, int cin, int cout,
                                               int H, const T* __restrict__ mask) {
  __shared__ float tile[16][18][18];
  const int tpr = (H + 15) / 16;
  const int ty0 = (blockIdx.x / tpr) * 16, tx0 = (blockIdx.x % tpr) * 16;
  const int ob = blockIdx.y * 16;
  const int n = blockIdx.z;
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  const int64_t HH = (int64_t)H * H;
  float acc[16];
#pragma unroll
  for (int o = 0; o < 16; ++o) acc[o] = 0.f;
  for (int c0 = 0; c0 < cin; c0 += 16) {
    const int cn = cin - c0 < 16 ? cin - c0 : 16;
    __syncthreads();
    for (int idx = threadIdx.x; idx < cn * 324; idx += 256) {
      const int c = idx / 324, r = idx - c * 324;
      const int yy = r / 18, xx = r - yy * 18;
      const int gy = ty0 + yy - 1, gx = tx0 + xx - 1;
      float v = 0.f;
      if (gy >= 0 && gy < H && gx >= 0 && gx < H) v = (float)in[((int64_t)n * cin + c0 + c) * HH + (int64_t)gy * H + gx];
      tile[c][yy][xx] = v;
    }
    __syncthreads();
    for (int c = 0; c < cn; ++c) {
      float v[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) v[k] = tile[c][ty + k / 3][tx + k % 3];
#pragma unroll
      for (int o = 0; o < 16; ++o) {
        if (ob + o < cout) { /* uniform */
          const float* w =
              TR ? W + ((int64_t)(c0 + c) * cout + ob + o) * 9 : W + ((int64_t)(ob + o) * cin + c0 + c) * 9;
#pragma unroll
          for (int k = 0; k < 9; ++k) acc[o] = fmaf(TR ? w[8 - k] : w[k], v[k], acc[o]);
        }
      }
    }
  }
  const int y = ty0 + ty, x = tx0 + tx;
  if (y >= H || x >= H) return;
#pragma unroll
  for (int o = 0; o < 16; ++o) {
    if (ob + o < cout) {
      float r = acc[o] + (B ? B[ob + o] : 0.f);
      if (RELU) r = r > 0.f ? r : 0.f;
      const int64_t oi = ((int64_t)n * cout + ob + o) * HH + (int64_t)y * H + x;
      if (mask && !((float)mask[oi] > 0.f)) r = 0.f;
      out[oi] = (T)r;
    }
  }
}

/* 3x3 convolution on the FP16 matrix cores (FP16 activations, cin a
 * multiple of 16): an implicit GEMM D[o][pixel] = sum_K Wk[o][K] X[K][pixel]
 * with K = tap * cin + i (tap-major).  Wk is the layer's weights re-laid in
 * FP16 as [o][tap][i] (forward) or the flipped transpose (data gradient), so
 * one kernel serves both.  A workgroup owns 16 output channels x a 16 x 16
 * pixel tile of one image; each wave four pixel rows (four 16-pixel column
 * blocks of v_mfma_f32_16x16x32_f16).  Input channels pass through LDS in
 * chunks of 32 (16 when cin is 16), channel-last, so a lane's 8 K-values (one
 * tap, 8 consecutive channels) are one 16-byte read. */
typedef _Float16 vw_h8 __attribute__((ext_vector_type(8)));
typedef float vw_f4u __attribute__((ext_vector_type(4)));
/* logical workgroup index with consecutive indices on one XCD (workgroups
 * are dispatched to the 8 XCDs round-robin) */
__device__ __forceinline__ int64_t vu_xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, r = nb % 8, x = b % 8, i = b / 8;
  return x * q + (x < r ? x : r) + i;
}
typedef _Float16 vw_h2 __attribute__((ext_vector_type(2)));
template <bool RELU, int CB, bool PF, bool VS, int OCB>
__global__ __launch_bounds__(256) void k_conv3_mfma(const _Float16* __restrict__ in, const _Float16* __restrict__ Wk,
                                                    const float* __restrict__ B, _Float16* __restrict__ out, int cin,
                                                    int cout, int H, const _Float16* __restrict__ mask) {
  static_assert(CB == 16 || CB == 32, "channels per LDS chunk");
  constexpr int STEPS = CB == 32 ? 9 : 5;
  /* [y][x][ci], pixel rows of CB channels + 8 halves of padding (the 16
   * pixels a 16-lane read group takes then fall on disjoint banks); 16-channel
   * chunks need only 15.5 KB, which lets 8 workgroups share a CU */
  constexpr int LR = CB + 8;
  __shared__ __attribute__((aligned(16))) _Float16 tile[18 * 18 * LR];
  static_assert(18 * 18 * LR >= 16 * 16 * 24, "output staging area");
  const int tpr = (H + 15) / 16;
  /* one-dimensional grid, workgroups ordered (output-channel block, tile x,
   * tile y, image) from fastest and mapped so that consecutive ones run on
   * one XCD: the blocks sharing an input tile, and the tiles sharing its
   * cache lines, meet in one L2 */
  const int nob = (cout + 16 * OCB - 1) / (16 * OCB); /* blocks of 16 OCB output channels */
  const int vb = (int)vu_xcd_block(blockIdx.x, gridDim.x);
  const int ob = (vb % nob) * 16 * OCB;
  const int tl = (vb / nob) % (tpr * tpr);
  const int n = vb / (nob * tpr * tpr);
  const int ty0 = (tl / tpr) * 16, tx0 = (tl % tpr) * 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  const int64_t HH = (int64_t)H * H;
  const int K = 9 * cin;
  constexpr int cb = CB; /* channels per LDS chunk (a divisor of cin) */
  /* acc[b][j]: output channels ob + 16 b .., pixel row 4 wave + j */
  vw_f4u acc[OCB][4];
#pragma unroll
  for (int b = 0; b < OCB; ++b)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[b][j] = vw_f4u{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < cin; c0 += cb) {
    /* the chunk's weight operands (k-steps: (tap, 32 channels) or, with
     * 16-channel chunks, (tap, 16 channels) for lanes g < 2 and the next tap
     * for g >= 2); with PF all of them are issued before the staging loads so
     * their latency overlaps it (for launches too small to hide it by
     * occupancy: PF costs registers), without PF one per k-step.  VS stages
     * the tile with 16-byte loads of channel pairs (the element-wise form is
     * kept for H % 8 != 0). */
    vw_h8 aw[PF ? STEPS : 1];
    if constexpr (PF) {
#pragma unroll
      for (int ks = 0; ks < STEPS; ++ks) {
        const int tap = CB == 32 ? ks : 2 * ks + (g >> 1);
        const int ci0 = CB == 32 ? 8 * g : 8 * (g & 1);
        aw[ks] = vw_h8{};
        if (tap < 9 && ob + r < cout) aw[ks] = *(const vw_h8*)(Wk + ((int64_t)(ob + r) * K + tap * cin + c0 + ci0));
      }
    }
    __syncthreads();
    if (VS && H % 8 == 0) {
      /* interior columns: a lane takes a channel pair and 8 pixels of one row
       * (two aligned 16-byte loads) and writes each pixel's pair as one dword;
       * lanes run over the pairs first, so a write phase covers consecutive
       * dwords */
      const int npair = cb >> 1;
      for (int idx = threadIdx.x; idx < npair * 36; idx += 256) {
        const int cp = idx % npair, rest = idx / npair, yy = rest >> 1, hx = rest & 1;
        const int gy = ty0 + yy - 1, gx = tx0 + 8 * hx;
        vw_h8 v0 = {}, v1 = {};
        if (gy >= 0 && gy < H && gx < H) {
          const _Float16* s0 = in + ((int64_t)n * cin + c0 + 2 * cp) * HH + (int64_t)gy * H + gx;
          v0 = *(const vw_h8*)s0;
          v1 = *(const vw_h8*)(s0 + HH);
        }
        _Float16* d = tile + (yy * 18 + 1 + 8 * hx) * LR + 2 * cp;
#pragma unroll
        for (int i = 0; i < 8; ++i) *(vw_h2*)(d + i * LR) = vw_h2{v0[i], v1[i]};
      }
      /* halo columns tx0 - 1 and tx0 + 16 */
      for (int idx = threadIdx.x; idx < cb * 36; idx += 256) {
        const int c = idx % cb, rest = idx / cb, yy = rest >> 1, side = rest & 1;
        const int gy = ty0 + yy - 1, gx = side ? tx0 + 16 : tx0 - 1;
        _Float16 v = (_Float16)0.f;
        if (gy >= 0 && gy < H && gx >= 0 && gx < H) v = in[((int64_t)n * cin + c0 + c) * HH + (int64_t)gy * H + gx];
        tile[(yy * 18 + (side ? 17 : 0)) * LR + c] = v;
      }
    } else {
      for (int idx = threadIdx.x; idx < 324 * cb; idx += 256) {
        const int c = idx / 324, q = idx - c * 324, yy = q / 18, xx = q - yy * 18;
        const int gy = ty0 + yy - 1, gx = tx0 + xx - 1;
        _Float16 v = (_Float16)0.f;
        if (gy >= 0 && gy < H && gx >= 0 && gx < H) v = in[((int64_t)n * cin + c0 + c) * HH + (int64_t)gy * H + gx];
        tile[(yy * 18 + xx) * LR + c] = v;
      }
    }
    __syncthreads();
#define VW_CONV_KSTEP(av_)                                                                      \
  do {                                                                                          \
    const int tap = CB == 32 ? ks : 2 * ks + (g >> 1);                                          \
    const int ci0 = CB == 32 ? 8 * g : 8 * (g & 1);                                             \
    const bool kon = tap < 9;                                                                   \
    const int ky = kon ? tap / 3 : 0, kx = kon ? tap % 3 : 0;                                   \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                             \
      const int py = 4 * wave + j; /* tile row of this column block */                          \
      vw_h8 bv = *(const vw_h8*)(tile + ((py + ky) * 18 + r + kx) * LR + ci0);                  \
      if (!kon) bv = vw_h8{};                                                                   \
      _Pragma("unroll") for (int b = 0; b < OCB; ++b)                                            \
        acc[b][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av_[b], bv, acc[b][j], 0, 0, 0);      \
    }                                                                                           \
  } while (0)
    if constexpr (PF) {
#pragma unroll
      for (int ks = 0; ks < STEPS; ++ks) {
        const vw_h8 av1[1] = {aw[ks]};
        VW_CONV_KSTEP(av1);
      }
    } else if constexpr (OCB == 1) {
#pragma unroll 1
      for (int ks = 0; ks < STEPS; ++ks) {
        const int tap = CB == 32 ? ks : 2 * ks + (g >> 1);
        const int ci0 = CB == 32 ? 8 * g : 8 * (g & 1);
        vw_h8 av[1] = {{}};
        if (tap < 9 && ob + r < cout) av[0] = *(const vw_h8*)(Wk + ((int64_t)(ob + r) * K + tap * cin + c0 + ci0));
        VW_CONV_KSTEP(av);
      }
    } else {
      /* OCB weight operands per k-step, one per 16-channel block; unrolled
       * so that their loads are issued ahead of the products */
#pragma unroll
      for (int ks = 0; ks < STEPS; ++ks) {
        const int tap = CB == 32 ? ks : 2 * ks + (g >> 1);
        const int ci0 = CB == 32 ? 8 * g : 8 * (g & 1);
        vw_h8 av[OCB];
#pragma unroll
        for (int b = 0; b < OCB; ++b) {
          av[b] = vw_h8{};
          const int o = ob + 16 * b + r;
          if (tap < 9 && o < cout) av[b] = *(const vw_h8*)(Wk + ((int64_t)o * K + tap * cin + c0 + ci0));
        }
        VW_CONV_KSTEP(av);
      }
    }
#undef VW_CONV_KSTEP
  }
  /* D[4 g + i][r]: output channel ob + 4 g + i, pixel (4 wave + j, r).  With
   * VS (H % 8 == 0: whole 16-pixel rows) the tile goes through LDS as
   * [16 channels][16 rows][16 + 8 pixels] halves and leaves as 16-byte
   * stores of 8 pixels (with 16-byte mask loads), instead of 2-byte stores
   * of 16 lanes per channel row. */
  if (VS && H % 8 == 0) {
    _Float16* ob_ = tile;
#pragma unroll
    for (int b = 0; b < OCB; ++b) {
      const int obb = ob + 16 * b;
      __syncthreads(); /* every wave is done with the input tile / the last block's stores */
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int o = obb + 4 * g + i;
          float v = acc[b][j][i] + ((B && o < cout) ? B[o] : 0.f);
          if (RELU) v = v > 0.f ? v : 0.f;
          ob_[((4 * g + i) * 16 + 4 * wave + j) * 24 + r] = (_Float16)v;
        }
      __syncthreads();
      /* 16 channels x 16 rows x 2 halves of 8 pixels: 512 stores, 2 per thread */
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int e = (int)threadIdx.x + 256 * q;
        const int hx = e & 1, yy = (e >> 1) & 15, oc = e >> 5;
        const int o = obb + oc, y = ty0 + yy, x = tx0 + 8 * hx;
        if (o >= cout || y >= H || x >= H) continue;
        vw_h8 v = *(const vw_h8*)(ob_ + (oc * 16 + yy) * 24 + 8 * hx);
        const int64_t oi = ((int64_t)n * cout + o) * HH + (int64_t)y * H + x;
        if (mask) {
          const vw_h8 mk = *(const vw_h8*)(mask + oi);
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (!((float)mk[k] > 0.f)) v[k] = (_Float16)0.f;
        }
        *(vw_h8*)(out + oi) = v;
      }
    }
    return;
  }
#pragma unroll
  for (int b = 0; b < OCB; ++b)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int y = ty0 + 4 * wave + j, x = tx0 + r;
      if (y >= H || x >= H) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = ob + 16 * b + 4 * g + i;
        if (o >= cout) continue;
        float v = acc[b][j][i] + (B ? B[o] : 0.f);
        if (RELU) v = v > 0.f ? v : 0.f;
        const int64_t oi = ((int64_t)n * cout + o) * HH + (int64_t)y * H + x;
        if (mask && !((float)mask[oi] > 0.f)) v = 0.f;
        out[oi] = (_Float16)v;
      }
    }
}

/* FP16 weight layouts for the matrix-core convolution, all 3x3 layers in
 * one launch: forward [co][tap][ci], data gradient [ci][tap][co] with the taps
 * flipped; base = prefix sums of the layers' weight counts */
struct WLayoutJobs {
  int64_t base[15];
  int64_t w[14], hf[14], hb[14];
  int ci[14], co[14];
};
__global__ void k_wlayout(WLayoutJobs J, const float* __restrict__ W, _Float16* __restrict__ wh) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= J.base[14]) return;
  int j = 0;
  while (j < 13 && t >= J.base[j + 1]) ++j;
  const int64_t i = t - J.base[j];
  const int64_t ci_n = J.ci[j], co_n = J.co[j];
  const int64_t co = i / (ci_n * 9), rem = i - co * ci_n * 9, ci = rem / 9, tap = rem - ci * 9;
  const _Float16 w = (_Float16)W[J.w[j] + i];
  wh[J.hf[j] + (co * 9 + tap) * ci_n + ci] = w;
  wh[J.hb[j] + (ci * 9 + (8 - tap)) * co_n + co] = w;
}

template <typename T>
__device__ __forceinline__ float vw_block_sum(float v, float* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}

/* pixels per chunk of the output layer's weight-gradient reduction */
#define VW_WG_CHUNK 16384

/* weight and bias gradients of a 3x3 layer: a workgroup owns 16 output x 16
 * input channels and one chunk of 16 x 16 pixel tiles; per tile it stages dz
 * of its output channels (16 x 16) and the input channels with their halo
 * (18 x 18) in LDS, and each lane accumulates 2 output x 2 input channels x
 * 9 taps (8 LDS reads per 36 FMAs), over a quarter of the tile's rows.
 * Partials [chunk][co][ci][10] (9 weights, then the bias term from the
 * ci = 0 lanes), folded in chunk order by k_fold_stage1/2. */
#define VW_WG_TILES 16 /* 16 x 16 pixel tiles per chunk */
template <typename T>
__global__ __launch_bounds__(256) void k_conv3_wgrad_t(const T* __restrict__ in, const T* __restrict__ dz,
                                                       float* __restrict__ part, int N, int ci_n, int co_n, int H) {
  /* one arena: the staged tiles (16 x 18 x 18 inputs, 16 x 16 x 17 gradients)
   * and, at the end, the quarter partials (3 x 64 lanes x 38) */
  __shared__ float arena[16 * 324 + 16 * 16 * 17];
  static_assert(3 * 64 * 38 <= 16 * 324 + 16 * 16 * 17, "reduction area");
  float (*sx)[18][18] = reinterpret_cast<float (*)[18][18]>(arena);
  float (*sg)[16][17] = reinterpret_cast<float (*)[16][17]>(arena + 16 * 324);
  const int cib = blockIdx.x * 16, cob = blockIdx.y * 16;
  const int tpr = (H + 15) / 16;
  const int64_t ntiles = (int64_t)N * tpr * tpr;
  const int64_t t0 = (int64_t)blockIdx.z * VW_WG_TILES;
  const int64_t t1 = t0 + VW_WG_TILES < ntiles ? t0 + VW_WG_TILES : ntiles;
  const int tid = threadIdx.x;
  const int oi = (tid >> 3) & 7, ii = tid & 7;  /* lane's channel pairs: co cob + 2 oi + {0,1}, ci cib + 2 ii + {0,1} */
  const int half = tid >> 6;                     /* 0..3: which quarter of the tile's rows the lane sums */
  const int64_t HH = (int64_t)H * H;
  float acc[2][2][9], ab[2] = {0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int k = 0; k < 9; ++k) acc[a][b][k] = 0.f;
  for (int64_t t = t0; t < t1; ++t) {
    const int n = (int)(t / (tpr * tpr));
    const int r = (int)(t - (int64_t)n * tpr * tpr);
    const int y0 = (r / tpr) * 16, x0 = (r % tpr) * 16;
    __syncthreads();
    for (int idx = tid; idx < 16 * 324; idx += 256) {
      const int c = idx / 324, q = idx - c * 324, yy = q / 18, xx = q - yy * 18;
      const int gy = y0 + yy - 1, gx = x0 + xx - 1;
      float v = 0.f;
      if (cib + c < ci_n && gy >= 0 && gy < H && gx >= 0 && gx < H)
        v = (float)in[((int64_t)n * ci_n + cib + c) * HH + (int64_t)gy * H + gx];
      sx[c][yy][xx] = v;
    }
    for (int idx = tid; idx < 16 * 256; idx += 256) {
      const int c = idx >> 8, q = idx & 255, yy = q >> 4, xx = q & 15;
      const int gy = y0 + yy, gx = x0 + xx;
      float v = 0.f;
      if (cob + c < co_n && gy < H && gx < H) v = (float)dz[((int64_t)n * co_n + cob + c) * HH + (int64_t)gy * H + gx];
      sg[c][yy][xx] = v;
    }
    __syncthreads();
    /* the lane sums rows 4 half .. 4 half + 3 of the tile (the four quarters
     * are added at the end), every column */
#pragma unroll 1
    for (int yy = 4 * half; yy < 4 * half + 4; ++yy) {
      float w0[2][3], w1[2][3], w2[2][3]; /* 3 x 3 input window rows for the two input channels */
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int kx = 0; kx < 2; ++kx) {
          w0[b][kx + 1] = sx[2 * ii + b][yy][kx];
          w1[b][kx + 1] = sx[2 * ii + b][yy + 1][kx];
          w2[b][kx + 1] = sx[2 * ii + b][yy + 2][kx];
        }
#pragma unroll 2
      for (int xx = 0; xx < 16; ++xx) {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          w0[b][0] = w0[b][1]; w0[b][1] = w0[b][2]; w0[b][2] = sx[2 * ii + b][yy][xx + 2];
          w1[b][0] = w1[b][1]; w1[b][1] = w1[b][2]; w1[b][2] = sx[2 * ii + b][yy + 1][xx + 2];
          w2[b][0] = w2[b][1]; w2[b][1] = w2[b][2]; w2[b][2] = sx[2 * ii + b][yy + 2][xx + 2];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const float gv = sg[2 * oi + a][yy][xx];
          if (ii == 0) ab[a] += gv;
#pragma unroll
          for (int b = 0; b < 2; ++b) {
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
              acc[a][b][kx] = fmaf(gv, w0[b][kx], acc[a][b][kx]);
              acc[a][b][3 + kx] = fmaf(gv, w1[b][kx], acc[a][b][3 + kx]);
              acc[a][b][6 + kx] = fmaf(gv, w2[b][kx], acc[a][b][6 + kx]);
            }
          }
        }
      }
    }
  }
  /* fold the four row quarters through the arena, lanes of quarter 0 write */
  __syncthreads();
  float* red = arena;
  const int lane64 = tid & 63;
  if (half > 0) {
    float* d = red + ((half - 1) * 64 + lane64) * 38;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int k = 0; k < 9; ++k) d[(a * 2 + b) * 9 + k] = acc[a][b][k];
    d[36] = ab[0];
    d[37] = ab[1];
  }
  __syncthreads();
  if (half == 0) {
    for (int h = 0; h < 3; ++h) {
      const float* d = red + (h * 64 + lane64) * 38;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int k = 0; k < 9; ++k) acc[a][b][k] += d[(a * 2 + b) * 9 + k];
      ab[0] += d[36];
      ab[1] += d[37];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int co = cob + 2 * oi + a;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int ci = cib + 2 * ii + b;
        if (co < co_n && ci < ci_n) {
          float* dst = part + (((int64_t)blockIdx.z * co_n + co) * ci_n + ci) * 10;
#pragma unroll
          for (int k = 0; k < 9; ++k) dst[k] = acc[a][b][k];
          if (ci == 0) dst[9] = ab[a];
        }
      }
    }
  }
}

/* weight and bias gradients of a 3x3 layer on the FP16 matrix cores:
 * dW[co][ci][tap] = sum over pixels p of dz[co][p] in[ci][p + tap] is a GEMM
 * with M = 16 output channels, N = 16 input channels, K = pixels, and one
 * accumulator per tap.  A workgroup owns a 16 x 16 channel block and a run of
 * tpc tiles of 16 x 32 pixels.  Per tile it stages dz (16 rows) and the input
 * with its halo (18 rows) in LDS as 16-byte chunks of 8 consecutive pixels,
 * chunk-major over the 16 channels (the 16 lanes of a read phase take 256
 * consecutive bytes); the input is stored three times, shifted by -1, 0 and +1
 * columns, so every operand is one aligned 16-byte read.  Each wave sums four
 * tile rows, a K-step being one row of 32 pixels: A = dz of the row (lane:
 * output channel lane % 16, chunk lane / 16), B = the shifted input rows
 * above, at and below it; the six input rows a wave needs are each read once
 * and applied to every (row, ky) they serve.  Partials [chunk][co][ci][10] as
 * for k_conv3_wgrad_t.  Requires H % 8 == 0 and co_n % 16 == 0; input
 * channels past ci_n are staged as zeros (the first layer has one). */
#define VW_WGM_SX (3 * 18 * 4 * 16) /* input chunks [kx][18 rows][4][16 ci] */
#define VW_WGM_SG (16 * 4 * 16)     /* gradient chunks [16 rows][4][16 co] */
__global__ __launch_bounds__(256) void k_conv3_wgrad_mfma(const _Float16* __restrict__ in,
                                                          const _Float16* __restrict__ dz, float* __restrict__ part,
                                                          int N, int ci_n, int co_n, int H, int tpc) {
  __shared__ vw_h8 arena[VW_WGM_SX + VW_WGM_SG];
  static_assert(3 * 64 * 37 * 4 <= (VW_WGM_SX + VW_WGM_SG) * 16, "reduction area");
  vw_h8* sx = arena;
  vw_h8* sg = arena + VW_WGM_SX;
  /* one-dimensional grid ordered (input-channel block, output-channel
   * block, chunk) from fastest, consecutive workgroups on one XCD: the
   * blocks of a chunk read the same input and gradient tiles */
  const int nib = (ci_n + 15) / 16, nob = co_n / 16;
  const int vb = (int)vu_xcd_block(blockIdx.x, gridDim.x);
  const int bz = vb / (nib * nob);
  const int cib = (vb % nib) * 16, cob = ((vb / nib) % nob) * 16;
  const int tpy = (H + 15) / 16, tpx = (H + 31) / 32;
  const int64_t ntiles = (int64_t)N * tpy * tpx;
  const int64_t t0 = (int64_t)bz * tpc;
  const int64_t t1 = t0 + tpc < ntiles ? t0 + tpc : ntiles;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r = lane & 15, g = lane >> 4;
  const int64_t HH = (int64_t)H * H;
  vw_f4u acc[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) acc[k] = vw_f4u{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  /* a tile's rows are loaded into registers one tile ahead (issued before
   * the products of the current tile) and written to LDS at the next step:
   * dz as lane -> (channel tid % 16, tile row tid / 16); the input as 288
   * (channel, halo row) items, columns x0 - 1 .. x0 + 32, item tid and, for
   * tid < 32, item tid + 256 */
  vw_h8 gvr[4], xva[4], xvb[4];
  _Float16 hla = (_Float16)0.f, hra = (_Float16)0.f, hlb = (_Float16)0.f, hrb = (_Float16)0.f;
#define VW_WGM_LDX(item_, xv_, hl_, hr_)                                                  \
  do {                                                                                    \
    const int c_ = (item_) & 15, gy_ = y0_ + ((item_) >> 4) - 1;                          \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) xv_[j] = vw_h8{};                       \
    hl_ = (_Float16)0.f;                                                                  \
    hr_ = (_Float16)0.f;                                                                  \
    if (gy_ >= 0 && gy_ < H && cib + c_ < ci_n) {                                         \
      const _Float16* src_ = in + ((int64_t)n_ * ci_n + cib + c_) * HH + (int64_t)gy_ * H + x0_; \
      _Pragma("unroll") for (int j = 0; j < 4; ++j) if (x0_ + 8 * j < H) xv_[j] = *(const vw_h8*)(src_ + 8 * j); \
      if (x0_ > 0) hl_ = src_[-1];                                                        \
      if (x0_ + 32 < H) hr_ = src_[32];                                                   \
    }                                                                                     \
  } while (0)
#define VW_WGM_LOAD(t_)                                                                   \
  do {                                                                                    \
    const int n_ = (int)((t_) / (tpy * tpx));                                             \
    const int rr_ = (int)((t_) - (int64_t)n_ * tpy * tpx);                                \
    const int y0_ = (rr_ / tpx) * 16, x0_ = (rr_ % tpx) * 32;                             \
    const int gyd_ = y0_ + (tid >> 4);                                                    \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                       \
      gvr[j] = vw_h8{};                                                                   \
      if (gyd_ < H && x0_ + 8 * j < H)                                                    \
        gvr[j] = *(const vw_h8*)(dz + ((int64_t)n_ * co_n + cob + (tid & 15)) * HH + (int64_t)gyd_ * H + x0_ + 8 * j); \
    }                                                                                     \
    VW_WGM_LDX(tid, xva, hla, hra);                                                       \
    if (tid < 32) VW_WGM_LDX(tid + 256, xvb, hlb, hrb);                                   \
  } while (0)
#define VW_WGM_STX(item_, xv_, hl_, hr_)                                                  \
  do {                                                                                    \
    const int c_ = (item_) & 15, h_ = (item_) >> 4;                                       \
    vw_h8 left_ = {}, right_ = {};                                                        \
    left_[7] = hl_;                                                                       \
    right_[0] = hr_;                                                                      \
    _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                       \
      const vw_h8 prev_ = j > 0 ? xv_[j > 0 ? j - 1 : 0] : left_;                         \
      const vw_h8 next_ = j < 3 ? xv_[j < 3 ? j + 1 : 3] : right_;                        \
      sx[((0 * 18 + h_) * 4 + j) * 16 + c_] = __builtin_shufflevector(prev_, xv_[j], 7, 8, 9, 10, 11, 12, 13, 14); \
      sx[((1 * 18 + h_) * 4 + j) * 16 + c_] = xv_[j];                                     \
      sx[((2 * 18 + h_) * 4 + j) * 16 + c_] = __builtin_shufflevector(xv_[j], next_, 1, 2, 3, 4, 5, 6, 7, 8); \
    }                                                                                     \
  } while (0)
  if (t0 < t1) VW_WGM_LOAD(t0);
  for (int64_t t = t0; t < t1; ++t) {
    __syncthreads(); /* every wave has finished reading the previous tile */
#pragma unroll
    for (int j = 0; j < 4; ++j) sg[((tid >> 4) * 4 + j) * 16 + (tid & 15)] = gvr[j];
    VW_WGM_STX(tid, xva, hla, hra);
    if (tid < 32) VW_WGM_STX(tid + 256, xvb, hlb, hrb);
    __syncthreads();
    if (t + 1 < t1) VW_WGM_LOAD(t + 1);
    vw_h8 A[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) A[q] = sg[((4 * wave + q) * 4 + g) * 16 + r];
    if (cib == 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 8; ++i) bsum += (float)A[q][i];
    }
    /* halo row 4 wave + h serves output rows 4 wave + q with ky = h - q */
#pragma unroll
    for (int h = 0; h < 6; ++h) {
      const int hr = 4 * wave + h;
      const vw_h8 B0 = sx[((0 * 18 + hr) * 4 + g) * 16 + r];
      const vw_h8 B1 = sx[((1 * 18 + hr) * 4 + g) * 16 + r];
      const vw_h8 B2 = sx[((2 * 18 + hr) * 4 + g) * 16 + r];
#pragma unroll
      for (int ky = 0; ky < 3; ++ky) {
        const int q = h - ky;
        if (q < 0 || q > 3) continue;
        acc[3 * ky + 0] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[q], B0, acc[3 * ky + 0], 0, 0, 0);
        acc[3 * ky + 1] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[q], B1, acc[3 * ky + 1], 0, 0, 0);
        acc[3 * ky + 2] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[q], B2, acc[3 * ky + 2], 0, 0, 0);
      }
    }
  }
#undef VW_WGM_LDX
#undef VW_WGM_LOAD
#undef VW_WGM_STX
  /* fold the four waves through the arena (waves 1..3 write, wave 0 adds) */
  __syncthreads();
  float* red = reinterpret_cast<float*>(arena);
  if (wave > 0) {
    float* d = red + ((wave - 1) * 64 + lane) * 37;
#pragma unroll
    for (int k = 0; k < 9; ++k)
#pragma unroll
      for (int i = 0; i < 4; ++i) d[4 * k + i] = acc[k][i];
    d[36] = bsum;
  }
  __syncthreads();
  if (wave == 0) {
    for (int w = 0; w < 3; ++w) {
      const float* d = red + (w * 64 + lane) * 37;
#pragma unroll
      for (int k = 0; k < 9; ++k)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[k][i] += d[4 * k + i];
      bsum += d[36];
    }
    /* D[4 g + i][r]: output channel cob + 4 g + i, input channel cib + r */
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float* dst = part + (((int64_t)bz * co_n + cob + 4 * g + i) * ci_n + cib + r) * 10;
      if (cib + r < ci_n) {
#pragma unroll
        for (int k = 0; k < 9; ++k) dst[k] = acc[k][i];
      }
    }
    /* the bias of output channel cob + r: the four chunk groups of the lanes */
    bsum += __shfl_xor(bsum, 16);
    bsum += __shfl_xor(bsum, 32);
    if (cib == 0 && g == 0) part[(((int64_t)bz * co_n + cob + r) * ci_n) * 10 + 9] = bsum;
  }
}

/* The weight-gradient folds of all 3x3 layers of a backward pass, two
 * launches after the last layer: stage 1 sums groups of VW_FOLD_GRP
 * consecutive chunk partials in chunk order, [group][per] per layer; stage 2
 * sums a layer's groups in order into dW [co][ci][9] and dB [co] (the ci = 0
 * partials).  The same order every call, so training is bit-reproducible. */
#define VW_FOLD_GRP 32
struct FoldJobs {
  int njob;
  int64_t base1[15], base2[15]; /* thread prefix sums: per x groups, per */
  const float* part[14];        /* [nch][per] */
  float* grp[14];               /* [ng][per]  */
  float *dW[14], *dB[14];
  int64_t per[14];
  int nch[14], ng[14], ci[14];
};
__device__ __forceinline__ int vw_fold_job(const int64_t* base, int njob, int64_t t) {
  int j = 0;
  while (j + 1 < njob && t >= base[j + 1]) ++j;
  return j;
}
__global__ void k_fold_stage1(FoldJobs J) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= J.base1[J.njob]) return;
  const int j = vw_fold_job(J.base1, J.njob, t);
  const int64_t per = J.per[j], u = t - J.base1[j];
  const int64_t gi = u / per, i = u - gi * per;
  const int c0 = (int)gi * VW_FOLD_GRP;
  const int c1 = c0 + VW_FOLD_GRP < J.nch[j] ? c0 + VW_FOLD_GRP : J.nch[j];
  const float* p = J.part[j];
  float s = 0.f;
  for (int c = c0; c < c1; ++c) s += p[c * per + i];
  J.grp[j][gi * per + i] = s;
}
__global__ void k_fold_stage2(FoldJobs J) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= J.base2[J.njob]) return;
  const int j = vw_fold_job(J.base2, J.njob, t);
  const int64_t per = J.per[j], i = t - J.base2[j];
  const float* g = J.grp[j];
  float s = 0.f;
  for (int k = 0; k < J.ng[j]; ++k) s += g[k * per + i];
  const int ci_n = J.ci[j];
  const int64_t pair = i / 10, k = i - pair * 10;
  const int64_t co = pair / ci_n, ci = pair - co * ci_n;
  if (k < 9) J.dW[j][(co * ci_n + ci) * 9 + k] = s;
  else if (ci == 0) J.dB[j][co] = s;
}

/* dz = dy * (y > 0) */
template <typename T>
__global__ void k_relu_bwd(const T* __restrict__ dy, const T* __restrict__ y, T* __restrict__ dz, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = (float)y[i] > 0.f ? dy[i] : (T)0.f;
}

/* 2x2 average pooling of [nc][H][H] to [nc][H/2][H/2] */
template <typename T>
__global__ void k_pool(const T* __restrict__ x, T* __restrict__ y, int64_t nc, int H) {
  const int h = H / 2;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * h * h) return;
  const int64_t c = i / ((int64_t)h * h);
  const int r = (int)(i - c * h * h), yy = r / h, xx = r - yy * h;
  const T* s = x + c * H * H + (int64_t)(2 * yy) * H + 2 * xx;
  y[i] = (T)(0.25f * (((float)s[0] + (float)s[1]) + ((float)s[H] + (float)s[H + 1])));
}
/* dx = base + pool^T(dy): each fine pixel gets a quarter of its coarse gradient */
template <typename T>
__global__ void k_pool_bwd_add(const T* __restrict__ dy, const T* __restrict__ base, T* __restrict__ dx, int64_t nc,
                               int H, const T* __restrict__ mask) {
  const int h = H / 2;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * H * H) return;
  const int64_t c = i / ((int64_t)H * H);
  const int r = (int)(i - c * H * H), y = r / H, x = r - y * H;
  const float g = 0.25f * (float)dy[c * h * h + (int64_t)(y / 2) * h + x / 2];
  const float v = (base ? (float)base[i] : 0.f) + g;
  dx[i] = (T)(mask && !((float)mask[i] > 0.f) ? 0.f : v);
}
/* nearest-neighbour upsampling of [nc][h][h] to [nc][2h][2h] */
template <typename T>
__global__ void k_up(const T* __restrict__ x, T* __restrict__ y, int64_t nc, int h) {
  const int H = 2 * h;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * H * H) return;
  const int64_t c = i / ((int64_t)H * H);
  const int r = (int)(i - c * H * H), yy = r / H, xx = r - yy * H;
  y[i] = x[c * h * h + (int64_t)(yy / 2) * h + xx / 2];
}
template <typename T>
__global__ void k_up_bwd(const T* __restrict__ dy, T* __restrict__ dx, int64_t nc, int h, const T* __restrict__ mask) {
  const int H = 2 * h;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * h * h) return;
  const int64_t c = i / ((int64_t)h * h);
  const int r = (int)(i - c * h * h), yy = r / h, xx = r - yy * h;
  const T* s = dy + c * H * H + (int64_t)(2 * yy) * H + 2 * xx;
  const float v = ((float)s[0] + (float)s[1]) + ((float)s[H] + (float)s[H + 1]);
  dx[i] = (T)(mask && !((float)mask[i] > 0.f) ? 0.f : v);
}
template <typename T>
__global__ void k_add(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ c, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = (T)((float)a[i] + (float)b[i]);
}

/* FP16 forms of the element-wise kernels above: a lane takes 8 consecutive
 * outputs along a row (16-byte accesses; the output row width a multiple of
 * 8), with the same arithmetic per element */
typedef _Float16 vw_h4 __attribute__((ext_vector_type(4)));
__global__ void k_relu_bwd_h8(const vw_h8* __restrict__ dy, const vw_h8* __restrict__ y, vw_h8* __restrict__ dz,
                              int64_t n8) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const vw_h8 a = dy[i], b = y[i];
  vw_h8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (float)b[k] > 0.f ? a[k] : (_Float16)0.f;
  dz[i] = r;
}
__global__ void k_add_h8(const vw_h8* __restrict__ a, const vw_h8* __restrict__ b, vw_h8* __restrict__ c, int64_t n8) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n8) return;
  const vw_h8 u = a[i], v = b[i];
  vw_h8 r;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = (_Float16)((float)u[k] + (float)v[k]);
  c[i] = r;
}
/* 2x2 pooling (up_bwd = 0: average) or the adjoint of upsampling (up_bwd = 1:
 * sum) of [nc][H][H] to [nc][H/2][H/2]; (H/2) % 8 == 0 */
template <bool SUM>
__global__ void k_pool_h8(const _Float16* __restrict__ x, _Float16* __restrict__ y, int64_t nc, int H,
                          const _Float16* __restrict__ mask) {
  const int h = H / 2, h8 = h / 8;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * h * h8) return;
  const int64_t c = i / ((int64_t)h * h8);
  const int r = (int)(i - c * h * h8), yy = r / h8, xx = 8 * (r - yy * h8);
  const _Float16* s = x + c * H * H + (int64_t)(2 * yy) * H + 2 * xx;
  const vw_h8 a0 = *(const vw_h8*)s, a1 = *(const vw_h8*)(s + 8);
  const vw_h8 b0 = *(const vw_h8*)(s + H), b1 = *(const vw_h8*)(s + H + 8);
  vw_h8 o;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float v0 = ((float)a0[2 * k] + (float)a0[2 * k + 1]) + ((float)b0[2 * k] + (float)b0[2 * k + 1]);
    const float v1 = ((float)a1[2 * k] + (float)a1[2 * k + 1]) + ((float)b1[2 * k] + (float)b1[2 * k + 1]);
    o[k] = (_Float16)(SUM ? v0 : 0.25f * v0);
    o[4 + k] = (_Float16)(SUM ? v1 : 0.25f * v1);
  }
  const int64_t oi = c * h * h + (int64_t)yy * h + xx;
  if (mask) {
    const vw_h8 m = *(const vw_h8*)(mask + oi);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (!((float)m[k] > 0.f)) o[k] = (_Float16)0.f;
  }
  *(vw_h8*)(y + oi) = o;
}
/* upsampling (base == nullptr, scale 1) or dx = base + pool^T(dy) (scale
 * 1/4) of [nc][h][h] to [nc][2h][2h]; (2h) % 8 == 0 */
template <bool POOLT>
__global__ void k_up_h8(const _Float16* __restrict__ x, const _Float16* __restrict__ base, _Float16* __restrict__ y,
                        int64_t nc, int h, const _Float16* __restrict__ mask) {
  const int H = 2 * h, H8 = H / 8;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc * H * H8) return;
  const int64_t c = i / ((int64_t)H * H8);
  const int r = (int)(i - c * H * H8), yy = r / H8, xx = 8 * (r - yy * H8);
  const vw_h4 s = *(const vw_h4*)(x + c * h * h + (int64_t)(yy / 2) * h + xx / 2);
  const int64_t o = c * H * H + (int64_t)yy * H + xx;
  vw_h8 v;
  if (POOLT) {
    const vw_h8 b = base ? *(const vw_h8*)(base + o) : vw_h8{};
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = (_Float16)((base ? (float)b[k] : 0.f) + 0.25f * (float)s[k / 2]);
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = s[k / 2];
  }
  if (mask) {
    const vw_h8 m = *(const vw_h8*)(mask + o);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (!((float)m[k] > 0.f)) v[k] = (_Float16)0.f;
  }
  *(vw_h8*)(y + o) = v;
}

/* 1x1 output convolution plus the input: out = b + sum_c w[c] d[c] + x */
template <typename T>
__global__ void k_out(const T* __restrict__ d, const T* __restrict__ x, const float* __restrict__ w,
                      const float* __restrict__ b, T* __restrict__ out, int N, int C, int64_t HH) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * HH) return;
  const int64_t n = i / HH, p = i - n * HH;
  float s = b[0];
  for (int c = 0; c < C; ++c) s = fmaf(w[c], (float)d[(n * C + c) * HH + p], s);
  out[i] = (T)(s + (float)x[i]);
}
/* masked squared error: per-block partial sums of the squared error and the
 * scaled output gradient dout = 2 (out - y) mask * gscale.  With gscale =
 * 1 / batch (not 1 / (batch * disk pixels)) the gradients stay O(1) per
 * pixel, clear of the FP16 subnormal range; every parameter gradient is then
 * the loss gradient times the disk's pixel count, which Adam's update does
 * not see and vw_unet_loss_grad divides out. */
template <typename T>
__global__ __launch_bounds__(256) void k_mse(const T* __restrict__ out, const T* __restrict__ y,
                                             const uint8_t* __restrict__ mask, int64_t HH, int64_t n, float gscale,
                                             T* __restrict__ dout, float* __restrict__ part) {
  __shared__ float red[256];
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.f;
  if (i < n) {
    const bool m = mask[i % HH] != 0;
    const float e = m ? (float)out[i] - (float)y[i] : 0.f;
    e2 = e * e;
    dout[i] = (T)(2.f * e * gscale);
  }
  const float s = vw_block_sum<T>(e2, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
/* ordered sum of the partials -> loss[slot] (FP64) */
__global__ __launch_bounds__(256) void k_loss_sum(const float* __restrict__ part, int64_t np, float inv_count,
                                                  double* __restrict__ loss) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += 256) s += (double)part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * (double)inv_count;
}
/* output layer backward: dd[c] = w[c] dout, and dw[c] / db by one workgroup per c */
template <typename T>
__global__ void k_out_bwd_data(const T* __restrict__ dout, const float* __restrict__ w, T* __restrict__ dd, int N,
                               int C, int64_t HH, const T* __restrict__ mask) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * C * HH) return;
  const int64_t n = i / (C * HH), r = i - n * C * HH, c = r / HH, p = r - c * HH;
  const float v = w[c] * (float)dout[n * HH + p];
  dd[i] = (T)(mask && !((float)mask[i] > 0.f) ? 0.f : v);
}
/* one workgroup per (channel, chunk of pixels); partials [chunk][C][2] */
template <typename T>
__global__ __launch_bounds__(256) void k_out_wgrad(const T* __restrict__ dout, const T* __restrict__ d,
                                                   float* __restrict__ part, int N, int C, int64_t HH) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  const int64_t p0 = (int64_t)blockIdx.y * VW_WG_CHUNK;
  const int64_t p1 = p0 + VW_WG_CHUNK < N * HH ? p0 + VW_WG_CHUNK : N * HH;
  float s = 0.f, sb = 0.f;
  for (int64_t i = p0 + threadIdx.x; i < p1; i += 256) {
    const int64_t n = i / HH, p = i - n * HH;
    const float g = (float)dout[i];
    s = fmaf(g, (float)d[(n * C + c) * HH + p], s);
    sb += g;
  }
  const float ts = vw_block_sum<T>(s, red);
  const float tb = vw_block_sum<T>(sb, red);
  if (threadIdx.x == 0) {
    part[((int64_t)blockIdx.y * C + c) * 2] = ts;
    part[((int64_t)blockIdx.y * C + c) * 2 + 1] = tb;
  }
}
__global__ void k_out_wgrad_fold(const float* __restrict__ part, int nchunk, int C, float* __restrict__ dw,
                                 float* __restrict__ db) {
  const int c = threadIdx.x;
  if (c >= C) return;
  float s = 0.f, sb = 0.f;
  for (int k = 0; k < nchunk; ++k) {
    s += part[((int64_t)k * C + c) * 2];
    sb += part[((int64_t)k * C + c) * 2 + 1];
  }
  dw[c] = s;
  if (c == 0) db[0] = sb;
}
/* batch gather from the resident set */
template <typename T>
__global__ void k_gather(const T* __restrict__ X, const T* __restrict__ Y, const int32_t* __restrict__ idx, int n,
                         int64_t HH, T* __restrict__ xb, T* __restrict__ yb) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * HH) return;
  const int64_t b = i / HH, p = i - b * HH;
  const int64_t src = (int64_t)idx[b] * HH + p;
  xb[i] = X[src];
  yb[i] = Y[src];
}
template <typename T>
__global__ void k_from_f32(const float* __restrict__ a, T* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}
template <typename T>
__global__ void k_to_f32(const T* __restrict__ a, float* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
/* Adam with bias correction; step t >= 1 */
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
}

/* ------------------------------------------------------------------ */
/* network                                                              */
/* ------------------------------------------------------------------ */

struct Conv {
  int ci, co, lvl;   /* channels and level (image side img_n >> lvl) */
  int64_t w, b;      /* offsets in the parameter vector */
  int64_t hf, hb;    /* offsets of the FP16 forward / data-gradient weight layouts */
};

struct Net {
  int N = 0, BM = 0;
  bool fp16 = false;
  std::string err;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  std::vector<Conv> conv; /* 14 3x3 layers in forward order, then the 1x1 output */
  int64_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  _Float16* wh = nullptr; /* FP16 weight layouts of the 3x3 layers (matrix-core path) */
  int64_t nwh = 0;
  int64_t adam_t = 0;
  uint8_t* mask = nullptr;
  int64_t nmask = 0;
  /* activations and gradients, each [BM][C][H][H] of the storage type */
  enum {
    X0, E0A, E0, P1, E1A, E1, P2, E2A, E2, P3, E3A, E3, U2, D2A, S2, D2, U1, D1A, S1, D1, U0, D0A, S0, D0, OUT, YT, NT
  };
  int ch[NT] = {}, lv[NT] = {};
  void* a[NT] = {};
  void* d[NT] = {};
  void* dz = nullptr; /* masked gradient scratch, the largest tensor */
  float* part = nullptr;
  float* wpart = nullptr;  /* weight-gradient chunk partials, a region per 3x3 layer, then the 1x1 layer's */
  float* wpart2 = nullptr; /* their sums over groups of VW_FOLD_GRP chunks, a region per 3x3 layer */
  int64_t woff[15] = {}, w2off[14] = {};
  FoldJobs jobs{};         /* the folds queued by the current backward pass */
  double* loss = nullptr; /* per-step losses on the device */
  int64_t loss_cap = 0;
  void *X = nullptr, *Y = nullptr;
  int64_t count = 0;
  int32_t* idx = nullptr;
  int64_t idx_cap = 0;
  double train_ms = 0.0;
  float* scratch_dev = nullptr; /* this device's address of scratch (vw_unet_predict_noalloc) */

  int fail(int code, const std::string& e) {
    err = e;
    return code;
  }
  int H(int l) const { return N >> l; }
  int64_t elems(int t) const { return (int64_t)ch[t] * H(lv[t]) * H(lv[t]); }
};

#define UC(x)                                                                              \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

/* weight-gradient chunking of a 3x3 layer for a batch of n: whether the
 * matrix-core kernel runs, its tiles per chunk, and the chunk count (about
 * 512 workgroups per launch, all resident at once at two per CU: fewer chunks also mean
 * fewer partials to write and fold; measured against 2,048, 1,024 and 256 at
 * batch 16, 256 x 256) */
struct WgPlan {
  bool mfma;
  int tpc;
  int64_t nch;
};
inline WgPlan wg_plan(bool fp16, const Conv& L, int H, int n) {
  WgPlan p;
  p.mfma = fp16 && L.co % 16 == 0 && H % 8 == 0;
  if (p.mfma) {
    const int64_t tiles = (int64_t)n * ((H + 15) / 16) * ((H + 31) / 32);
    const int64_t nb = (int64_t)((L.ci + 15) / 16) * (L.co / 16);
    p.tpc = (int)std::max<int64_t>(1, tiles * nb / 512);
    p.nch = (tiles + p.tpc - 1) / p.tpc;
  } else {
    const int64_t tpr = (H + 15) / 16;
    p.tpc = VW_WG_TILES;
    p.nch = ((int64_t)n * tpr * tpr + VW_WG_TILES - 1) / VW_WG_TILES;
  }
  return p;
}

template <typename P>
int ualloc(Net* c, P** p, size_t bytes) {
  void* q = nullptr;
  UC(hipMalloc(&q, bytes ? bytes : 1));
  c->allocs.push_back(q);
  *p = (P*)q;
  return VW_OK;
}

/* matrix-core convolution launch: 32-channel LDS chunks when cin allows;
 * weight prefetch for launches of fewer than 4,096 workgroups; vector
 * staging (choices measured at batch 16, 256 x 256: with the XCD-grouped
 * workgroup order the vector staging is faster on every layer) */
#ifndef VW_CONV_MIN_WG
#define VW_CONV_MIN_WG 512
#endif
template <bool RELU, int CB>
void conv3_mfma_cb(Net* c, dim3 grid, const _Float16* in, const _Float16* Wk, const float* B, _Float16* out, int cin,
                   int cout, int H, const _Float16* mask) {
  /* grid = (tiles, 16-channel output blocks, images); a workgroup takes
   * OCB of the blocks (each wave all of them on its 4 pixel rows: the input
   * operand read from LDS serves OCB products) while that leaves at least
   * VW_CONV_MIN_WG workgroups */
  const unsigned nb16 = grid.y, rest = grid.x * grid.z;
  int ocb = 1;
  if (nb16 % 4 == 0 && rest * (nb16 / 4) >= VW_CONV_MIN_WG) ocb = 4;
  else if (nb16 % 2 == 0 && rest * (nb16 / 2) >= VW_CONV_MIN_WG) ocb = 2;
  const unsigned nwg = rest * (nb16 / ocb);
  const bool pf = ocb == 1 && nwg < 4096;
  if (pf) k_conv3_mfma<RELU, CB, true, true, 1><<<nwg, 256, 0, c->st>>>(in, Wk, B, out, cin, cout, H, mask);
  else if (ocb == 1) k_conv3_mfma<RELU, CB, false, true, 1><<<nwg, 256, 0, c->st>>>(in, Wk, B, out, cin, cout, H, mask);
  else if (ocb == 2) k_conv3_mfma<RELU, CB, false, true, 2><<<nwg, 256, 0, c->st>>>(in, Wk, B, out, cin, cout, H, mask);
  else k_conv3_mfma<RELU, CB, false, true, 4><<<nwg, 256, 0, c->st>>>(in, Wk, B, out, cin, cout, H, mask);
}
template <bool RELU>
void conv3_mfma(Net* c, dim3 grid, const _Float16* in, const _Float16* Wk, const float* B, _Float16* out, int cin,
                int cout, int H, const _Float16* mask = nullptr) {
  if (cin % 32 == 0) conv3_mfma_cb<RELU, 32>(c, grid, in, Wk, B, out, cin, cout, H, mask);
  else conv3_mfma_cb<RELU, 16>(c, grid, in, Wk, B, out, cin, cout, H, mask);
}

template <typename T>
struct Ops {
  static constexpr bool kHalf = std::is_same<T, _Float16>::value;
  static int conv3(Net* c, const Conv& L, int src, int dst, int n) {
    const int H = c->H(L.lvl);
    const int tpr = (H + 15) / 16;
    dim3 grid((unsigned)(tpr * tpr), (unsigned)((L.co + 15) / 16), (unsigned)n);
    if (kHalf && L.ci % 16 == 0) { /* matrix cores */
      conv3_mfma<true>(c, grid, (const _Float16*)c->a[src], c->wh + L.hf, c->w + L.b, (_Float16*)c->a[dst], L.ci,
                       L.co, H);
      UC(hipGetLastError());
      return VW_OK;
    }
    k_conv3<T, false, true><<<grid, 256, 0, c->st>>>((const T*)c->a[src], c->w + L.w, c->w + L.b, (T*)c->a[dst], L.ci,
                                                     L.co, H, nullptr);
    UC(hipGetLastError());
    return VW_OK;
  }
  /* backward of layer L whose input is tensor src and output (a ReLU output)
   * tensor dst.  The masked output gradient dz = d[dst] (y > 0) is d[dst]
   * itself, masked by the kernel that produced it, when dst's gradient has
   * no other use; with gdst >= 0 it is formed here from d[gdst], the equal
   * gradient of a skip sum that is used elsewhere too.  Writes the weight
   * gradients and, if dsrc >= 0, d[dsrc] (masked by a[dsrc] when mask_dsrc:
   * dsrc is the ReLU output of the layer below and this its only use). */
  static int conv3_bwd(Net* c, const Conv& L, int src, int dst, int dsrc, int n, int gdst = -1,
                       bool mask_dsrc = false) {
    const int H = c->H(L.lvl);
    const int64_t ne = (int64_t)n * c->elems(dst);
    const void* dzp = c->d[dst];
    if (gdst >= 0) {
      const void* gd = c->d[gdst];
      if (kHalf && ne % 8 == 0)
        k_relu_bwd_h8<<<nblk(ne / 8), 256, 0, c->st>>>((const vw_h8*)gd, (const vw_h8*)c->a[dst], (vw_h8*)c->dz,
                                                        ne / 8);
      else
        k_relu_bwd<T><<<nblk(ne), 256, 0, c->st>>>((const T*)gd, (const T*)c->a[dst], (T*)c->dz, ne);
      UC(hipGetLastError());
      dzp = c->dz;
    }
    const T* dmask = mask_dsrc ? (const T*)c->a[dsrc] : nullptr;
    const WgPlan wp = wg_plan(kHalf, L, H, n);
    const int li = (int)(&L - c->conv.data()); /* layer index: its partial regions */
    float* part = c->wpart + c->woff[li];
    const dim3 wgrid((unsigned)((L.ci + 15) / 16), (unsigned)((L.co + 15) / 16), (unsigned)wp.nch);
    if (wp.mfma)
      k_conv3_wgrad_mfma<<<wgrid.x * wgrid.y * wgrid.z, 256, 0, c->st>>>(
          (const _Float16*)c->a[src], (const _Float16*)dzp, part, n, L.ci, L.co, H, wp.tpc);
    else
      k_conv3_wgrad_t<T><<<wgrid, 256, 0, c->st>>>((const T*)c->a[src], (const T*)dzp, part, n, L.ci, L.co, H);
    UC(hipGetLastError());
    { /* queue the fold (launched once per backward pass) */
      FoldJobs& J = c->jobs;
      if (J.njob >= 14) return c->fail(VW_ERR_RANGE, "internal: more than 14 weight-gradient folds in one pass");
      const int j = J.njob++;
      const int64_t per = (int64_t)L.co * L.ci * 10;
      J.part[j] = part;
      J.grp[j] = c->wpart2 + c->w2off[li];
      J.dW[j] = c->g + L.w;
      J.dB[j] = c->g + L.b;
      J.per[j] = per;
      J.nch[j] = (int)wp.nch;
      J.ng[j] = (int)((wp.nch + VW_FOLD_GRP - 1) / VW_FOLD_GRP);
      J.ci[j] = L.ci;
      J.base1[j + 1] = J.base1[j] + per * J.ng[j];
      J.base2[j + 1] = J.base2[j] + per;
    }
    if (dsrc >= 0) {
      const int tpr = (H + 15) / 16;
      dim3 grid((unsigned)(tpr * tpr), (unsigned)((L.ci + 15) / 16), (unsigned)n);
      if (kHalf && L.co % 16 == 0) { /* matrix cores, flipped transpose */
        conv3_mfma<false>(c, grid, (const _Float16*)dzp, c->wh + L.hb, nullptr, (_Float16*)c->d[dsrc], L.co, L.ci,
                          H, (const _Float16*)dmask);
        UC(hipGetLastError());
        return VW_OK;
      }
      k_conv3<T, true, false><<<grid, 256, 0, c->st>>>((const T*)dzp, c->w + L.w, nullptr, (T*)c->d[dsrc], L.co,
                                                       L.ci, H, dmask);
      UC(hipGetLastError());
    }
    return VW_OK;
  }
  static int pool(Net* c, int s, int t, int n) {
    const int64_t nc = (int64_t)n * c->ch[s];
    const int H = c->H(c->lv[s]);
    if (kHalf && (H / 2) % 8 == 0)
      k_pool_h8<false><<<nblk(nc * (H / 2) * (H / 16)), 256, 0, c->st>>>((const _Float16*)c->a[s],
                                                                         (_Float16*)c->a[t], nc, H, nullptr);
    else
      k_pool<T><<<nblk(nc * (H / 2) * (H / 2)), 256, 0, c->st>>>((const T*)c->a[s], (T*)c->a[t], nc, H);
    UC(hipGetLastError());
    return VW_OK;
  }
  static int up(Net* c, int s, int t, int n) {
    const int64_t nc = (int64_t)n * c->ch[s];
    const int h = c->H(c->lv[s]);
    if (kHalf && (2 * h) % 8 == 0)
      k_up_h8<false><<<nblk(nc * 2 * h * (2 * h / 8)), 256, 0, c->st>>>((const _Float16*)c->a[s], nullptr,
                                                                        (_Float16*)c->a[t], nc, h, nullptr);
    else
      k_up<T><<<nblk(nc * 4 * h * h), 256, 0, c->st>>>((const T*)c->a[s], (T*)c->a[t], nc, h);
    UC(hipGetLastError());
    return VW_OK;
  }
  static int add(Net* c, int x, int y, int z, int n) {
    const int64_t ne = (int64_t)n * c->elems(z);
    if (kHalf && ne % 8 == 0)
      k_add_h8<<<nblk(ne / 8), 256, 0, c->st>>>((const vw_h8*)c->a[x], (const vw_h8*)c->a[y], (vw_h8*)c->a[z],
                                                 ne / 8);
    else
      k_add<T><<<nblk(ne), 256, 0, c->st>>>((const T*)c->a[x], (const T*)c->a[y], (T*)c->a[z], ne);
    UC(hipGetLastError());
    return VW_OK;
  }

  static int forward(Net* c, int n) {
    const std::vector<Conv>& L = c->conv;
    int rc;
    if (kHalf) { /* FP16 layouts of the current weights */
      WLayoutJobs J{};
      for (int k = 0; k < 14; ++k) {
        J.w[k] = L[k].w;
        J.hf[k] = L[k].hf;
        J.hb[k] = L[k].hb;
        J.ci[k] = L[k].ci;
        J.co[k] = L[k].co;
        J.base[k + 1] = J.base[k] + (int64_t)L[k].co * L[k].ci * 9;
      }
      k_wlayout<<<nblk(J.base[14]), 256, 0, c->st>>>(J, c->w, c->wh);
      UC(hipGetLastError());
    }
#define F(x) if ((rc = (x))) return rc
    F(conv3(c, L[0], Net::X0, Net::E0A, n));
    F(conv3(c, L[1], Net::E0A, Net::E0, n));
    F(pool(c, Net::E0, Net::P1, n));
    F(conv3(c, L[2], Net::P1, Net::E1A, n));
    F(conv3(c, L[3], Net::E1A, Net::E1, n));
    F(pool(c, Net::E1, Net::P2, n));
    F(conv3(c, L[4], Net::P2, Net::E2A, n));
    F(conv3(c, L[5], Net::E2A, Net::E2, n));
    F(pool(c, Net::E2, Net::P3, n));
    F(conv3(c, L[6], Net::P3, Net::E3A, n));
    F(conv3(c, L[7], Net::E3A, Net::E3, n));
    F(up(c, Net::E3, Net::U2, n));
    F(conv3(c, L[8], Net::U2, Net::D2A, n));
    F(add(c, Net::D2A, Net::E2, Net::S2, n));
    F(conv3(c, L[9], Net::S2, Net::D2, n));
    F(up(c, Net::D2, Net::U1, n));
    F(conv3(c, L[10], Net::U1, Net::D1A, n));
    F(add(c, Net::D1A, Net::E1, Net::S1, n));
    F(conv3(c, L[11], Net::S1, Net::D1, n));
    F(up(c, Net::D1, Net::U0, n));
    F(conv3(c, L[12], Net::U0, Net::D0A, n));
    F(add(c, Net::D0A, Net::E0, Net::S0, n));
    F(conv3(c, L[13], Net::S0, Net::D0, n));
    const Conv& O = L[14];
    const int64_t HH = (int64_t)c->N * c->N;
    k_out<T><<<nblk(n * HH), 256, 0, c->st>>>((const T*)c->a[Net::D0], (const T*)c->a[Net::X0], c->w + O.w,
                                              c->w + O.b, (T*)c->a[Net::OUT], n, 16, HH);
    UC(hipGetLastError());
    return VW_OK;
  }

  /* loss into loss[slot], gradients into c->g */
  /* the skip additions S = DA + E pass their gradient unchanged to DA: the
   * layer producing DA reads dS directly (no copy) */
  static int backward(Net* c, int n, int64_t slot) {
    const std::vector<Conv>& L = c->conv;
    const int64_t HH = (int64_t)c->N * c->N;
    const int64_t ne = n * HH;
    const unsigned nb = nblk(ne);
    const float inv = 1.0f / (float)(c->nmask * n);
    k_mse<T><<<nb, 256, 0, c->st>>>((const T*)c->a[Net::OUT], (const T*)c->a[Net::YT], c->mask, HH, ne,
                                    1.0f / (float)n, (T*)c->d[Net::OUT], c->part);
    UC(hipGetLastError());
    k_loss_sum<<<1, 256, 0, c->st>>>(c->part, nb, inv, c->loss + slot);
    UC(hipGetLastError());
    const Conv& O = L[14];
    const int onch = (int)((ne + VW_WG_CHUNK - 1) / VW_WG_CHUNK);
    float* opart = c->wpart + c->woff[14];
    k_out_wgrad<T><<<dim3(16, (unsigned)onch), 256, 0, c->st>>>((const T*)c->d[Net::OUT], (const T*)c->a[Net::D0],
                                                                opart, n, 16, HH);
    UC(hipGetLastError());
    k_out_wgrad_fold<<<1, 64, 0, c->st>>>(opart, onch, 16, c->g + O.w, c->g + O.b);
    UC(hipGetLastError());
    k_out_bwd_data<T><<<nblk(ne * 16), 256, 0, c->st>>>((const T*)c->d[Net::OUT], c->w + O.w, (T*)c->d[Net::D0], n,
                                                        16, HH, (const T*)c->a[Net::D0]); /* masked: dD0 feeds only L13 */
    UC(hipGetLastError());
    int rc;
    c->jobs.njob = 0;
    c->jobs.base1[0] = c->jobs.base2[0] = 0;
    /* decoder level 0 */
    F(conv3_bwd(c, L[13], Net::S0, Net::D0, Net::S0, n));            /* dS0 */
    F(conv3_bwd(c, L[12], Net::U0, Net::D0A, Net::U0, n, Net::S0));  /* dU0 (dD0A = dS0) */
    F(up_bwd(c, Net::U0, Net::D1, n));                               /* dD1 */
    F(conv3_bwd(c, L[11], Net::S1, Net::D1, Net::S1, n));
    F(conv3_bwd(c, L[10], Net::U1, Net::D1A, Net::U1, n, Net::S1));
    F(up_bwd(c, Net::U1, Net::D2, n));
    F(conv3_bwd(c, L[9], Net::S2, Net::D2, Net::S2, n));
    F(conv3_bwd(c, L[8], Net::U2, Net::D2A, Net::U2, n, Net::S2));
    F(up_bwd(c, Net::U2, Net::E3, n));                               /* dE3 */
    /* encoder, deepest first; the skip gradients dS* seed dE0..dE2 */
    F(conv3_bwd(c, L[7], Net::E3A, Net::E3, Net::E3A, n, -1, true));
    F(conv3_bwd(c, L[6], Net::P3, Net::E3A, Net::P3, n));
    F(pool_bwd(c, Net::P3, Net::S2, Net::E2, n));                    /* dE2 = dS2 + pool^T dP3 */
    F(conv3_bwd(c, L[5], Net::E2A, Net::E2, Net::E2A, n, -1, true));
    F(conv3_bwd(c, L[4], Net::P2, Net::E2A, Net::P2, n));
    F(pool_bwd(c, Net::P2, Net::S1, Net::E1, n));
    F(conv3_bwd(c, L[3], Net::E1A, Net::E1, Net::E1A, n, -1, true));
    F(conv3_bwd(c, L[2], Net::P1, Net::E1A, Net::P1, n));
    F(pool_bwd(c, Net::P1, Net::S0, Net::E0, n));
    F(conv3_bwd(c, L[1], Net::E0A, Net::E0, Net::E0A, n, -1, true));
    F(conv3_bwd(c, L[0], Net::X0, Net::E0A, -1, n));
#undef F
    const FoldJobs& J = c->jobs;
    k_fold_stage1<<<nblk(J.base1[J.njob]), 256, 0, c->st>>>(J);
    UC(hipGetLastError());
    k_fold_stage2<<<nblk(J.base2[J.njob]), 256, 0, c->st>>>(J);
    UC(hipGetLastError());
    return VW_OK;
  }
  static int up_bwd(Net* c, int s_up, int t, int n) {
    const int64_t nc = (int64_t)n * c->ch[t];
    const int h = c->H(c->lv[t]);
    if (kHalf && h % 8 == 0)
      k_pool_h8<true><<<nblk(nc * h * (h / 8)), 256, 0, c->st>>>((const _Float16*)c->d[s_up], (_Float16*)c->d[t],
                                                                 nc, 2 * h, (const _Float16*)c->a[t]);
    else
      k_up_bwd<T><<<nblk(nc * h * h), 256, 0, c->st>>>((const T*)c->d[s_up], (T*)c->d[t], nc, h, (const T*)c->a[t]);
    UC(hipGetLastError());
    return VW_OK;
  }
  /* d[t] = d[base] + pool^T d[s] (t, base at the fine level) */
  static int pool_bwd(Net* c, int s, int base, int t, int n) {
    const int64_t nc = (int64_t)n * c->ch[t];
    const int H = c->H(c->lv[t]);
    if (kHalf && H % 8 == 0)
      k_up_h8<true><<<nblk(nc * H * (H / 8)), 256, 0, c->st>>>((const _Float16*)c->d[s], (const _Float16*)c->d[base],
                                                               (_Float16*)c->d[t], nc, H / 2,
                                                               (const _Float16*)c->a[t]);
    else
      k_pool_bwd_add<T><<<nblk(nc * H * H), 256, 0, c->st>>>((const T*)c->d[s], (const T*)c->d[base], (T*)c->d[t],
                                                             nc, H, (const T*)c->a[t]);
    UC(hipGetLastError());
    return VW_OK;
  }
  /* host FP32 -> tensor t of the storage type, through a staging buffer:
   * allocated per call, or the caller's (staging, no allocation) */
  static int upload(Net* c, const float* host, int t, int n, bool grad = false, float* staging = nullptr) {
    const int64_t ne = n * c->elems(t);
    float* tmp = staging;
    if (!tmp) UC(hipMalloc(&tmp, sizeof(float) * ne));
    hipError_t e = hipMemcpyAsync(tmp, host, sizeof(float) * ne, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) {
      k_from_f32<T><<<nblk(ne), 256, 0, c->st>>>(tmp, (T*)(grad ? c->d[t] : c->a[t]), ne);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (!staging) (void)hipFree(tmp);
    if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("upload: ") + hipGetErrorString(e));
    return VW_OK;
  }
  static int download(Net* c, int t, int n, float* host, float* staging = nullptr) {
    const int64_t ne = n * c->elems(t);
    float* tmp = staging;
    if (!tmp) UC(hipMalloc(&tmp, sizeof(float) * ne));
    k_to_f32<T><<<nblk(ne), 256, 0, c->st>>>((const T*)c->a[t], tmp, ne);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, sizeof(float) * ne, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (!staging) (void)hipFree(tmp);
    if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("download: ") + hipGetErrorString(e));
    return VW_OK;
  }
  static int gather(Net* c, const int32_t* idx, int n) {
    const int64_t HH = (int64_t)c->N * c->N;
    k_gather<T><<<nblk(n * HH), 256, 0, c->st>>>((const T*)c->X, (const T*)c->Y, idx, n, HH, (T*)c->a[Net::X0],
                                                 (T*)c->a[Net::YT]);
    UC(hipGetLastError());
    return VW_OK;
  }
};

int build(Net* c, int img_n, int bm, bool fp16, uint64_t seed) {
  c->N = img_n;
  c->BM = bm;
  c->fp16 = fp16;
  /* (ci, co, level) of the 14 3x3 layers in forward order, then the 1x1 */
  const int spec[15][3] = {{1, 16, 0},   {16, 16, 0},  {16, 32, 1},  {32, 32, 1},  {32, 64, 2},
                           {64, 64, 2},  {64, 128, 3}, {128, 128, 3}, {128, 64, 2}, {64, 64, 2},
                           {64, 32, 1},  {32, 32, 1},  {32, 16, 0},  {16, 16, 0},  {16, 1, 0}};
  int64_t off = 0;
  for (int k = 0; k < 15; ++k) {
    Conv L;
    L.ci = spec[k][0];
    L.co = spec[k][1];
    L.lvl = spec[k][2];
    const int ks = k < 14 ? 9 : 1;
    L.w = off;
    off += (int64_t)L.co * L.ci * ks;
    L.b = off;
    off += L.co;
    L.hf = c->nwh;
    L.hb = c->nwh + (int64_t)L.co * L.ci * 9;
    if (k < 14) c->nwh += 2 * (int64_t)L.co * L.ci * 9;
    c->conv.push_back(L);
  }
  c->nparam = off;
  /* tensor channels and levels */
  const int tc[Net::NT][2] = {{1, 0},  {16, 0},  {16, 0},  {16, 1}, {32, 1}, {32, 1}, {32, 2}, {64, 2}, {64, 2},
                              {64, 3}, {128, 3}, {128, 3}, {128, 2}, {64, 2}, {64, 2}, {64, 2}, {64, 1}, {32, 1},
                              {32, 1}, {32, 1},  {32, 0},  {16, 0}, {16, 0}, {16, 0}, {1, 0},  {1, 0}};
  for (int t = 0; t < Net::NT; ++t) {
    c->ch[t] = tc[t][0];
    c->lv[t] = tc[t][1];
  }
  hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("stream: ") + hipGetErrorString(e));
  const size_t ts = fp16 ? 2 : 4;
  int rc;
  int64_t maxe = 0;
  for (int t = 0; t < Net::NT; ++t) {
    const size_t bytes = (size_t)bm * c->elems(t) * ts;
    maxe = std::max<int64_t>(maxe, (int64_t)bm * c->elems(t));
    if ((rc = ualloc(c, &c->a[t], bytes)) || (rc = ualloc(c, &c->d[t], bytes))) return rc;
  }
  if ((rc = ualloc(c, &c->dz, (size_t)maxe * ts))) return rc;
  /* weight-gradient partials: the largest chunks x pairs x 10 of any layer
   * and batch size (the chunk count is not monotone in the batch) */
  {
    int64_t wp = 0, wp2 = 0;
    for (int k = 0; k < 14; ++k) {
      const Conv& L = c->conv[k];
      int64_t r = 0, r2 = 0;
      for (int n = 1; n <= bm; ++n) {
        const int64_t nch = wg_plan(fp16, L, img_n >> L.lvl, n).nch;
        const int64_t per = (int64_t)L.ci * L.co * 10;
        r = std::max<int64_t>(r, nch * per);
        r2 = std::max<int64_t>(r2, (nch + VW_FOLD_GRP - 1) / VW_FOLD_GRP * per);
      }
      c->woff[k] = wp;
      c->w2off[k] = wp2;
      wp += (r + 63) & ~(int64_t)63;
      wp2 += (r2 + 63) & ~(int64_t)63;
    }
    c->woff[14] = wp; /* the output layer: pixel chunks x 16 channels x 2 */
    wp += ((int64_t)bm * img_n * img_n + VW_WG_CHUNK - 1) / VW_WG_CHUNK * 32;
    if ((rc = ualloc(c, &c->wpart, sizeof(float) * wp)) || (rc = ualloc(c, &c->wpart2, sizeof(float) * wp2)))
      return rc;
  }
  const int64_t HH = (int64_t)img_n * img_n;
  if ((rc = ualloc(c, &c->wh, sizeof(_Float16) * c->nwh))) return rc;
  if ((rc = ualloc(c, &c->w, sizeof(float) * c->nparam)) || (rc = ualloc(c, &c->g, sizeof(float) * c->nparam)) ||
      (rc = ualloc(c, &c->m, sizeof(float) * c->nparam)) || (rc = ualloc(c, &c->v, sizeof(float) * c->nparam)) ||
      (rc = ualloc(c, &c->mask, (size_t)HH)) || (rc = ualloc(c, &c->part, sizeof(float) * (bm * HH / 256 + 1))))
    return rc;
  /* lm-grid disk: pixel centres (2j + 1 - N) / N, inside when l^2 + m^2 < 1 */
  std::vector<uint8_t> mk((size_t)HH);
  c->nmask = 0;
  for (int i = 0; i < img_n; ++i)
    for (int j = 0; j < img_n; ++j) {
      const int64_t a = 2 * j + 1 - img_n, b = 2 * i + 1 - img_n;
      mk[(size_t)i * img_n + j] = (a * a + b * b < (int64_t)img_n * img_n) ? 1 : 0;
      c->nmask += mk[(size_t)i * img_n + j];
    }
  UC(hipMemcpy(c->mask, mk.data(), (size_t)HH, hipMemcpyHostToDevice));
  /* He-normal weights, zero biases; the 1x1 output layer starts at zero so
   * the untrained network returns its input */
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> w((size_t)c->nparam, 0.f);
  for (int k = 0; k < 14; ++k) {
    const Conv& L = c->conv[k];
    const double sd = std::sqrt(2.0 / (9.0 * L.ci));
    for (int64_t i = 0; i < (int64_t)L.co * L.ci * 9; ++i) w[(size_t)(L.w + i)] = (float)(sd * nd(rng));
  }
  UC(hipMemcpy(c->w, w.data(), sizeof(float) * c->nparam, hipMemcpyHostToDevice));
  UC(hipMemset(c->m, 0, sizeof(float) * c->nparam));
  UC(hipMemset(c->v, 0, sizeof(float) * c->nparam));
  return VW_OK;
}

Net* as_net(void* p) { return (Net*)p; }

/* Address of scratch on the current device, resolved once per device (with
 * that device current) and kept in a table by device ordinal. */
std::mutex g_scratch_mu;
std::vector<float*> g_scratch_addr;

int resolve_scratch(Net* c) {
  int dev = 0;
  UC(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_addr[(size_t)dev]) {
    void* a = nullptr;
    UC(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    if (!a) return c->fail(VW_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_addr[(size_t)dev] = (float*)a;
  }
  c->scratch_dev = g_scratch_addr[(size_t)dev];
  return VW_OK;
}

template <typename F>
int dispatch(Net* c, F&& f) {
  return c->fp16 ? f(Ops<_Float16>{}) : f(Ops<float>{});
}

}  // namespace

VW_API int vw_unet_create(int32_t img_n, int32_t batch_max, int32_t act_fp16, uint64_t seed, void** out) {
  if (!out) return VW_ERR_ARG;
  *out = nullptr;
  if (img_n < 8 || img_n % 8 != 0 || img_n > 4096 || batch_max < 1 || batch_max > 4096) {
    g_unet_err = "img_n must be a multiple of 8 in [8, 4096] and batch_max in [1, 4096]";
    return VW_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    g_unet_err = "no HIP device";
    return VW_ERR_DEVICE;
  }
  Net* c = new Net();
  int rc = build(c, img_n, batch_max, act_fp16 != 0, seed);
  if (!rc) rc = resolve_scratch(c);
  if (rc) {
    g_unet_err = c->err;
    vw_unet_destroy(c);
    return rc;
  }
  *out = c;
  return VW_OK;
}

VW_API void vw_unet_destroy(void* p) {
  Net* c = as_net(p);
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void* q : c->allocs) (void)hipFree(q);
  if (c->X) (void)hipFree(c->X);
  if (c->Y) (void)hipFree(c->Y);
  if (c->idx) (void)hipFree(c->idx);
  if (c->loss) (void)hipFree(c->loss);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

VW_API const char* vw_unet_last_error(void* p) { return p ? as_net(p)->err.c_str() : g_unet_err.c_str(); }

VW_API int64_t vw_unet_param_count(void* p) { return p ? as_net(p)->nparam : -1; }

VW_API int vw_unet_get_params(void* p, float* out) {
  Net* c = as_net(p);
  if (!c || !out) return VW_ERR_ARG;
  UC(hipMemcpy(out, c->w, sizeof(float) * c->nparam, hipMemcpyDeviceToHost));
  return VW_OK;
}

VW_API int vw_unet_set_params(void* p, const float* in) {
  Net* c = as_net(p);
  if (!c || !in) return VW_ERR_ARG;
  UC(hipMemcpy(c->w, in, sizeof(float) * c->nparam, hipMemcpyHostToDevice));
  return VW_OK;
}

namespace {
int ensure_loss(Net* c, int64_t n) {
  if (n <= c->loss_cap) return VW_OK;
  if (c->loss) (void)hipFree(c->loss);
  c->loss = nullptr;
  c->loss_cap = 0;
  UC(hipMalloc(&c->loss, sizeof(double) * n));
  c->loss_cap = n;
  return VW_OK;
}
}  // namespace

VW_API int vw_unet_loss_grad(void* p, const float* X, const float* Y, int32_t n, double* loss, float* grad) {
  Net* c = as_net(p);
  if (!c || !X || !Y) return VW_ERR_ARG;
  if (n < 1 || n > c->BM) return c->fail(VW_ERR_ARG, "n must be in [1, batch_max]");
  int rc;
  if ((rc = ensure_loss(c, 1))) return rc;
  rc = dispatch(c, [&](auto ops) {
    using O = decltype(ops);
    int r;
    if ((r = O::upload(c, X, Net::X0, n)) || (r = O::upload(c, Y, Net::YT, n)) || (r = O::forward(c, n)) ||
        (r = O::backward(c, n, 0)))
      return r;
    return VW_OK;
  });
  if (rc) return rc;
  UC(hipStreamSynchronize(c->st));
  if (loss) UC(hipMemcpy(loss, c->loss, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) {
    UC(hipMemcpy(grad, c->g, sizeof(float) * c->nparam, hipMemcpyDeviceToHost));
    const float un = 1.0f / (float)c->nmask; /* the backward pass works at gscale = 1 / n */
    for (int64_t i = 0; i < c->nparam; ++i) grad[i] *= un;
  }
  return VW_OK;
}

VW_API int vw_unet_set_data(void* p, const float* X, const float* Y, int64_t count) {
  Net* c = as_net(p);
  if (!c || !X || !Y || count < 1) return VW_ERR_ARG;
  const int64_t HH = (int64_t)c->N * c->N;
  const size_t ts = c->fp16 ? 2 : 4;
  if (c->X) (void)hipFree(c->X);
  if (c->Y) (void)hipFree(c->Y);
  c->X = c->Y = nullptr;
  c->count = 0;
  UC(hipMalloc(&c->X, (size_t)(count * HH) * ts));
  UC(hipMalloc(&c->Y, (size_t)(count * HH) * ts));
  /* convert in slices through a device FP32 staging buffer */
  const int64_t slice = std::max<int64_t>(1, std::min<int64_t>(count, (int64_t)(64 << 20) / HH));
  float* tmp = nullptr;
  UC(hipMalloc(&tmp, sizeof(float) * slice * HH));
  hipError_t e = hipSuccess;
  for (int64_t s0 = 0; s0 < count && e == hipSuccess; s0 += slice) {
    const int64_t k = std::min(slice, count - s0);
    for (int which = 0; which < 2 && e == hipSuccess; ++which) {
      e = hipMemcpy(tmp, (which ? Y : X) + s0 * HH, sizeof(float) * k * HH, hipMemcpyHostToDevice);
      if (e != hipSuccess) break;
      void* dst = which ? c->Y : c->X;
      if (c->fp16)
        k_from_f32<_Float16><<<nblk(k * HH), 256, 0, nullptr>>>(tmp, (_Float16*)dst + s0 * HH, k * HH);
      else
        k_from_f32<float><<<nblk(k * HH), 256, 0, nullptr>>>(tmp, (float*)dst + s0 * HH, k * HH);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipDeviceSynchronize();
    }
  }
  (void)hipFree(tmp);
  if (e != hipSuccess) return c->fail(VW_ERR_DEVICE, std::string("set_data: ") + hipGetErrorString(e));
  c->count = count;
  return VW_OK;
}

VW_API int vw_unet_train(void* p, int64_t steps, int32_t batch, float lr, uint64_t seed, float* losses) {
  Net* c = as_net(p);
  if (!c || steps < 1 || !(lr > 0.f)) return VW_ERR_ARG;
  if (batch < 1 || batch > c->BM) return c->fail(VW_ERR_ARG, "batch must be in [1, batch_max]");
  if (c->count < batch) return c->fail(VW_ERR_STATE, "vw_unet_train needs vw_unet_set_data with at least one batch");
  /* batch indices for every step: a fresh shuffle of the set per epoch */
  std::vector<int32_t> idx((size_t)(steps * batch));
  std::vector<int32_t> perm((size_t)c->count);
  for (int64_t i = 0; i < c->count; ++i) perm[(size_t)i] = (int32_t)i;
  std::mt19937_64 rng(seed);
  int64_t pos = c->count;
  for (int64_t s = 0; s < steps * batch; ++s) {
    if (pos + 1 > c->count) {
      for (int64_t i = c->count - 1; i > 0; --i) {
        const int64_t j = (int64_t)(rng() % (uint64_t)(i + 1));
        std::swap(perm[(size_t)i], perm[(size_t)j]);
      }
      pos = 0;
    }
    idx[(size_t)s] = perm[(size_t)pos++];
  }
  if (steps * batch > c->idx_cap) {
    if (c->idx) (void)hipFree(c->idx);
    c->idx = nullptr;
    c->idx_cap = 0;
    UC(hipMalloc(&c->idx, sizeof(int32_t) * steps * batch));
    c->idx_cap = steps * batch;
  }
  UC(hipMemcpy(c->idx, idx.data(), sizeof(int32_t) * steps * batch, hipMemcpyHostToDevice));
  int rc;
  if ((rc = ensure_loss(c, steps))) return rc;
  hipEvent_t e0, e1;
  UC(hipEventCreate(&e0));
  UC(hipEventCreate(&e1));
  UC(hipEventRecord(e0, c->st));
  for (int64_t s = 0; s < steps; ++s) {
    rc = dispatch(c, [&](auto ops) {
      using O = decltype(ops);
      int r;
      if ((r = O::gather(c, c->idx + s * batch, batch)) || (r = O::forward(c, batch)) || (r = O::backward(c, batch, s)))
        return r;
      return VW_OK;
    });
    if (rc) break;
    ++c->adam_t;
    const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)c->adam_t));
    const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)c->adam_t));
    k_adam<<<nblk(c->nparam), 256, 0, c->st>>>(c->w, c->g, c->m, c->v, c->nparam, lr, c1, c2);
    if (hipGetLastError() != hipSuccess) {
      rc = c->fail(VW_ERR_DEVICE, "adam launch");
      break;
    }
  }
  (void)hipEventRecord(e1, c->st);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  c->train_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (losses) {
    std::vector<double> l((size_t)steps);
    UC(hipMemcpy(l.data(), c->loss, sizeof(double) * steps, hipMemcpyDeviceToHost));
    for (int64_t s = 0; s < steps; ++s) losses[s] = (float)l[(size_t)s];
  }
  return VW_OK;
}

VW_API int vw_unet_predict(void* p, const float* X, int32_t n, float* out) {
  Net* c = as_net(p);
  if (!c || !X || !out || n < 0) return VW_ERR_ARG;
  const int64_t HH = (int64_t)c->N * c->N;
  for (int32_t b0 = 0; b0 < n; b0 += c->BM) {
    const int k = std::min(c->BM, n - b0);
    int rc = dispatch(c, [&](auto ops) {
      using O = decltype(ops);
      int r;
      if ((r = O::upload(c, X + b0 * HH, Net::X0, k)) || (r = O::forward(c, k)) || (r = O::download(c, Net::OUT, k, out + b0 * HH)))
        return r;
      return VW_OK;
    });
    if (rc) return rc;
  }
  return VW_OK;
}

/* vw_unet_predict without allocation: the FP32 staging of the images in and
 * out goes through this device's scratch array (the kernels and the forward
 * pass are those of vw_unet_predict).  Batches of batch_max images; the
 * largest, min(batch_max, n) x img_n^2 values, must fit the array.  Calls on
 * one device share the array: calls that overlap there race. */
VW_API int vw_unet_predict_noalloc(void* p, const float* X, int32_t n, float* out) {
  Net* c = as_net(p);
  if (!c || !X || !out || n < 0) return VW_ERR_ARG;
  if (!c->scratch_dev) return c->fail(VW_ERR_STATE, "scratch address not resolved");
  const int64_t HH = (int64_t)c->N * c->N;
  const int64_t kmax = std::min<int64_t>(c->BM, n);
  if (kmax * HH > VW_UNET_SCRATCH_FLOATS)
    return c->fail(VW_ERR_RANGE, "vw_unet_predict_noalloc: " + std::to_string(kmax) + " images of " +
                                     std::to_string(c->N) + "^2 exceed the staging array of " +
                                     std::to_string(VW_UNET_SCRATCH_FLOATS) + " values");
  for (int32_t b0 = 0; b0 < n; b0 += c->BM) {
    const int k = std::min(c->BM, n - b0);
    int rc = dispatch(c, [&](auto ops) {
      using O = decltype(ops);
      int r;
      if ((r = O::upload(c, X + b0 * HH, Net::X0, k, false, c->scratch_dev)) || (r = O::forward(c, k)) ||
          (r = O::download(c, Net::OUT, k, out + b0 * HH, c->scratch_dev)))
        return r;
      return VW_OK;
    });
    if (rc) return rc;
  }
  return VW_OK;
}

VW_API int vw_unet_train_ms(void* p, double* ms) {
  Net* c = as_net(p);
  if (!c || !ms) return VW_ERR_ARG;
  *ms = c->train_ms;
  return VW_OK;
}
