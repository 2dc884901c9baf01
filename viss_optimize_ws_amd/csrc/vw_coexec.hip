/*
 * vw_coexec -- how much vector-ALU work of a wave overlaps the matrix core
 * on this device.  The batched Viss and recon kernels interleave per-term
 * phasor VALU work (FP64 phase, turn reduction, FP32 sin / cos, conversions,
 * BF16 splits) with v_mfma_f32_16x16x32_bf16 products; their time is
 * max(VALU, MFMA) where the two overlap and VALU + MFMA where they do not.
 *
 * Method: a loop of 8 independent products (8 accumulators) per iteration,
 * alone (MFMA), the same loop's VALU work alone (NV instructions of one
 * class per product, 8 independent chains), and both in one loop with
 * scheduling groups that issue each product followed by its NV VALU
 * instructions (checked in the build's AMD assembly: the loop bodies hold
 * nothing else).  hidden = (t_mfma + t_valu - t_both) / min(t_mfma, t_valu):
 * 100% is full overlap, 0% none.  Accumulators live in VGPRs (built with
 * -amdgpu-mfma-vgpr-form, as the kernels' are) so that no accumulator moves
 * enter the loops.  Output: one JSON line.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef __bf16 vw_bf8 __attribute__((ext_vector_type(8)));
typedef float vw_f4 __attribute__((ext_vector_type(4)));

enum { C_FMA32 = 0, C_FMA64, C_SIN32, C_FRACT64, C_CVT, C_N };
static const char* kName[C_N] = {"fma_f32", "fma_f64", "sin_f32", "fract_f64", "cvt_f32_f64"};

template <int C>
__device__ __forceinline__ void vop(float& x, double& y) {
  if (C == C_FMA32) x = fmaf(x, 0.999f, 0.001f);
  if (C == C_FMA64) y = fma(y, 0.999, 0.001);
  if (C == C_SIN32) x = __builtin_amdgcn_sinf(x);
  if (C == C_FRACT64) y = __builtin_amdgcn_fract(y + 0.37);
  if (C == C_CVT) x = x + (float)y;
}

#define VW_MF(c) c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0)
#define VW_V8                                                                                 \
  vop<C>(x0, y0); vop<C>(x1, y1); vop<C>(x2, y2); vop<C>(x3, y3);                             \
  vop<C>(x4, y4); vop<C>(x5, y5); vop<C>(x6, y6); vop<C>(x7, y7)

/* MODE 0: products only; 1: VALU only; 2: both, interleaved */
template <int MODE, int C, int NV>
__global__ __launch_bounds__(256) void k_coexec(float* out, int iters) {
  vw_f4 c0 = {}, c1 = {}, c2 = {}, c3 = {}, c4 = {}, c5 = {}, c6 = {}, c7 = {};
  vw_bf8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)(0.001f * (float)(threadIdx.x + j));
    b[j] = (__bf16)(0.002f * (float)j);
  }
  const float t = 1e-3f * (float)threadIdx.x;
  float x0 = t, x1 = t + 0.1f, x2 = t + 0.2f, x3 = t + 0.3f, x4 = t + 0.4f, x5 = t + 0.5f, x6 = t + 0.6f,
        x7 = t + 0.7f;
  double y0 = x0, y1 = x1, y2 = x2, y3 = x3, y4 = x4, y5 = x5, y6 = x6, y7 = x7;
  for (int i = 0; i < iters; ++i) {
    if (MODE != 1) {
      VW_MF(c0); VW_MF(c1); VW_MF(c2); VW_MF(c3); VW_MF(c4); VW_MF(c5); VW_MF(c6); VW_MF(c7);
    }
    if (MODE != 0) {
#pragma unroll
      for (int r = 0; r < NV; ++r) { VW_V8; }
    }
    if (MODE == 2) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* one product */
        __builtin_amdgcn_sched_group_barrier(0x002, NV, 0); /* NV VALU */
      }
    }
  }
  const vw_f4 c = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
  out[(size_t)blockIdx.x * 256 + threadIdx.x] =
      c[0] + c[1] + c[2] + c[3] + x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7 +
      (float)(y0 + y1 + y2 + y3 + y4 + y5 + y6 + y7);
}

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      exit(1);                                                                  \
    }                                                                           \
  } while (0)

template <int MODE, int C, int NV>
double timed(float* out, unsigned blocks, int iters) {
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  k_coexec<MODE, C, NV><<<blocks, 256>>>(out, iters / 8); /* warm-up */
  CK(hipGetLastError());
  CK(hipEventRecord(e0));
  k_coexec<MODE, C, NV><<<blocks, 256>>>(out, iters);
  CK(hipGetLastError());
  CK(hipEventRecord(e1));
  CK(hipEventSynchronize(e1));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, e0, e1));
  CK(hipEventDestroy(e0));
  CK(hipEventDestroy(e1));
  return ms;
}

template <int C, int NV>
void row(float* out, unsigned blocks, int iters, double t_mfma, bool& first) {
  const double tv = timed<1, C, NV>(out, blocks, iters), tb = timed<2, C, NV>(out, blocks, iters);
  const double lo = tv < t_mfma ? tv : t_mfma;
  printf("%s{\"class\": \"%s\", \"valu_per_product\": %d, \"ms_valu\": %.4f, \"ms_both\": %.4f, \"hidden\": %.3f}",
         first ? "" : ", ", kName[C], NV, tv, tb, (t_mfma + tv - tb) / lo);
  first = false;
}

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 20000;
  if (iters < 8 || iters > 10000000) {
    fprintf(stderr, "iterations must be in [8, 1e7]\n");
    return 2;
  }
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  printf("{\"device\": \"%s\", \"cus\": %d, \"iters\": %d, \"runs\": [", prop.gcnArchName, prop.multiProcessorCount,
         iters);
  bool first_run = true;
  for (int wps : {2, 4}) {
    const unsigned blocks = (unsigned)prop.multiProcessorCount * (unsigned)wps; /* 4 waves per block */
    float* out = nullptr;
    CK(hipMalloc(&out, sizeof(float) * (size_t)blocks * 256));
    const double tm = timed<0, C_FMA32, 1>(out, blocks, iters);
    /* products per SIMD: iters x 8 per wave, wps waves */
    printf("%s{\"waves_per_simd\": %d, \"ms_mfma\": %.4f, \"ns_per_product_per_simd\": %.4f, \"rows\": [",
           first_run ? "" : ", ", wps, tm, tm * 1e6 / ((double)iters * 8.0 * wps));
    first_run = false;
    bool first = true;
    row<C_FMA32, 2>(out, blocks, iters, tm, first);
    row<C_FMA32, 4>(out, blocks, iters, tm, first);
    row<C_FMA64, 2>(out, blocks, iters, tm, first);
    row<C_FMA64, 4>(out, blocks, iters, tm, first);
    row<C_SIN32, 1>(out, blocks, iters, tm, first);
    row<C_SIN32, 2>(out, blocks, iters, tm, first);
    row<C_FRACT64, 2>(out, blocks, iters, tm, first);
    row<C_CVT, 2>(out, blocks, iters, tm, first);
    printf("]}");
    CK(hipFree(out));
  }
  printf("]}\n");
  return 0;
}
