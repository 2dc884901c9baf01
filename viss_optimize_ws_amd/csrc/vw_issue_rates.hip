/*
 * vw_issue_rates -- issue cost, on this device, of the vector-ALU instruction
 * classes that the visibility and reconstruction kernels are made of, in
 * units of one FP32 FMA (on gfx950 a wave64 FP32 FMA issues in 2 cycles: 32
 * lanes per SIMD, the published 157.3 TFLOP/s vector FP32 rate).  The VALU
 * rooflines of bench.py weigh each kernel's instructions per term by these
 * costs.
 *
 * Method: 8 independent dependency chains per lane, 8 waves per SIMD on
 * every CU (latency hidden), timed with HIP events.  A class X is measured
 * as a loop of (X, FP32 FMA) pairs, the FMA keeping the values bounded; its
 * cost is (t[X + fma] - t[fma]) / t[fma] with t[fma] a loop of FP32 FMAs
 * alone.  FP64 classes are measured against FP64 FMA loops the same way.
 * Output: one JSON line.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CH 8

enum { OP_FMA32 = 0, OP_SIN32, OP_COS32, OP_FMA64, OP_FRACT64, OP_CVT2, OP_MUL64, OP_ADD64, OP_N };

template <int OP>
__global__ __launch_bounds__(256) void k_rate(float* out, int iters) {
  float x[CH];
  double y[CH];
#pragma unroll
  for (int k = 0; k < CH; ++k) {
    x[k] = 0.1f + 0.01f * (float)(threadIdx.x & 63) + 0.03f * (float)k;
    y[k] = (double)x[k];
  }
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int k = 0; k < CH; ++k) {
        if (OP == OP_FMA32) x[k] = fmaf(x[k], 0.5f, 0.25f);
        if (OP == OP_SIN32) x[k] = fmaf(__builtin_amdgcn_sinf(x[k]), 0.5f, 0.25f);
        if (OP == OP_COS32) x[k] = fmaf(__builtin_amdgcn_cosf(x[k]), 0.5f, 0.25f);
        if (OP == OP_FMA64) y[k] = fma(y[k], 0.5, 0.25);
        if (OP == OP_FRACT64) y[k] = fma(__builtin_amdgcn_fract(y[k]), 0.5, 0.25);
        if (OP == OP_CVT2) y[k] = fma((double)(float)y[k], 0.5, 0.25);
        if (OP == OP_MUL64) y[k] = fma(y[k] * 1.25, 0.5, 0.25);
        if (OP == OP_ADD64) y[k] = fma(y[k] + 0.125, 0.5, 0.25);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < CH; ++k) s += x[k] + (float)y[k];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

#define CK(x)                                                                   \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      exit(1);                                                                  \
    }                                                                           \
  } while (0)

template <int OP>
double time_op(float* out, unsigned blocks, int iters, hipEvent_t a, hipEvent_t b) {
  k_rate<OP><<<blocks, 256>>>(out, iters / 8); /* warm-up */
  CK(hipGetLastError());
  CK(hipEventRecord(a));
  k_rate<OP><<<blocks, 256>>>(out, iters);
  CK(hipGetLastError());
  CK(hipEventRecord(b));
  CK(hipEventSynchronize(b));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, a, b));
  return ms;
}

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 20000;
  if (iters < 8 || iters > 10000000) {
    fprintf(stderr, "iterations must be in [8, 1e7]\n");
    return 2;
  }
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  const int cus = prop.multiProcessorCount;
  const unsigned blocks = (unsigned)cus * 8; /* 4 waves per block: 8 waves per SIMD */
  float* out = nullptr;
  CK(hipMalloc(&out, sizeof(float) * (size_t)blocks * 256));
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  double t[OP_N];
  t[OP_FMA32] = time_op<OP_FMA32>(out, blocks, iters, a, b);
  t[OP_SIN32] = time_op<OP_SIN32>(out, blocks, iters, a, b);
  t[OP_COS32] = time_op<OP_COS32>(out, blocks, iters, a, b);
  t[OP_FMA64] = time_op<OP_FMA64>(out, blocks, iters, a, b);
  t[OP_FRACT64] = time_op<OP_FRACT64>(out, blocks, iters, a, b);
  t[OP_CVT2] = time_op<OP_CVT2>(out, blocks, iters, a, b);
  t[OP_MUL64] = time_op<OP_MUL64>(out, blocks, iters, a, b);
  t[OP_ADD64] = time_op<OP_ADD64>(out, blocks, iters, a, b);
  /* lane instructions of one launch: blocks x 256 lanes x iters x 4 x CH */
  const double lane_ops = (double)blocks * 256.0 * iters * 4.0 * CH;
  const double rate = lane_ops / (t[OP_FMA32] * 1e-3); /* FP32 FMA lane-instructions / s */
  /* gfx950 issues a wave64 FP32 FMA in 2 cycles (32 lanes per SIMD, 4 SIMDs per CU) */
  const double clk_ghz = rate / ((double)cus * 128.0) / 1e9;
  const double u = t[OP_FMA32];
  printf("{\"device\": \"%s\", \"cus\": %d, \"fma32_lane_instr_per_s\": %.6e, \"effective_clock_ghz\": %.4f, "
         "\"slots\": {\"fma_f32\": 1.0, \"sin_f32\": %.3f, \"cos_f32\": %.3f, \"fma_f64\": %.3f, "
         "\"fract_f64\": %.3f, \"cvt_f32_f64_pair\": %.3f, \"mul_f64\": %.3f, \"add_f64\": %.3f}, "
         "\"ms\": {\"fma32\": %.3f, \"sin32+fma32\": %.3f, \"cos32+fma32\": %.3f, \"fma64\": %.3f, "
         "\"fract64+fma64\": %.3f, \"cvt_pair+fma64\": %.3f, \"mul64+fma64\": %.3f, \"add64+fma64\": %.3f}, "
         "\"iters\": %d, \"blocks\": %u}\n",
         prop.gcnArchName, cus, rate, clk_ghz, (t[OP_SIN32] - u) / u, (t[OP_COS32] - u) / u, t[OP_FMA64] / u,
         (t[OP_FRACT64] - t[OP_FMA64]) / u, (t[OP_CVT2] - t[OP_FMA64]) / u, (t[OP_MUL64] - t[OP_FMA64]) / u,
         (t[OP_ADD64] - t[OP_FMA64]) / u, t[OP_FMA32], t[OP_SIN32], t[OP_COS32], t[OP_FMA64], t[OP_FRACT64],
         t[OP_CVT2], t[OP_MUL64], t[OP_ADD64], iters, blocks);
  CK(hipFree(out));
  return 0;
}
