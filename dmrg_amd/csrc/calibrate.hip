// Calibration of the matrix-core ceilings that bound the H_eff kernels: each
// wavefront issues a long run of one MFMA instruction on independent
// accumulators held in registers (no memory traffic inside the loop), so the
// measured rate is what the instruction itself sustains on this device at its
// clock under load.  The roofline figures are still taken against the
// published peaks; these rates say how much of a peak an instruction can
// reach at all.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace dmrgx {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int CAL_CHAINS = 4;  // independent accumulators per wavefront

// kind 0: v_mfma_f32_32x32x16_bf16, 1: v_mfma_f32_16x16x32_bf16,
// 2: v_mfma_f64_16x16x4f64, 3: v_mfma_f32_16x16x4f32; n rounds of CAL_CHAINS
// instructions per wavefront; out gets one value per thread (keeps the work)
template <int KIND>
__global__ __launch_bounds__(256) void k_mfma_run(float* __restrict__ out, int n, float seed) {
  const float x = seed * (float)(threadIdx.x + 1);
  float s = 0.0f;
  if (KIND == 0 || KIND == 1) {
    bf8v a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      a[i] = (__bf16)(x + i);
      b[i] = (__bf16)(seed - i);
    }
    if (KIND == 0) {
      f16v c[CAL_CHAINS];
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q) c[q] = f16v{};
      for (int it = 0; it < n; ++it)
#pragma unroll
        for (int q = 0; q < CAL_CHAINS; ++q) c[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c[q], 0, 0, 0);
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q)
        for (int i = 0; i < 16; ++i) s += c[q][i];
    } else {
      f4v c[CAL_CHAINS];
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q) c[q] = f4v{};
      for (int it = 0; it < n; ++it)
#pragma unroll
        for (int q = 0; q < CAL_CHAINS; ++q) c[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c[q], 0, 0, 0);
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q)
        for (int i = 0; i < 4; ++i) s += c[q][i];
    }
  } else if (KIND == 2) {
    const double a = x, b = seed - 1.0;
    d4v c[CAL_CHAINS];
#pragma unroll
    for (int q = 0; q < CAL_CHAINS; ++q) c[q] = d4v{};
    for (int it = 0; it < n; ++it)
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q) c[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < CAL_CHAINS; ++q)
      for (int i = 0; i < 4; ++i) s += (float)c[q][i];
  } else {
    const float a = x, b = seed - 1.0f;
    f4v c[CAL_CHAINS];
#pragma unroll
    for (int q = 0; q < CAL_CHAINS; ++q) c[q] = f4v{};
    for (int it = 0; it < n; ++it)
#pragma unroll
      for (int q = 0; q < CAL_CHAINS; ++q) c[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c[q], 0, 0, 0);
#pragma unroll
    for (int q = 0; q < CAL_CHAINS; ++q)
      for (int i = 0; i < 4; ++i) s += c[q][i];
  }
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

// flop of one instruction of each kind
constexpr double kCalFlop[4] = {2.0 * 32 * 32 * 16, 2.0 * 16 * 16 * 32, 2.0 * 16 * 16 * 4, 2.0 * 16 * 16 * 4};

}  // namespace
}  // namespace dmrgx

using namespace dmrgx;

#define CAPI_BEGIN try {
#define CAPI_END                                   \
  return 0;                                        \
  }                                                \
  catch (const dmrgx::Error& e) {                  \
    set_last_error(e.what());                      \
    return e.code;                                 \
  }                                                \
  catch (const std::exception& e) {                \
    set_last_error(e.what());                      \
    return -2;                                     \
  }

extern "C" {

int dmrgx_calibrate_mfma(int kind, int waves_per_simd, int rounds, double* tflops) {
  CAPI_BEGIN
  DMRGX_CHECK(kind >= 0 && kind < 4, "calibrate_mfma: kind must be 0..3");
  DMRGX_CHECK(waves_per_simd >= 1 && waves_per_simd <= 8, "calibrate_mfma: waves per SIMD must be 1..8");
  DMRGX_CHECK(rounds >= 1 && rounds <= (1 << 20), "calibrate_mfma: rounds must be 1..2^20");
  DMRGX_CHECK(tflops != nullptr, "calibrate_mfma: null result pointer");
  ensure_init();
  hipStream_t s = rt().stream;
  // one 256-thread work-group holds one wavefront per SIMD of a CU
  const int grid = rt().num_cu * waves_per_simd;
  DevBuf out((size_t)grid * 256 * sizeof(float));
  hipEvent_t ev[2];
  for (auto& e : ev) DMRGX_HIP(hipEventCreate(&e));
  auto launch = [&]() {
    switch (kind) {
      case 0: hipLaunchKernelGGL(k_mfma_run<0>, dim3(grid), dim3(256), 0, s, (float*)out.ptr, rounds, 1e-3f); break;
      case 1: hipLaunchKernelGGL(k_mfma_run<1>, dim3(grid), dim3(256), 0, s, (float*)out.ptr, rounds, 1e-3f); break;
      case 2: hipLaunchKernelGGL(k_mfma_run<2>, dim3(grid), dim3(256), 0, s, (float*)out.ptr, rounds, 1e-3f); break;
      default: hipLaunchKernelGGL(k_mfma_run<3>, dim3(grid), dim3(256), 0, s, (float*)out.ptr, rounds, 1e-3f); break;
    }
    DMRGX_HIP(hipGetLastError());
  };
  launch();  // warm-up (clocks ramp)
  DMRGX_HIP(hipEventRecord(ev[0], s));
  launch();
  DMRGX_HIP(hipEventRecord(ev[1], s));
  DMRGX_HIP(hipEventSynchronize(ev[1]));
  float ms = 0;
  DMRGX_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto& e : ev) DMRGX_HIP(hipEventDestroy(e));
  const double flop = kCalFlop[kind] * CAL_CHAINS * (double)rounds * grid * 4.0;  // 4 wavefronts per group
  *tflops = flop / (ms * 1e-3) / 1e12;
  CAPI_END
}

}  // extern "C"
