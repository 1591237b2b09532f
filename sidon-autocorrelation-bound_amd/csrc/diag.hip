// diag.hip -- measured FP32 ceilings of the two engines the O(n^2) products
// could use on this device, under this library's compile flags.
//
//  * VALU: independent v_fma_f32 chains in registers (packed FP32 is compiled
//    off for the whole library, so this is the ceiling of any VALU
//    formulation of the convolution / correlation);
//  * MFMA: independent v_mfma_f32_16x16x4f32 accumulation chains.
// Both run with every CU busy and no memory traffic; the result decides the
// engine of the local-search kernels (the matrix-core path is used because it
// measures well above the VALU ceiling).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "runtime.hpp"

namespace sidon {
namespace {

constexpr int kChains = 8;

__global__ __launch_bounds__(256) void k_valu_fma(float* out, int reps, float x, float y) {
  float acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) acc[c] = x + (float)(threadIdx.x + c);
  for (int r = 0; r < reps; ++r) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int c = 0; c < kChains; ++c) acc[c] = __builtin_fmaf(acc[c], y, x);
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += acc[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mfma_f32(float* out, int reps, float x) {
  f32x4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = f32x4{x, x, x, x};
  const float a = x + (float)threadIdx.x, b = x - (float)threadIdx.x;
  for (int r = 0; r < reps; ++r) {
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] = mfma16x16x4(a, b, acc[c]);
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace
}  // namespace sidon

using namespace sidon;

extern "C" int sidon_measure_fp32_ceilings(float* scratch, int scratch_len, double* valu_tflops,
                                           double* mfma_tflops, hipStream_t stream) {
  int rc = SIDON_OK;
  DeviceState* ds = device_state(&rc);
  if (!ds) return rc;
  const int blocks = ds->cu_count * 8;
  if (!scratch || scratch_len < blocks * 256 || !valu_tflops || !mfma_tflops) return SIDON_ERR_ARG;
  hipEvent_t e0, e1;
  SIDON_HIP_TRY(hipEventCreate(&e0));
  SIDON_HIP_TRY(hipEventCreate(&e1));
  const int reps = 4096;
  float ms = 0.0f;
  // VALU: blocks * 256 threads * reps * 16 * chains FMAs, 2 FLOP each
  hipLaunchKernelGGL(k_valu_fma, dim3(blocks), dim3(256), 0, stream, scratch, 16, 1.0f, 0.999f);
  SIDON_HIP_TRY(hipEventRecord(e0, stream));
  hipLaunchKernelGGL(k_valu_fma, dim3(blocks), dim3(256), 0, stream, scratch, reps, 1.0f, 0.999f);
  SIDON_HIP_TRY(hipEventRecord(e1, stream));
  SIDON_HIP_TRY(hipEventSynchronize(e1));
  SIDON_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  *valu_tflops = 2.0 * blocks * 256.0 * reps * 16.0 * kChains / (ms * 1e-3) / 1e12;
  // MFMA: per wave reps * 8 * 4 instructions of 16*16*4 MACs
  hipLaunchKernelGGL(k_mfma_f32, dim3(blocks), dim3(256), 0, stream, scratch, 16, 1.0f);
  SIDON_HIP_TRY(hipEventRecord(e0, stream));
  hipLaunchKernelGGL(k_mfma_f32, dim3(blocks), dim3(256), 0, stream, scratch, reps, 1.0f);
  SIDON_HIP_TRY(hipEventRecord(e1, stream));
  SIDON_HIP_TRY(hipEventSynchronize(e1));
  SIDON_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  const double waves = blocks * 4.0;
  *mfma_tflops = 2.0 * waves * reps * 8.0 * 4.0 * 1024.0 / (ms * 1e-3) / 1e12;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  SIDON_HIP_TRY(hipGetLastError());
  return SIDON_OK;
}
