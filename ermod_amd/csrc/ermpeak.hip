// Peak-rate microbenchmarks for the roofline: the device's FP32 FMA rate
// (v_fma_f32, one lane op per FMA: the library is compiled with the packed
// FP32 forms disabled) and its FP64 matrix rate (v_mfma_f64_16x16x4f64, the
// instruction the correlation kernel uses).  Each thread or wave runs
// independent dependency chains long enough to hide the instruction
// latency; the results are stored so that nothing is optimised away.
// Built as libermpeak.so beside libermgpu.so; nothing in libermgpu.so uses
// it.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

constexpr int FMA_CHAINS = 8;

// every thread: FMA_CHAINS chains of iters x 16 dependent FMAs
__global__ __launch_bounds__(256) void k_peak_fma_f32(int iters, float x, float y, float *out) {
  float a[FMA_CHAINS];
#pragma unroll
  for (int c = 0; c < FMA_CHAINS; ++c) a[c] = (float)(threadIdx.x + c);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
#pragma unroll
      for (int c = 0; c < FMA_CHAINS; ++c) a[c] = __builtin_fmaf(a[c], x, y);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < FMA_CHAINS; ++c) s += a[c];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef double pk_dbl4 __attribute__((ext_vector_type(4)));
constexpr int MFMA_CHAINS = 4;

// every wave: MFMA_CHAINS chains of iters x 8 dependent 16x16x4 FP64 MFMAs
// (2048 flop each)
__global__ __launch_bounds__(256) void k_peak_mfma_f64(int iters, double x, double y, double *out) {
  pk_dbl4 acc[MFMA_CHAINS];
#pragma unroll
  for (int c = 0; c < MFMA_CHAINS; ++c) acc[c] = (pk_dbl4){0.0, 0.0, 0.0, (double)c};
  const double a = x + threadIdx.x, b = y - threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
#pragma unroll
      for (int c = 0; c < MFMA_CHAINS; ++c) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < MFMA_CHAINS; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// Launches on stream (NULL: the null stream) blocks x 256 threads; out holds
// blocks x 256 floats.  Returns the flops launched (2 per FMA), or -1 when
// the launch fails.
double erm_peak_fma_f32(void *stream, int32_t blocks, int32_t iters, float *out) {
  hipLaunchKernelGGL(k_peak_fma_f32, dim3(blocks), dim3(256), 0, (hipStream_t)stream, iters, 0.999f, 1e-3f, out);
  if (hipGetLastError() != hipSuccess) return -1.0;
  return 2.0 * FMA_CHAINS * 16.0 * iters * 256.0 * blocks;
}

// As erm_peak_fma_f32 for the FP64 MFMA kernel; out holds blocks x 256
// doubles.  Returns the flops launched (2048 per MFMA per wave).
double erm_peak_mfma_f64(void *stream, int32_t blocks, int32_t iters, double *out) {
  hipLaunchKernelGGL(k_peak_mfma_f64, dim3(blocks), dim3(256), 0, (hipStream_t)stream, iters, 1e-3, 2e-3, out);
  if (hipGetLastError() != hipSuccess) return -1.0;
  return 2048.0 * MFMA_CHAINS * 8.0 * iters * 4.0 * blocks;
}

}
