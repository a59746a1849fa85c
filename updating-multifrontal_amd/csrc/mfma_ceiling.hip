// mfma_ceiling.hip -- register-only throughput of the instructions the kernels use on this GPU:
// v_mfma_f64_16x16x4, v_mfma_f32_16x16x4, v_mfma_f32_16x16x32_bf16 (independent accumulator
// chains, no memory) and FP64 vector FMA, at 1-8 waves per SIMD.  The attainable rates set the compute roofline the
// kernels are judged against beside the published peaks.
#include <hip/hip_runtime.h>
#include <cstdio>
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
template <int NA>
__global__ __launch_bounds__(256) void k64(double* out, int iters) {
  d4 acc[NA];
  for (int i = 0; i < NA; i++) acc[i] = (d4){0, 0, 0, 0};
  double a = 1.0 + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < NA; i++) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0;
  for (int i = 0; i < NA; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
template <int NA>
__global__ __launch_bounds__(256) void k32(float* out, int iters) {
  f4 acc[NA];
  for (int i = 0; i < NA; i++) acc[i] = (f4){0, 0, 0, 0};
  float a = 1.0f + threadIdx.x * 1e-7f, b = 1.0f - threadIdx.x * 1e-7f;
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < NA; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[i], 0, 0, 0);
  }
  float s = 0;
  for (int i = 0; i < NA; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
template <int NA>
__global__ __launch_bounds__(256) void kbf(float* out, int iters) {
  f4 acc[NA];
  for (int i = 0; i < NA; i++) acc[i] = (f4){0, 0, 0, 0};
  bf8 a, b;
  for (int j = 0; j < 8; j++) {
    a[j] = (__bf16)(1.0f + (threadIdx.x + j) * 1e-3f);
    b[j] = (__bf16)(1.0f - (threadIdx.x + j) * 1e-3f);
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < NA; i++) acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i], 0, 0, 0);
  }
  float s = 0;
  for (int i = 0; i < NA; i++) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
// FP64 vector FMA throughput for comparison
__global__ __launch_bounds__(256) void kfma(double* out, int iters) {
  double x[8];
  for (int i = 0; i < 8; i++) x[i] = 1.0 + (threadIdx.x + i) * 1e-9;
  const double m = 0.999999, c = 1e-9;
  for (int it = 0; it < iters; it++)
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = fma(x[i], m, c);
  double s = 0;
  for (int i = 0; i < 8; i++) s += x[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
int main() {
  int ncu = 0; hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0);
  int clk = 0; hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, 0);
  printf("CUs %d, clock %d kHz\n", ncu, clk);
  const int maxb = ncu * 8;
  double* d; if (hipMalloc(&d, sizeof(double) * maxb * 256)) return 1;
  hipEvent_t e0, e1; hipEventCreate(&e0); hipEventCreate(&e1);
  const int iters = 4096;
  for (int wpc : {1, 2, 4, 8}) {  // workgroups (4 waves each) per CU -> waves per SIMD
    const int nb = ncu * wpc;
    float ms;
    hipLaunchKernelGGL(k64<4>, dim3(nb), dim3(256), 0, 0, d, 16); hipDeviceSynchronize();
    hipEventRecord(e0); hipLaunchKernelGGL(k64<4>, dim3(nb), dim3(256), 0, 0, d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    double fl = (double)nb * 4 * iters * 4 * 2048.0;
    printf("f64 mfma 16x16x4, %d waves/SIMD, 4 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(k64<8>, dim3(nb), dim3(256), 0, 0, d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 4 * iters * 8 * 2048.0;
    printf("f64 mfma 16x16x4, %d waves/SIMD, 8 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(k32<4>, dim3(nb), dim3(256), 0, 0, (float*)d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 4 * iters * 4 * 2048.0;
    printf("f32 mfma 16x16x4, %d waves/SIMD, 4 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(k32<8>, dim3(nb), dim3(256), 0, 0, (float*)d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 4 * iters * 8 * 2048.0;
    printf("f32 mfma 16x16x4, %d waves/SIMD, 8 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(kbf<4>, dim3(nb), dim3(256), 0, 0, (float*)d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 4 * iters * 4 * 16384.0;
    printf("bf16 mfma 16x16x32, %d waves/SIMD, 4 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(kbf<8>, dim3(nb), dim3(256), 0, 0, (float*)d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 4 * iters * 8 * 16384.0;
    printf("bf16 mfma 16x16x32, %d waves/SIMD, 8 chains: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
    hipEventRecord(e0); hipLaunchKernelGGL(kfma, dim3(nb), dim3(256), 0, 0, d, iters); hipEventRecord(e1); hipEventSynchronize(e1);
    hipEventElapsedTime(&ms, e0, e1);
    fl = (double)nb * 256 * iters * 8 * 2.0;
    printf("f64 vector fma,   %d waves/SIMD: %.1f TFLOP/s\n", wpc, fl / ms / 1e9);
  }
  return 0;
}
