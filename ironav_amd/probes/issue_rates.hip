// issue_rates.hip — the VALU issue rates that bound the SAD kernels, measured on the device.
//
// Each kernel runs 8 independent chains of one instruction per lane in a long loop over 4096
// workgroups of 256 threads and prints lane-operations per second and per CU-clock (at 2.4 GHz).
// A full-rate wave64 VALU instruction occupies a 16-lane SIMD for 4 clocks, i.e. 64 lane-ops per
// CU-clock over the 4 SIMDs; v_sad_u8 does 4 absolute differences per lane-op and v_qsad_pk_u16_u8 /
// v_mqsad_u32_u8 16 (4 adjacent 4-sample SADs).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
typedef unsigned long long u64;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
constexpr int ITERS = 4096, CH = 8;

__global__ void k_sad(uint32_t *out, uint32_t s) {
  uint32_t a[CH];
  for (int i = 0; i < CH; i++) a[i] = threadIdx.x + i;
  uint32_t r = s ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = __builtin_amdgcn_sad_u8(r + i, s, a[i]);
  }
  uint32_t t = 0;
  for (int i = 0; i < CH; i++) t += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}
__global__ void k_sadhi(uint32_t *out, uint32_t s) {
  uint32_t a[CH];
  for (int i = 0; i < CH; i++) a[i] = threadIdx.x + i;
  uint32_t r = s ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = __builtin_amdgcn_sad_hi_u8(r + i, s, a[i]);
  }
  uint32_t t = 0;
  for (int i = 0; i < CH; i++) t += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}
__global__ void k_qsad(uint32_t *out, uint32_t s) {
  u64 a[CH];
  for (int i = 0; i < CH; i++) a[i] = threadIdx.x + i;
  u64 r = ((u64)s << 32) ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = __builtin_amdgcn_qsad_pk_u16_u8(r + i, s, a[i]);
  }
  u64 t = 0;
  for (int i = 0; i < CH; i++) t += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)t ^ (uint32_t)(t >> 32);
}
__global__ void k_mqsad(uint32_t *out, uint32_t s) {
  u32x4 a[CH];
  for (int i = 0; i < CH; i++) a[i] = u32x4{threadIdx.x, (uint32_t)i, 0u, 1u};
  u64 r = ((u64)s << 32) ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = __builtin_amdgcn_mqsad_u32_u8(r + i, s, a[i]);
  }
  uint32_t t = 0;
  for (int i = 0; i < CH; i++) t += a[i][0] ^ a[i][1] ^ a[i][2] ^ a[i][3];
  out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}
__global__ void k_min3(uint32_t *out, uint32_t s) {
  uint32_t a[CH];
  for (int i = 0; i < CH; i++) a[i] = threadIdx.x + i;
  uint32_t r = s ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = min(a[i], min(r + i, s - (uint32_t)it));
  }
  uint32_t t = 0;
  for (int i = 0; i < CH; i++) t += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}
__global__ void k_add(uint32_t *out, uint32_t s) {
  uint32_t a[CH];
  for (int i = 0; i < CH; i++) a[i] = threadIdx.x + i;
  uint32_t r = s ^ threadIdx.x;
  for (int it = 0; it < ITERS; it++) {
#pragma unroll
    for (int i = 0; i < CH; i++) {
      a[i] = a[i] + (r ^ (uint32_t)i);
      asm volatile("" : "+v"(a[i]));
    }
  }
  uint32_t t = 0;
  for (int i = 0; i < CH; i++) t += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = t;
}

typedef void (*kfn)(uint32_t *, uint32_t);
int main() {
  const int blocks = 256 * 16, threads = 256;
  uint32_t *out;
  if (hipMalloc(&out, (size_t)blocks * threads * 4) != hipSuccess) return 1;  // one word per thread
  struct { const char *n; kfn f; } ks[] = {{"v_sad_u8", k_sad}, {"v_sad_hi_u8", k_sadhi}, {"v_qsad_pk_u16_u8", k_qsad},
                                           {"v_mqsad_u32_u8", k_mqsad}, {"v_min3_u32", k_min3}, {"v_add_u32", k_add}};
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  for (auto &k : ks) {
    hipLaunchKernelGGL(k.f, dim3(blocks), dim3(threads), 0, 0, out, 12345u);
    (void)hipEventRecord(e0, 0);
    for (int rep = 0; rep < 5; rep++) hipLaunchKernelGGL(k.f, dim3(blocks), dim3(threads), 0, 0, out, 777u + rep);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double ops = 5.0 * blocks * threads * (double)ITERS * CH;
    printf("%-18s %8.3f ms  %8.1f G lane-ops/s  (%.3f per CU-clock at 2.4 GHz)\n", k.n, ms, ops / (ms * 1e6),
           ops / (ms * 1e-3) / 256 / 2.4e9);
  }
  (void)hipFree(out);
  return 0;
}
