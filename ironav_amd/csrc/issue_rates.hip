// issue_rates.hip — the VALU issue rates that bound the SAD kernels, measured on the device
// (libironav_probe.so, run by `bench.py --config issue-rates`).
//
// Each kernel runs 8 independent chains of one instruction per lane in a long loop over 4096
// workgroups of 256 threads; the result is lane-operations per second.
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

// The issue rate of each probed instruction in lane-operations per second, in the order of
// ia_probe_issue_rate_name: v_sad_u8, v_sad_hi_u8, v_qsad_pk_u16_u8, v_mqsad_u32_u8, v_min3_u32,
// v_add_u32. Returns the count written (at most n), or -1 if a HIP call failed.
extern "C" int ia_probe_issue_rates(double *out, int n) {
  const int blocks = 256 * 16, threads = 256;
  const kfn ks[] = {k_sad, k_sadhi, k_qsad, k_mqsad, k_min3, k_add};
  const int nk = (int)(sizeof ks / sizeof ks[0]);
  uint32_t *buf;
  if (hipMalloc(&buf, (size_t)blocks * threads * 4) != hipSuccess) return -1;  // one word per thread
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1;
  int w = 0;
  for (int k = 0; k < nk && w < n; k++) {
    hipLaunchKernelGGL(ks[k], dim3(blocks), dim3(threads), 0, 0, buf, 12345u);
    if (hipEventRecord(e0, 0) != hipSuccess) return -1;
    for (int rep = 0; rep < 5; rep++) hipLaunchKernelGGL(ks[k], dim3(blocks), dim3(threads), 0, 0, buf, 777u + rep);
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1;
    out[w++] = 5.0 * blocks * threads * (double)ITERS * CH / (ms * 1e-3);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  return w;
}

extern "C" const char *ia_probe_issue_rate_name(int k) {
  static const char *names[] = {"v_sad_u8", "v_sad_hi_u8", "v_qsad_pk_u16_u8", "v_mqsad_u32_u8", "v_min3_u32", "v_add_u32"};
  return k >= 0 && k < 6 ? names[k] : nullptr;
}
