// hostwave.h — runs a 64-lane HIP workgroup's device code on the host, one fiber per lane.
//
// TEST INFRASTRUCTURE: a CPU build of the analysis kernel (ironav_amd/csrc/analyse.hip) under the
// address and undefined-behaviour sanitizers, so that every private array, LDS field and global
// buffer the kernel indexes is bounds-checked on the host. Lanes run as fibers on one host thread
// (ucontext, with the address sanitizer's fiber annotations); every cross-lane operation the kernel
// uses (__syncthreads, readlane, readfirstlane, the DPP row shifts and row broadcasts of its wave
// reductions) is a barrier over the 64 lanes followed by a read of the values the lanes published.
// Workgroups of a launch run one after another, so `__shared__` variables are function statics
// shared by the lanes of the running workgroup.
//
// A lane that reaches a cross-lane operation the others do not reach is a kernel bug (on the GPU
// such a reduction would read lanes that never wrote); the scheduler reports it and aborts.
#pragma once

#include <assert.h>
#include <math.h>
#include <ucontext.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#if defined(__SANITIZE_ADDRESS__)
#define HOSTWAVE_ASAN 1
#include <sanitizer/common_interface_defs.h>
#else
#define HOSTWAVE_ASAN 0
#endif

#include <algorithm>
#include <atomic>
#include <type_traits>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
// the kernels' global-memory pointer qualifier (address_space(1) for the GPU): plain host memory here
#define IA_GLOBAL_AS
#define __constant__
// the occupancy attribute of the GPU build means nothing on the host
#define amdgpu_waves_per_eu(...) unused

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
typedef void *hipStream_t;
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct hipDeviceProp_t {
  int multiProcessorCount;
};
struct HwIdx {
  unsigned x, y, z;
};

namespace hostwave {

constexpr int kLanes = 64;
constexpr size_t kStack = 8u << 20;  // per lane (sanitizer frames are large)

// The lanes of the running workgroup are fibers on one host thread, switched at every cross-lane
// operation: a lane runs until it reaches the next barrier, then the next lane runs. A barrier
// completes when every lane has reached it.
struct Wave {
  ucontext_t sched, lane[kLanes];
  char *stack[kLanes] = {};
  int done[kLanes] = {};
  int at_barrier[kLanes] = {};
  int cur = 0;
  HwIdx block{0, 0, 0};
  void *fake_stack = nullptr;
  const void *sched_bottom = nullptr;
  size_t sched_size = 0;
};

inline Wave &wave() {
  static Wave w;
  return w;
}

inline void to_sched() {
  Wave &w = wave();
#if HOSTWAVE_ASAN
  void *fs = nullptr;
  __sanitizer_start_switch_fiber(&fs, w.sched_bottom, w.sched_size);
  swapcontext(&w.lane[w.cur], &w.sched);
  __sanitizer_finish_switch_fiber(fs, nullptr, nullptr);
#else
  swapcontext(&w.lane[w.cur], &w.sched);
#endif
}

inline void barrier_wait() {
  Wave &w = wave();
  w.at_barrier[w.cur] = 1;
  to_sched();
}

inline int64_t *slots() {
  static int64_t s[kLanes];
  return s;
}

inline int cur_lane() { return wave().cur; }

// Every lane publishes v; returns the values of all lanes after the exchange (copied before the
// second barrier lets the slots be reused).
inline void exchange(int64_t v, int64_t *all) {
  slots()[cur_lane()] = v;
  barrier_wait();
  memcpy(all, slots(), sizeof(int64_t) * kLanes);
  barrier_wait();
}

inline void syncthreads() { barrier_wait(); }

template <class T>
inline T readlane(T v, int k) {
  static_assert(sizeof(T) <= 8, "readlane of a value wider than 64 bits");
  int64_t all[kLanes], x = 0;
  memcpy(&x, &v, sizeof(T));
  exchange(x, all);
  if (k < 0 || k >= kLanes) {
    fprintf(stderr, "hostwave: readlane of lane %d\n", k);
    abort();
  }
  T r;
  memcpy(&r, &all[k], sizeof(T));
  return r;
}

// __builtin_amdgcn_update_dpp(old, src, ctrl, row_mask, bank_mask, bound_ctrl) for the controls the
// kernels use: quad_perm (0x00..0xff), row_shr:1..15 (0x111..0x11f), row_mirror (0x140),
// row_half_mirror (0x141), row_bcast:15 (0x142), row_bcast:31 (0x143). A lane
// whose row or bank is masked off, or whose source lane is outside the row (shifts) or below the
// broadcast source, keeps `old` (bound_ctrl 0) or takes 0 (bound_ctrl 1).
inline int update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  int64_t all[kLanes];
  exchange(src, all);
  const int l = cur_lane(), row = l >> 4, bank = (l >> 2) & 3;
  if (!((row_mask >> row) & 1) || !((bank_mask >> bank) & 1)) return old;
  int from = -1;
  if (ctrl >= 0 && ctrl <= 0xff) {
    from = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);  // quad_perm
  } else if (ctrl >= 0x111 && ctrl <= 0x11f) {
    const int n = ctrl - 0x110;
    if ((l & 15) >= n) from = l - n;
  } else if (ctrl == 0x140) {
    from = (l & ~15) | (15 - (l & 15));  // row_mirror
  } else if (ctrl == 0x141) {
    from = (l & ~7) | (7 - (l & 7));  // row_half_mirror
  } else if (ctrl == 0x142) {
    if (row >= 1) from = row * 16 - 1;
  } else if (ctrl == 0x143) {
    if (row >= 2) from = 31;
  } else {
    fprintf(stderr, "hostwave: DPP control 0x%x is not emulated\n", ctrl);
    abort();
  }
  if (from < 0) return bound_ctrl ? 0 : old;
  return (int)all[from];
}

// __ballot(pred): the lanes whose predicate holds, as a 64-bit mask
inline uint64_t ballot(int pred) {
  int64_t all[kLanes];
  exchange(pred != 0, all);
  uint64_t m = 0;
  for (int i = 0; i < kLanes; i++) m |= (uint64_t)(all[i] != 0) << i;
  return m;
}

struct Thunk {
  void (*fn)(void *);
  void *arg;
};

inline void lane_entry(unsigned lo, unsigned hi) {
  Thunk *t = (Thunk *)(((uintptr_t)hi << 32) | lo);
  Wave &w = wave();
#if HOSTWAVE_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &w.sched_bottom, &w.sched_size);
#endif
  t->fn(t->arg);
  w.done[w.cur] = 1;
#if HOSTWAVE_ASAN
  __sanitizer_start_switch_fiber(nullptr, w.sched_bottom, w.sched_size);
#endif
  setcontext(&w.sched);
}

// Runs fn(arg) as the 64 lanes of workgroup `block`.
inline void run_workgroup(unsigned block, void (*fn)(void *), void *arg) {
  Wave &w = wave();
  w.block = HwIdx{block, 0, 0};
  Thunk t{fn, arg};
  const uintptr_t tp = (uintptr_t)&t;
  for (int l = 0; l < kLanes; l++) {
    if (!w.stack[l]) w.stack[l] = (char *)malloc(kStack);
    w.done[l] = 0;
    w.at_barrier[l] = 0;
    getcontext(&w.lane[l]);
    w.lane[l].uc_stack.ss_sp = w.stack[l];
    w.lane[l].uc_stack.ss_size = kStack;
    w.lane[l].uc_link = nullptr;
    makecontext(&w.lane[l], (void (*)())lane_entry, 2, (unsigned)(tp & 0xffffffffu), (unsigned)(tp >> 32));
  }
  for (;;) {
    int ndone = 0, nbar = 0;
    for (int l = 0; l < kLanes; l++) {
      if (w.done[l]) {
        ndone++;
        continue;
      }
      w.cur = l;
      w.at_barrier[l] = 0;
#if HOSTWAVE_ASAN
      __sanitizer_start_switch_fiber(&w.fake_stack, w.stack[l], kStack);
      swapcontext(&w.sched, &w.lane[l]);
      __sanitizer_finish_switch_fiber(w.fake_stack, nullptr, nullptr);
#else
      swapcontext(&w.sched, &w.lane[l]);
#endif
      if (w.done[l]) ndone++;
      else nbar++;
    }
    if (ndone == kLanes) return;
    if (ndone != 0 && nbar != 0) {
      fprintf(stderr, "hostwave: lanes diverged at a cross-lane operation (%d lanes returned, %d waiting)\n", ndone, nbar);
      abort();
    }
  }
}

}  // namespace hostwave

#define threadIdx (HwIdx{(unsigned)hostwave::cur_lane(), 0, 0})
#define blockIdx (hostwave::wave().block)
#define __syncthreads() hostwave::syncthreads()
#define __threadfence_block() std::atomic_thread_fence(std::memory_order_seq_cst)
#define __threadfence() std::atomic_thread_fence(std::memory_order_seq_cst)
#define __builtin_amdgcn_readlane(v, k) hostwave::readlane((v), (k))
#define __builtin_amdgcn_readfirstlane(v) hostwave::readlane((v), 0)
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) hostwave::update_dpp((old), (src), (ctrl), (rm), (bm), (bc))

inline int __clz(int x) { return x ? __builtin_clz((unsigned)x) : 32; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }

namespace hostwave {
// v_alignbyte_b32: the 4 bytes at byte offset s & 3 of the 8-byte value hi:lo
inline uint32_t alignbyte(uint32_t hi, uint32_t lo, uint32_t s) {
  return (uint32_t)(((((uint64_t)hi) << 32) | lo) >> (8 * (s & 3)));
}
// v_sad_u8: acc plus the sum of the absolute differences of the four bytes of a and b
inline uint32_t sad_u8(uint32_t a, uint32_t b, uint32_t acc) {
  for (int k = 0; k < 4; k++) {
    const int x = (int)((a >> (8 * k)) & 0xff), y = (int)((b >> (8 * k)) & 0xff);
    acc += (uint32_t)(x > y ? x - y : y - x);
  }
  return acc;
}
}  // namespace hostwave
namespace hostwave {
// ds_bpermute_b32: the value v of lane (addr / 4) mod 64, for each lane its own
inline int bpermute(int addr, int v) {
  int64_t all[kLanes];
  exchange(v, all);
  return (int)all[(addr >> 2) & (kLanes - 1)];
}
}  // namespace hostwave
#define __builtin_amdgcn_ds_bpermute(addr, v) hostwave::bpermute((addr), (v))
#define __builtin_amdgcn_alignbyte(hi, lo, s) hostwave::alignbyte((hi), (lo), (s))
#define __builtin_amdgcn_sad_u8(a, b, acc) hostwave::sad_u8((a), (b), (acc))
#define __ballot(p) hostwave::ballot((p))
inline int atomicAdd(int *p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
  return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
  return a > b ? a : b;
}

// A launch: workgroups one after another, each as 64 lane fibers.
template <class K, class... Args>
inline void hostwave_launch(K kernel, dim3 grid, dim3 block, Args... args) {
  if (block.x != hostwave::kLanes || block.y != 1 || block.z != 1) {
    fprintf(stderr, "hostwave: only 64-lane workgroups are emulated\n");
    abort();
  }
  auto call = [&]() { kernel(args...); };
  using Call = decltype(call);
  for (unsigned b = 0; b < grid.x; b++)
    hostwave::run_workgroup(b, [](void *c) { (*(Call *)c)(); }, &call);
}

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) hostwave_launch(kernel, grid, block, __VA_ARGS__)

inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
  p->multiProcessorCount = 1;
  return hipSuccess;
}
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, K, int, int) {
  *n = 1;
  return hipSuccess;
}
