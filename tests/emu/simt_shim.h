// SIMT emulation shim: the device code of kraken2_amd/csrc/k2g_kernels.hip
// compiled for the host.  A workgroup is 64 std::threads; ballots, shuffles
// and __syncthreads are barrier-synchronised collectives, so a collective
// reached by only part of the wave deadlocks here (and is reported by the
// watchdog) instead of silently reading stale lanes as a GPU would.
#pragma once
#include <atomic>
#include <barrier>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <thread>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__ static

namespace emu {
struct Dim3 {
  unsigned x = 1, y = 1, z = 1;
};
extern thread_local Dim3 threadIdx_, blockIdx_;
extern Dim3 blockDim_, gridDim_;
extern std::barrier<> *bar;
extern unsigned long long vote_slot[64];
extern uint64_t shfl_slot[64];
extern std::atomic<unsigned long long> collectives;
inline void wait() {
  collectives.fetch_add(1, std::memory_order_relaxed);
  bar->arrive_and_wait();
}
}  // namespace emu

#define threadIdx emu::threadIdx_
#define blockIdx emu::blockIdx_
#define blockDim emu::blockDim_
#define gridDim emu::gridDim_

inline void __syncthreads() { emu::wait(); }
inline void __builtin_amdgcn_wave_barrier() { emu::wait(); }
inline void __threadfence() { __atomic_thread_fence(__ATOMIC_SEQ_CST); }

inline unsigned long long __ballot(int pred) {
  const unsigned lane = threadIdx.x;
  emu::vote_slot[lane] = pred ? 1ULL : 0ULL;
  emu::wait();
  unsigned long long m = 0;
  for (int i = 0; i < 64; i++) m |= emu::vote_slot[i] << i;
  emu::wait();
  return m;
}

template <class T>
inline T __shfl(T v, int src) {
  const unsigned lane = threadIdx.x;
  emu::shfl_slot[lane] = (uint64_t) v;
  emu::wait();
  T r = (T) emu::shfl_slot[src & 63];
  emu::wait();
  return r;
}

template <class T>
inline T __shfl_xor(T v, int mask) {
  return __shfl(v, (int) (threadIdx.x ^ (unsigned) mask));
}

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long) x) : 64; }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline uint64_t __umul64hi(uint64_t a, uint64_t b) {
  return (uint64_t) (((unsigned __int128) a * b) >> 64);
}

template <class T, class U>
inline T atomicAdd(T *p, U v) { return __atomic_fetch_add(p, (T) v, __ATOMIC_SEQ_CST); }
template <class T, class U>
inline T atomicOr(T *p, U v) { return __atomic_fetch_or(p, (T) v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicCAS(T *p, T expected, T desired) {
  __atomic_compare_exchange_n(p, &expected, desired, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return expected;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < (T) v && !__atomic_compare_exchange_n(p, &cur, (T) v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
#define __ATOMIC_RELAXED_ __ATOMIC_RELAXED
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_SEQ_CST)

inline int __builtin_amdgcn_readfirstlane(int v) { return __shfl(v, 0); }
inline int __builtin_amdgcn_readlane(int v, int lane) { return __shfl(v, lane); }
inline bool __builtin_amdgcn_inverse_ballot_w64(unsigned long long m) { return (m >> threadIdx.x) & 1ULL; }
// Only the wave shift right by one lane (dpp_ctrl 0x138) is emulated: lane i
// gets lane i-1's src, lane 0 keeps `old`.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int, int, bool) {
  if (ctrl != 0x138) {
    std::fprintf(stderr, "emu: unsupported dpp_ctrl %#x\n", ctrl);
    std::abort();
  }
  const int up = __shfl(src, (int) threadIdx.x - 1);
  return threadIdx.x == 0 ? old : up;
}
