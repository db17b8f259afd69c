// device_util.h — device helpers shared by the traversal and pivot-selection
// kernels: wave reductions, device-scope (cross-XCD) loads and stores, and
// the bounded grid barrier of the persistent kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace phde {

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Exact sum of squared distances as a 128-bit integer (lo, hi words).  A
// weighted distance reaches 2^30 - 2, its square 2^60, so a few dozen
// squares already pass 2^64; path-like unweighted graphs pass it above
// n = 3.8M.  Each word is added with its carry, so the total is exact at
// any size; the host rounds it to a double once (u128_to_double).
struct U128 {
  unsigned long long lo, hi;
};
__device__ __forceinline__ void u128_add(U128 &a, unsigned long long v) {
  a.lo += v;
  a.hi += a.lo < v ? 1ull : 0ull;
}
__device__ __forceinline__ U128 wave_sum_u128(U128 a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long lo = __shfl_xor(a.lo, o), hi = __shfl_xor(a.hi, o);
    a.hi += hi;
    u128_add(a, lo);
  }
  return a;
}
// acc[0] += v.lo, acc[1] += v.hi + carry: the carry out of the low word is
// known from the value the atomic returns.
__device__ __forceinline__ void atomic_add_u128(unsigned long long *acc, U128 v) {
  if (v.lo) {
    const unsigned long long old = atomicAdd(&acc[0], v.lo);
    if (old + v.lo < old) v.hi++;
  }
  if (v.hi) atomicAdd(&acc[1], v.hi);
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Data written by one block and read by another within the persistent
// kernel go through device-scope (agent) relaxed atomic loads and stores:
// those are coherent across the XCDs' L2 caches without the L2 write-back /
// invalidate that an agent-scope fence costs, so the read-only graph stays
// cached from level to level.
__device__ __forceinline__ int32_t ld_dev(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_dev(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_dev(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned int *p, unsigned int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Grid barrier for co-resident blocks (a persistent launch): every block
// adds one to a monotone arrival counter and waits until it reaches
// `target` (= blocks x barriers so far).  All of this thread's memory
// operations complete before the arrival (s_waitcnt), and every
// cross-block datum is accessed with device-scope operations, so no cache
// maintenance is needed.  The wait is bounded: on timeout the block flags
// an error and continues, so the kernel always ends.  Returns true in the
// blocks whose own wait timed out (the others then time out at their next
// barrier), so callers can stop without loading the flag after every
// barrier.
__device__ __forceinline__ bool grid_barrier(unsigned int *arrive, unsigned int target,
                                             unsigned int *err, int backoff) {
  __shared__ int timed_out;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long spins = 0;
    int to = 0;
    while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      // polling every block on one word congests its memory channel and
      // delays unrelated loads: back off between polls
      for (int k = 0; k < backoff; k++) __builtin_amdgcn_s_sleep(2);
      if (++spins > (1L << 25)) {
        atomicOr(err, 1u);
        to = 1;
        break;
      }
    }
    timed_out = to;
  }
  __syncthreads();
  return timed_out != 0;
}

// Exclusive prefix sum over the wave; *total gets the sum of all lanes.
__device__ __forceinline__ int wave_excl_scan(int x, int *total) {
  const int lane = __lane_id();
  int incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  *total = __shfl(incl, 63);
  return incl - x;
}

}  // namespace phde
