// levels.h — host-side driver for level-synchronous traversals run as one
// kernel launch per level.
//
// Measured on MI355X: a back-to-back kernel launch costs ~2.6 us of GPU
// time, a cooperative grid.sync() 33-78 us and a hand-rolled atomic grid
// barrier 4-13 us; kernel boundaries also give cross-XCD coherence without
// explicit L2 write-back / invalidate.  So traversals enqueue level kernels
// ahead of the GPU; each level kernel publishes its level number (progress)
// and, once the frontier is empty, the final depth (done) into mapped host
// memory, which the host polls without touching the stream.  At most
// `lead` kernels past the end are enqueued; they return at once.
#pragma once
#include <chrono>
#include <thread>

#include "phde_common.h"

namespace phde {

// flags layout in mapped host memory (ctx->h_flags / ctx->d_flags)
enum { kFlagDone = 0, kFlagProgress = 1 };

// Mapped host flags of a context (allocated on first use).
int flags_setup(phde_ctx *c);

// Enqueue level kernels launch(L) for L = 0, 1, ... until the device reports
// done.  Returns PHDE_OK, or PHDE_ERR_DEVICE when the stream reports an error
// or no progress is seen for `stall_s` seconds.
template <class Launch>
int run_levels(phde_ctx *c, Launch launch, int lead = 48, int chunk = 16, double stall_s = 60.0) {
  volatile int *h = c->h_flags;
  int enqueued = 0;
  int last_prog = -1;
  auto t_prog = std::chrono::steady_clock::now();
  for (;;) {
    const int done = h[kFlagDone];
    if (done >= 0) return PHDE_OK;
    const int prog = h[kFlagProgress];
    if (prog != last_prog) {
      last_prog = prog;
      t_prog = std::chrono::steady_clock::now();
    }
    if (enqueued - prog <= lead) {
      for (int i = 0; i < chunk; i++) {
        int rc = launch(enqueued);
        if (rc) return rc;
        enqueued++;
      }
      continue;
    }
    const hipError_t q = hipStreamQuery(c->stream);
    if (q != hipSuccess && q != hipErrorNotReady) {
      set_error(std::string("traversal stream: ") + hipGetErrorString(q));
      return PHDE_ERR_DEVICE;
    }
    if (q == hipSuccess && h[kFlagDone] < 0) {
      // everything enqueued has finished without reaching the end: keep going
      continue;
    }
    const double stalled =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_prog).count();
    if (stalled > stall_s) {
      set_error("traversal made no progress");
      return PHDE_ERR_DEVICE;
    }
    std::this_thread::yield();
  }
}

// Device side.  `done_dev` is a device-memory copy of the final depth (-1
// while running) that later level kernels read cheaply; `host_flags` is the
// mapped host copy the driver polls.  A level kernel returns at once when
// the traversal finished before its level; the kernel of level depth + 1
// (the first to see an empty frontier) runs the epilogue.
__device__ __forceinline__ bool level_finished(const int *done_dev, int L) {
  const int done = *done_dev;
  return done >= 0 && done + 1 < L;
}

__device__ __forceinline__ void level_publish(int *host_flags, int L) {
  __hip_atomic_store(&host_flags[kFlagProgress], L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void level_done(int *done_dev, int *host_flags, int depth) {
  *done_dev = depth;
  __hip_atomic_store(&host_flags[kFlagDone], depth, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace phde
