// traverse.hip — weighted shortest paths (delta-stepping) and 64-wide
// bit-parallel multi-source BFS for gfx950, each as one persistent
// cooperative kernel per call; plus the weighted Gram product S^T D S.
//
// SSSP follows DeltaStep (gapbs/sssp.cc:53-140) in what it computes: exact
// int32 distances, 2^30-1 for unreached vertices.  The bucket structure is
// the two-pile (near / far) form of delta-stepping: the near queue holds
// vertices whose tentative distance is below the current bucket bound and is
// relaxed round after round; improvements beyond the bound go to the far
// pile, which is filtered into the next bucket when the near queue empties.
// Relaxation is an atomicMin on the distance, so the result is the exact
// shortest-path distance for any delta (label correcting).  Vertices enter a
// queue at most once per round thanks to per-vertex round stamps.
//
// Multi-source BFS carries up to 64 searches in one 64-bit mask per vertex
// (bit k = source k): a vertex's `fresh` mask holds the searches that first
// reached it at the current level, pushed to neighbours with atomicOr; the
// level of search k at v is written when bit k first appears.  No reference
// counterpart; it equals k single-source BFS.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <climits>
#include <cstdio>

#include "phde_common.h"
#include "traverse.h"

namespace cg = cooperative_groups;

namespace phde {

constexpr int kTBlock = 256;
constexpr int32_t kDistInf = INT32_MAX / 2;  // sssp.cc:50

__device__ __forceinline__ unsigned long long lanemask_lt_t() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Wave-aggregated append of flagged items to q (count at *cnt).
__device__ __forceinline__ void wave_push(bool f, uint32_t v, uint32_t *q, unsigned int *cnt) {
  const unsigned long long mask = __ballot(f);
  if (!mask) return;
  const int leader = __ffsll((long long)mask) - 1;
  unsigned int base = 0;
  if (__lane_id() == leader) base = atomicAdd(cnt, (unsigned int)__popcll(mask));
  base = __shfl(base, leader);
  if (f) q[base + __popcll(mask & lanemask_lt_t())] = v;
}

__device__ __forceinline__ unsigned int aload(unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct SsspArgs {
  DevGraph g;
  int32_t *dist;
  uint32_t *near0, *near1, *far0, *far1;
  int32_t *nstamp, *fstamp;  // last round / epoch a vertex was queued
  unsigned int *cnt;         // [0..2] near counts (rotating), [3..5] far counts, [6] min scratch
  int64_t src;
  int32_t delta;
  int32_t *min_dist;
  KcResult *kc;
};

__global__ __launch_bounds__(kTBlock) void sssp_persistent_kernel(SsspArgs a) {
  cg::grid_group grid = cg::this_grid();
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave, nwaves = gsize / kWave;
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;
  const int32_t *wt = a.g.w;

  for (int64_t i = gtid; i < n; i += gsize) {
    a.dist[i] = (i == a.src) ? 0 : kDistInf;
    a.nstamp[i] = -1;
    a.fstamp[i] = -1;
  }
  if (gtid == 0) {
    for (int k = 0; k < 8; k++) a.cnt[k] = 0;
    a.cnt[0] = 1;
    a.near0[0] = (uint32_t)a.src;
    a.kc->key = 0;
    a.kc->sumsq = 0;
    a.kc->unreached = 0;
    a.kc->depth = 0;
  }
  grid.sync();

  long long bound = a.delta;  // current bucket: distances < bound
  int round = 0, epoch = 0;
  int nsel = 0, fsel = 0;     // which near / far buffer is current
  int steps = 0;
  for (;;) {
    unsigned int *ncur_cnt = &a.cnt[round % 3];
    unsigned int *nnext_cnt = &a.cnt[(round + 1) % 3];
    const unsigned int nq = aload(ncur_cnt);
    const unsigned int nf = aload(&a.cnt[3 + (epoch % 2)]);
    if (nq == 0 && nf == 0) break;
    const uint32_t *ncur = nsel ? a.near1 : a.near0;
    uint32_t *nnext = nsel ? a.near0 : a.near1;
    uint32_t *fcur = fsel ? a.far1 : a.far0;
    if (gtid == 0) a.cnt[(round + 2) % 3] = 0;
    if (nq > 0) {
      // ---- relax the near queue (sssp.cc:74-108)
      for (int64_t base = gwave * kWave; base < (int64_t)nq; base += nwaves * kWave) {
        const int64_t idx = base + lane;
        const bool act = idx < (int64_t)nq;
        const uint32_t u = act ? ncur[idx] : 0u;
        const uint32_t b = act ? rowptr[u] : 0u, e = act ? rowptr[u + 1] : 0u;
        const int32_t du = act ? a.dist[u] : 0;
        int d = (int)(e - b);
        // wave-cooperative expansion of long lists
        const bool big = d >= 32;
        unsigned long long bmask = __ballot(big);
        while (bmask) {
          const int l = __ffsll((long long)bmask) - 1;
          bmask &= bmask - 1;
          const uint32_t bb = __shfl(b, l), be = __shfl(e, l);
          const int32_t bd = __shfl(du, l);
          for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
            const uint32_t k = k0 + lane;
            bool pn = false, pf = false;
            uint32_t v = 0;
            if (k < be) {
              v = adj[k];
              const int32_t nd = bd + wt[k];
              if (nd < a.dist[v]) {
                const int32_t old = atomicMin(&a.dist[v], nd);
                if (nd < old) {
                  if (nd < bound)
                    pn = atomicExch(&a.nstamp[v], round + 1) != round + 1;
                  else
                    pf = atomicExch(&a.fstamp[v], epoch) != epoch;
                }
              }
            }
            wave_push(pn, v, nnext, nnext_cnt);
            wave_push(pf, v, fcur, &a.cnt[3 + (epoch % 2)]);
          }
        }
        const int ds = big ? 0 : d;
        int maxd = ds;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
        for (int t = 0; t < maxd; t++) {
          bool pn = false, pf = false;
          uint32_t v = 0;
          if (t < ds) {
            v = adj[b + t];
            const int32_t nd = du + wt[b + t];
            if (nd < a.dist[v]) {
              const int32_t old = atomicMin(&a.dist[v], nd);
              if (nd < old) {
                if (nd < bound)
                  pn = atomicExch(&a.nstamp[v], round + 1) != round + 1;
                else
                  pf = atomicExch(&a.fstamp[v], epoch) != epoch;
              }
            }
          }
          wave_push(pn, v, nnext, nnext_cnt);
          wave_push(pf, v, fcur, &a.cnt[3 + (epoch % 2)]);
        }
      }
      nsel ^= 1;
      round++;
    } else {
      // ---- near queue empty: advance the bucket bound past the smallest
      // far distance and split the far pile (sssp.cc:109-131).
      // pass 1: minimum tentative distance in the far pile
      if (gtid == 0) a.cnt[6] = 0x7fffffff;
      grid.sync();
      int32_t mn = 0x7fffffff;
      for (int64_t i = gtid; i < (int64_t)nf; i += gsize) mn = min(mn, a.dist[fcur[i]]);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mn = min(mn, __shfl_xor(mn, o));
      if (lane == 0 && mn != 0x7fffffff) atomicMin((int *)&a.cnt[6], mn);
      grid.sync();
      const int32_t fmin = (int32_t)aload(&a.cnt[6]);
      bound = ((long long)fmin / a.delta + 1) * (long long)a.delta;
      // pass 2: split the pile into the near queue and the next far pile
      uint32_t *fnext = fsel ? a.far0 : a.far1;
      unsigned int *fnext_cnt = &a.cnt[3 + ((epoch + 1) % 2)];
      for (int64_t base = gwave * kWave; base < (int64_t)nf; base += nwaves * kWave) {
        const int64_t idx = base + lane;
        bool pn = false, pf = false;
        uint32_t v = 0;
        if (idx < (int64_t)nf) {
          v = fcur[idx];
          if (a.dist[v] < bound)
            pn = atomicExch(&a.nstamp[v], round + 1) != round + 1;
          else
            pf = atomicExch(&a.fstamp[v], epoch + 1) != epoch + 1;
        }
        wave_push(pn, v, nnext, nnext_cnt);
        wave_push(pf, v, fnext, fnext_cnt);
      }
      if (gtid == 0) a.cnt[3 + (epoch % 2)] = 0;
      fsel ^= 1;
      epoch++;
      nsel ^= 1;
      round++;
    }
    steps++;
    grid.sync();
  }

  // ---- fused k-centers epilogue (as in the BFS kernel)
  unsigned long long key = 0, sumsq = 0, unreached = 0;
  for (int64_t i = gtid; i < n; i += gsize) {
    const int32_t d = a.dist[i];
    if (d >= kDistInf)
      unreached++;
    else
      sumsq += (unsigned long long)d * (unsigned long long)d;
    if (a.min_dist) {
      int32_t md = a.min_dist[i];
      if (d < md) {
        md = d;
        a.min_dist[i] = md;
      }
      const unsigned long long kk =
          ((unsigned long long)(uint32_t)md << 32) | (0xffffffffull - (unsigned long long)i);
      key = kk > key ? kk : key;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ok = __shfl_xor(key, o);
    key = ok > key ? ok : key;
    sumsq += __shfl_xor(sumsq, o);
    unreached += __shfl_xor(unreached, o);
  }
  if (lane == 0) {
    if (a.min_dist) atomicMax(&a.kc->key, key);
    atomicAdd(&a.kc->sumsq, sumsq);
    if (unreached) atomicAdd(&a.kc->unreached, unreached);
  }
  if (gtid == 0) a.kc->depth = steps;
}

// ------------------------------------------------------------------ MS-BFS

struct MsArgs {
  DevGraph g;
  int k;
  unsigned long long *visit, *fr0, *fr1;
  uint32_t *q0, *q1;
  unsigned int *cnt;  // rotating [3]
  const int64_t *src;
  int32_t *D;  // n*k column-major
};

__global__ __launch_bounds__(kTBlock) void msbfs_persistent_kernel(MsArgs a) {
  cg::grid_group grid = cg::this_grid();
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave, nwaves = gsize / kWave;
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;

  for (int64_t i = gtid; i < n; i += gsize) {
    a.visit[i] = 0;
    a.fr0[i] = 0;
    a.fr1[i] = 0;
  }
  for (int64_t e = gtid; e < n * a.k; e += gsize) a.D[e] = -1;
  if (gtid == 0) a.cnt[0] = a.cnt[1] = a.cnt[2] = 0;
  grid.sync();
  if (gtid < a.k) {
    const int64_t s = a.src[gtid];
    atomicOr(&a.visit[s], 1ull << gtid);
    const unsigned long long old = atomicOr(&a.fr0[s], 1ull << gtid);
    a.D[(int64_t)gtid * n + s] = 0;
    if (old == 0) a.q0[atomicAdd(&a.cnt[0], 1u)] = (uint32_t)s;
  }
  grid.sync();
  for (int L = 0;; L++) {
    unsigned int *ccur = &a.cnt[L % 3], *cnext = &a.cnt[(L + 1) % 3];
    const unsigned int nq = aload(ccur);
    if (nq == 0) break;
    if (gtid == 0) a.cnt[(L + 2) % 3] = 0;
    const uint32_t *qc = (L & 1) ? a.q1 : a.q0;
    uint32_t *qn = (L & 1) ? a.q0 : a.q1;
    unsigned long long *fc = (L & 1) ? a.fr1 : a.fr0;
    unsigned long long *fn = (L & 1) ? a.fr0 : a.fr1;
    // expand: push fresh bits of every queued vertex to its neighbours
    for (int64_t base = gwave * kWave; base < (int64_t)nq; base += nwaves * kWave) {
      const int64_t idx = base + lane;
      const bool act = idx < (int64_t)nq;
      const uint32_t u = act ? qc[idx] : 0u;
      const unsigned long long bits = act ? fc[u] : 0ull;
      const uint32_t b = act ? rowptr[u] : 0u, e = act ? rowptr[u + 1] : 0u;
      const int d = (int)(e - b);
      const bool big = d >= 32;
      unsigned long long bmask = __ballot(big);
      while (bmask) {
        const int l = __ffsll((long long)bmask) - 1;
        bmask &= bmask - 1;
        const uint32_t bb = __shfl(b, l), be = __shfl(e, l);
        const unsigned long long bbits = __shfl(bits, l);
        for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
          const uint32_t k = k0 + lane;
          bool push = false;
          uint32_t v = 0;
          if (k < be) {
            v = adj[k];
            const unsigned long long nb = bbits & ~a.visit[v];
            if (nb) {
              const unsigned long long old = atomicOr(&a.visit[v], nb);
              const unsigned long long nw = nb & ~old;
              if (nw) {
                const unsigned long long pf = atomicOr(&fn[v], nw);
                push = pf == 0;
                unsigned long long t = nw;
                while (t) {
                  const int s = __ffsll((long long)t) - 1;
                  t &= t - 1;
                  a.D[(int64_t)s * n + v] = L + 1;
                }
              }
            }
          }
          wave_push(push, v, qn, cnext);
        }
      }
      const int ds = big ? 0 : d;
      int maxd = ds;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
      for (int t = 0; t < maxd; t++) {
        bool push = false;
        uint32_t v = 0;
        if (t < ds) {
          v = adj[b + t];
          const unsigned long long nb = bits & ~a.visit[v];
          if (nb) {
            const unsigned long long old = atomicOr(&a.visit[v], nb);
            const unsigned long long nw = nb & ~old;
            if (nw) {
              const unsigned long long pf = atomicOr(&fn[v], nw);
              push = pf == 0;
              unsigned long long tt = nw;
              while (tt) {
                const int s = __ffsll((long long)tt) - 1;
                tt &= tt - 1;
                a.D[(int64_t)s * n + v] = L + 1;
              }
            }
          }
        }
        wave_push(push, v, qn, cnext);
      }
    }
    grid.sync();
    // retire this level's fresh masks (they were read above)
    for (int64_t i = gtid; i < (int64_t)nq; i += gsize) fc[qc[i]] = 0;
    grid.sync();
  }
}

// ------------------------------------------------------------------ host side

static int coop_grid(phde_ctx *c, const void *fn, int want_per_cu, int *out) {
  int per_cu = 0;
  PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kTBlock, 0));
  if (per_cu < 1) {
    set_error("traversal kernel cannot be resident");
    return PHDE_ERR_DEVICE;
  }
  *out = c->num_cus * std::min(per_cu, want_per_cu);
  return PHDE_OK;
}

int sssp_run(phde_ctx *c, int64_t src, int32_t delta, int32_t *dist_out, int32_t *min_dist,
             bool sync_result) {
  if (!c->g.w) {
    set_error("SSSP needs edge weights");
    return PHDE_ERR_WEIGHTS;
  }
  int rc = bfs_setup(c);
  if (rc) return rc;
  const int64_t n = c->g.n;
  if (!c->ts_far[0]) {
    for (int i = 0; i < 2; i++) PHDE_HIP(hipMalloc(&c->ts_far[i], sizeof(uint32_t) * (n + 1)));
    PHDE_HIP(hipMalloc(&c->ts_nstamp, sizeof(int32_t) * n));
    PHDE_HIP(hipMalloc(&c->ts_fstamp, sizeof(int32_t) * n));
    PHDE_HIP(hipMalloc(&c->ts_cnt, sizeof(unsigned int) * 8));
    PHDE_TRY_RC(coop_grid(c, (const void *)sssp_persistent_kernel, 4, &c->sssp_grid));
  }
  if (delta <= 0) delta = c->default_delta;
  SsspArgs a;
  a.g = c->g;
  a.dist = dist_out;
  a.near0 = c->d_q[0];
  a.near1 = c->d_q[1];
  a.far0 = c->ts_far[0];
  a.far1 = c->ts_far[1];
  a.nstamp = c->ts_nstamp;
  a.fstamp = c->ts_fstamp;
  a.cnt = c->ts_cnt;
  a.src = src;
  a.delta = delta;
  a.min_dist = min_dist;
  a.kc = c->d_kc;
  void *args[] = {&a};
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_SSSP, &ev);
  PHDE_HIP(hipLaunchCooperativeKernel((const void *)sssp_persistent_kernel, dim3(c->sssp_grid),
                                      dim3(kTBlock), args, 0, c->stream));
  c->timer.end(c->stream, PHDE_KERNEL_SSSP, ev);
  if (sync_result) {
    PHDE_HIP(hipMemcpyAsync(c->h_kc, c->d_kc, sizeof(KcResult), hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  }
  return PHDE_OK;
}

int msbfs_run(phde_ctx *c, const int64_t *sources, int k, int32_t *D) {
  int rc = bfs_setup(c);
  if (rc) return rc;
  const int64_t n = c->g.n;
  if (!c->ms_visit) {
    PHDE_HIP(hipMalloc(&c->ms_visit, sizeof(unsigned long long) * n));
    PHDE_HIP(hipMalloc(&c->ms_fr[0], sizeof(unsigned long long) * n));
    PHDE_HIP(hipMalloc(&c->ms_fr[1], sizeof(unsigned long long) * n));
    PHDE_HIP(hipMalloc(&c->ms_src, sizeof(int64_t) * 64));
    if (!c->ts_cnt) PHDE_HIP(hipMalloc(&c->ts_cnt, sizeof(unsigned int) * 8));
    PHDE_TRY_RC(coop_grid(c, (const void *)msbfs_persistent_kernel, 4, &c->ms_grid));
  }
  PHDE_HIP(hipMemcpyAsync(c->ms_src, sources, sizeof(int64_t) * k, hipMemcpyHostToDevice, c->stream));
  MsArgs a;
  a.g = c->g;
  a.k = k;
  a.visit = c->ms_visit;
  a.fr0 = c->ms_fr[0];
  a.fr1 = c->ms_fr[1];
  a.q0 = c->d_q[0];
  a.q1 = c->d_q[1];
  a.cnt = c->ts_cnt;
  a.src = c->ms_src;
  a.D = D;
  void *args[] = {&a};
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_MSBFS, &ev);
  PHDE_HIP(hipLaunchCooperativeKernel((const void *)msbfs_persistent_kernel, dim3(c->ms_grid),
                                      dim3(kTBlock), args, 0, c->stream));
  c->timer.end(c->stream, PHDE_KERNEL_MSBFS, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

void traverse_release(phde_ctx *c) {
  void *p[] = {c->ts_far[0], c->ts_far[1], c->ts_nstamp, c->ts_fstamp, c->ts_cnt,
               c->ms_visit, c->ms_fr[0], c->ms_fr[1], c->ms_src};
  for (void *x : p)
    if (x) (void)hipFree(x);
}

}  // namespace phde
