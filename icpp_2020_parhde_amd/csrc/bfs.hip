// bfs.hip — direction-optimising breadth-first search for gfx950, one
// persistent cooperative kernel per traversal, with the k-centers update
// fused into its epilogue.
//
// Reference behaviour followed: DOBFS (gapbs/bfs.cc:144-185).  Levels are
// exact hop counts and do not depend on scheduling; the top-down /
// bottom-up switch uses the reference's alpha/beta rule (bfs.cc:161,
// 172-173), which only affects speed.  The epilogue restates the k-centers
// step of ParHDE.cpp:124-140 (min over columns, first maximiser).
//
// Design (MI355X): all blocks stay resident (cooperative launch) and meet
// at a grid barrier once per level, so a 2000-level traversal costs 2000
// barriers instead of 2000 kernel launches plus host round trips.  The
// frontier is a vertex queue in HBM for top-down levels; bottom-up levels
// test frontier membership in a bitmap that the previous bottom-up level
// wrote with one 64-bit ballot store per wave (no atomics, no clearing).
// The first bottom-up level after a top-down level tests `level == L`
// directly on the distance array, so top-down levels never touch a
// bitmap.  Queue appends are wave-aggregated (one atomic per wave).
// Vertices of degree >= kHugeDegree are expanded by the whole grid, those
// of degree >= kBigDegree by one wave, the rest one lane each.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cstdio>

#include "phde_common.h"

namespace cg = cooperative_groups;

namespace phde {

constexpr int kBfsBlock = 256;
constexpr int kHugeDegree = 8192;
constexpr int kBigDegree = 32;

__device__ __forceinline__ unsigned long long lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__device__ __forceinline__ unsigned int atomic_read_u32(unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atomic_read_u64(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Append the lanes' `found` vertices to the next frontier (queue, huge list,
// degree sum).  Must be called by all 64 lanes of the wave.
__device__ __forceinline__ void wave_append(bool found, uint32_t v, const int32_t *deg32,
                                            uint32_t *q_next, uint32_t *h_next, BfsSlot *slot) {
  const unsigned long long mask = __ballot(found);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  const int dv = found ? deg32[v] : 0;
  const bool huge = found && dv >= kHugeDegree;
  const unsigned long long hmask = __ballot(huge);
  const unsigned long long dsum = wave_sum_u64((unsigned long long)dv);
  unsigned int base = 0, hbase = 0;
  if (lane == leader) {
    base = atomicAdd(&slot->q, (unsigned int)__popcll(mask));
    atomicAdd(&slot->scout, dsum);
    if (hmask) hbase = atomicAdd(&slot->huge, (unsigned int)__popcll(hmask));
  }
  base = __shfl(base, leader);
  hbase = __shfl(hbase, leader);
  const unsigned long long lt = lanemask_lt();
  if (found) q_next[base + __popcll(mask & lt)] = v;
  if (huge) h_next[hbase + __popcll(hmask & lt)] = v;
}

__device__ __forceinline__ bool discover(int32_t *dist, uint32_t v, int32_t next_level) {
  if (dist[v] != kUnvisited) return false;
  return atomicCAS(&dist[v], kUnvisited, next_level) == kUnvisited;
}

struct BfsArgs {
  DevGraph g;
  int32_t *dist;
  uint32_t *q0, *q1, *h0, *h1;
  uint64_t *bm0, *bm1;
  BfsSlot *slots;  // 3
  int64_t src;
  int alpha, beta;
  int32_t *min_dist;  // nullable: fused k-centers update
  KcResult *kc;
};

__global__ __launch_bounds__(kBfsBlock) void bfs_persistent_kernel(BfsArgs a) {
  cg::grid_group grid = cg::this_grid();
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave;
  const int64_t nwaves = gsize / kWave;
  const uint32_t *rowptr = a.g.rowptr;
  const uint32_t *adj = a.g.adj;
  const int32_t *deg32 = a.g.deg32;

  // ---- initialisation: every vertex unvisited, frontier = {src}
  for (int64_t i = gtid; i < n; i += gsize) a.dist[i] = (i == a.src) ? 0 : kUnvisited;
  if (gtid == 0) {
    const int dsrc = deg32[a.src];
    a.slots[0].q = 1;
    a.slots[0].scout = (unsigned long long)dsrc;
    a.slots[0].huge = dsrc >= kHugeDegree ? 1u : 0u;
    a.q0[0] = (uint32_t)a.src;
    a.h0[0] = (uint32_t)a.src;
    for (int s = 1; s < 3; s++) {
      a.slots[s].q = 0;
      a.slots[s].huge = 0;
      a.slots[s].scout = 0;
    }
    a.kc->key = 0;
    a.kc->sumsq = 0;
    a.kc->unreached = 0;
    a.kc->depth = 0;
  }
  grid.sync();

  // Direction state machine, evaluated identically by every thread from the
  // same global counters (bfs.cc:159-183).
  bool in_bu = false, prev_bu = false, after_bu = false;
  long long edges_to_check = a.g.m;
  long long awake_old = 0;
  int L = 0;
  for (;; L++) {
    BfsSlot *cur = &a.slots[L % 3];
    BfsSlot *nxt = &a.slots[(L + 1) % 3];
    BfsSlot *old = &a.slots[(L + 2) % 3];
    const unsigned int nq = atomic_read_u32(&cur->q);
    if (nq == 0) break;
    const unsigned int nh = atomic_read_u32(&cur->huge);
    const unsigned long long sc = atomic_read_u64(&cur->scout);
    bool bu;
    if (in_bu) {
      const long long awake = nq;
      if (awake >= awake_old || awake > n / (long long)a.beta) {
        awake_old = awake;
        bu = true;
      } else {
        // leave the bottom-up phase with scout_count = 1 (bfs.cc:174-175)
        in_bu = false;
        after_bu = true;
        bu = false;
      }
    } else {
      bu = false;
    }
    if (!in_bu) {
      const long long scout = after_bu ? 1 : (long long)sc;
      after_bu = false;
      if (scout > edges_to_check / a.alpha) {
        in_bu = true;
        awake_old = nq;
        bu = true;
      } else {
        edges_to_check -= scout;
        bu = false;
      }
    }
    if (gtid == 0) {
      old->q = 0;
      old->huge = 0;
      old->scout = 0;
    }
    const uint32_t *q_cur = (L & 1) ? a.q1 : a.q0;
    uint32_t *q_next = (L & 1) ? a.q0 : a.q1;
    const uint32_t *h_cur = (L & 1) ? a.h1 : a.h0;
    uint32_t *h_next = (L & 1) ? a.h0 : a.h1;
    const uint64_t *bm_cur = (L & 1) ? a.bm1 : a.bm0;
    uint64_t *bm_next = (L & 1) ? a.bm0 : a.bm1;
    const int32_t next_level = L + 1;

    if (!bu) {
      // ---------------- top-down (bfs.cc:76-106)
      // grid-wide expansion of huge-degree frontier vertices
      for (unsigned int hi = 0; hi < nh; hi++) {
        const uint32_t u = h_cur[hi];
        const uint32_t b = rowptr[u], e = rowptr[u + 1];
        for (int64_t k0 = (int64_t)b + gwave * kWave; k0 < (int64_t)e; k0 += nwaves * kWave) {
          const int64_t k = k0 + lane;
          const bool ok = k < (int64_t)e;
          const uint32_t v = ok ? adj[k] : 0u;
          const bool f = ok && discover(a.dist, v, next_level);
          wave_append(f, v, deg32, q_next, h_next, nxt);
        }
      }
      for (int64_t base = gwave * kWave; base < (int64_t)nq; base += nwaves * kWave) {
        const int64_t idx = base + lane;
        const bool act = idx < (int64_t)nq;
        const uint32_t u = act ? q_cur[idx] : 0u;
        uint32_t b = act ? rowptr[u] : 0u;
        uint32_t e = act ? rowptr[u + 1] : 0u;
        int d = (int)(e - b);
        if (d >= kHugeDegree) d = 0;  // expanded above
        const bool big = d >= kBigDegree;
        unsigned long long bmask = __ballot(big);
        while (bmask) {
          const int l = __ffsll((long long)bmask) - 1;
          bmask &= bmask - 1;
          const uint32_t bb = __shfl(b, l), be = __shfl(e, l);
          for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
            const uint32_t k = k0 + lane;
            const bool ok = k < be;
            const uint32_t v = ok ? adj[k] : 0u;
            const bool f = ok && discover(a.dist, v, next_level);
            wave_append(f, v, deg32, q_next, h_next, nxt);
          }
        }
        const int ds = big ? 0 : d;
        const int maxd = wave_max_i32(ds);
        for (int t = 0; t < maxd; t++) {
          const bool ok = t < ds;
          const uint32_t v = ok ? adj[b + t] : 0u;
          const bool f = ok && discover(a.dist, v, next_level);
          wave_append(f, v, deg32, q_next, h_next, nxt);
        }
      }
    } else {
      // ---------------- bottom-up (bfs.cc:48-73)
      const int64_t nwords = (n + 63) / 64;
      for (int64_t w = gwave; w < nwords; w += nwaves) {
        const int64_t u = w * 64 + lane;
        bool f = false;
        if (u < n && a.dist[u] == kUnvisited) {
          const uint32_t b = rowptr[u], e = rowptr[u + 1];
          for (uint32_t k = b; k < e; k++) {
            const uint32_t v = adj[k];
            const bool in_front =
                prev_bu ? ((bm_cur[v >> 6] >> (v & 63)) & 1ull) != 0 : a.dist[v] == L;
            if (in_front) {
              f = true;
              break;
            }
          }
          if (f) a.dist[u] = next_level;
        }
        const unsigned long long mask = __ballot(f);
        if (lane == 0) bm_next[w] = mask;
        wave_append(f, (uint32_t)u, deg32, q_next, h_next, nxt);
      }
    }
    prev_bu = bu;
    grid.sync();
  }

  // ---- fused epilogue: k-centers update + argmax, sum of squares, unreached
  unsigned long long key = 0, sumsq = 0, unreached = 0;
  for (int64_t i = gtid; i < n; i += gsize) {
    const int32_t d = a.dist[i];
    if (d == kUnvisited) {
      unreached++;
    } else {
      sumsq += (unsigned long long)d * (unsigned long long)d;
    }
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
  __shared__ unsigned long long s_key[kBfsBlock / kWave], s_sq[kBfsBlock / kWave],
      s_un[kBfsBlock / kWave];
  const int wid = threadIdx.x / kWave;
  if (lane == 0) {
    s_key[wid] = key;
    s_sq[wid] = sumsq;
    s_un[wid] = unreached;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBfsBlock / kWave; w++) {
      key = s_key[w] > key ? s_key[w] : key;
      sumsq += s_sq[w];
      unreached += s_un[w];
    }
    if (a.min_dist) atomicMax(&a.kc->key, key);
    atomicAdd(&a.kc->sumsq, sumsq);
    if (unreached) atomicAdd(&a.kc->unreached, unreached);
    if (blockIdx.x == 0) a.kc->depth = L - 1;
  }
}

__global__ void dist_finalize_kernel(const int32_t *in, int32_t *out, int64_t n, int32_t from,
                                     int32_t to) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = in[i];
    out[i] = d == from ? to : d;
  }
}

int bfs_setup(phde_ctx *c) {
  if (c->bfs_grid > 0) return PHDE_OK;
  int per_cu = 0;
  PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bfs_persistent_kernel,
                                                         kBfsBlock, 0));
  if (per_cu < 1) {
    set_error("bfs kernel cannot be resident");
    return PHDE_ERR_DEVICE;
  }
  int want = 4;
  if (const char *e = getenv("PHDE_BFS_BLOCKS_PER_CU")) want = std::max(1, atoi(e));
  c->bfs_grid = c->num_cus * std::min(per_cu, want);
  const int64_t n = c->g.n;
  const int64_t nwords = (n + 63) / 64;
  const int64_t hcap = c->g.m / kHugeDegree + 2;
  PHDE_HIP(hipMalloc(&c->d_dist, sizeof(int32_t) * n));
  for (int i = 0; i < 2; i++) {
    PHDE_HIP(hipMalloc(&c->d_q[i], sizeof(uint32_t) * (n + 1)));
    PHDE_HIP(hipMalloc(&c->d_hq[i], sizeof(uint32_t) * hcap));
    PHDE_HIP(hipMalloc(&c->d_bm[i], sizeof(uint64_t) * nwords));
  }
  PHDE_HIP(hipMalloc(&c->d_slots, sizeof(BfsSlot) * 3));
  PHDE_HIP(hipMalloc(&c->d_kc, sizeof(KcResult)));
  PHDE_HIP(hipHostMalloc(&c->h_kc, sizeof(KcResult), hipHostMallocDefault));
  return PHDE_OK;
}

int bfs_run(phde_ctx *c, int64_t src, int alpha, int beta, int32_t *dist_out, int32_t *min_dist,
            bool sync_result) {
  int rc = bfs_setup(c);
  if (rc) return rc;
  if (src < 0 || src >= c->g.n) {
    set_error("bfs source out of range");
    return PHDE_ERR_ARG;
  }
  BfsArgs a;
  a.g = c->g;
  a.dist = dist_out ? dist_out : c->d_dist;
  a.q0 = c->d_q[0];
  a.q1 = c->d_q[1];
  a.h0 = c->d_hq[0];
  a.h1 = c->d_hq[1];
  a.bm0 = c->d_bm[0];
  a.bm1 = c->d_bm[1];
  a.slots = c->d_slots;
  a.src = src;
  a.alpha = alpha > 0 ? alpha : 15;
  a.beta = beta > 0 ? beta : 18;
  a.min_dist = min_dist;
  a.kc = c->d_kc;
  void *args[] = {&a};
  hipEvent_t ev = nullptr;
  c->timer.begin(c->stream, PHDE_KERNEL_BFS_LEVEL, &ev);
  PHDE_HIP(hipLaunchCooperativeKernel((const void *)bfs_persistent_kernel, dim3(c->bfs_grid),
                                      dim3(kBfsBlock), args, 0, c->stream));
  c->timer.end(c->stream, PHDE_KERNEL_BFS_LEVEL, ev);
  if (sync_result) {
    PHDE_HIP(hipMemcpyAsync(c->h_kc, c->d_kc, sizeof(KcResult), hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  }
  return PHDE_OK;
}

int dist_finalize(phde_ctx *c, const int32_t *dist_in, int32_t *dist_out, int32_t from,
                  int32_t to) {
  const int64_t n = c->g.n;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  dist_finalize_kernel<<<blocks, 256, 0, c->stream>>>(dist_in, dist_out, n, from, to);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

}  // namespace phde
