// bfs.hip — direction-optimising breadth-first search for gfx950, one
// kernel launch per level, with the k-centers update fused into the last
// level kernel.
//
// Reference behaviour followed: DOBFS (gapbs/bfs.cc:144-184).  Levels are
// exact hop counts and do not depend on scheduling; the top-down /
// bottom-up switch uses the reference's alpha/beta rule (bfs.cc:161,
// 172-173), which only affects speed.  The epilogue restates the k-centers
// step of ParHDE.cpp:124-140 (min over columns, first maximiser).
//
// Design (MI355X): level kernels are enqueued ahead by the host
// (levels.h); each reads the frontier counts of its level (complete at the
// kernel boundary), evaluates the direction rule redundantly in every
// thread, and processes the level.  Block 0 carries the rule's state to the
// next level through a parity-indexed slot.  The frontier is a vertex queue
// for top-down levels; bottom-up levels test membership in a bitmap that the
// previous bottom-up level wrote with one 64-bit ballot store per wave (no
// atomics, no clearing); the first bottom-up level after a top-down one tests
// `level == L` on the distance array instead, so top-down levels never touch
// a bitmap.  Queue appends are wave-aggregated (one atomic per wave).
// Vertices of degree >= kHugeDegree are expanded by the whole grid, those of
// degree >= kBigDegree by one wave, the rest one lane each.  The same levels
// run in one persistent launch (every block resident) with a grid barrier between them
// (bfs_persistent_kernel), and graphs with a deep BFS take the multi-hop
// traversal (bfs_hops_kernel: several hops per barrier, label-correcting,
// exact levels).
#include <algorithm>
#include <cstdio>

#include "device_util.h"
#include "hops.h"
#include "levels.h"
#include "pbfs.h"
#include "phde_common.h"

namespace phde {

constexpr int kBfsBlock = 256;
constexpr int kHugeDegree = 8192;
constexpr int kBigDegree = 32;

// Append the lanes' `found` vertices (with row ranges) to the next frontier
// (queue, huge list, degree sum).  Must be called by all 64 lanes.
__device__ __forceinline__ void wave_append(bool found, uint32_t v, uint32_t b, uint32_t e,
                                            QEnt *q_next, uint32_t *h_next, BfsSlot *slot) {
  const unsigned long long mask = __ballot(found);
  if (mask == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  const int dv = found ? (int)(e - b) : 0;
  const bool huge = found && dv >= kHugeDegree;
  const unsigned long long hmask = __ballot(huge);
  const unsigned long long dsum = wave_sum_u64((unsigned long long)dv);
  unsigned int base = 0, hbase = 0;
  if (lane == leader) {
    base = (unsigned int)(atomicAdd(&slot->qs, ((unsigned long long)__popcll(mask) << 32) | dsum) >> 32);
    if (hmask) hbase = atomicAdd(&slot->huge, (unsigned int)__popcll(hmask));
  }
  base = __shfl(base, leader);
  hbase = __shfl(hbase, leader);
  const unsigned long long lt = lanemask_lt();
  if (found) st_qent(&q_next[base + __popcll(mask & lt)], QEnt{v, b, e, 0u});
  if (huge) st_dev(&h_next[hbase + __popcll(hmask & lt)], v);
}

// Claim v for the next level; on success also fetch its row range.
__device__ __forceinline__ bool discover(int32_t *dist, const uint32_t *rowptr, uint32_t v,
                                         int32_t next_level, uint32_t *b, uint32_t *e) {
  if (dist[v] != kUnvisited) return false;  // plain load: see the small-list path
  if (atomicCAS(&dist[v], kUnvisited, next_level) != kUnvisited) return false;
  *b = rowptr[v];
  *e = rowptr[v + 1];
  return true;
}

constexpr int kBuGroup = 8;  // neighbours a lane tests per round trip (bottom-up)
constexpr int kWideFrontier = 65536;  // frontier size from which a level reads through the caches

// Direction-rule state carried from level to level (bfs.cc:159-183).
struct DirState {
  int in_bu, prev_bu, after_bu, pad;
  long long edges_to_check, awake_old;
};

struct BfsArgs {
  DevGraph g;
  int32_t *dist;
  QEnt *q0, *q1;
  uint32_t *h0, *h1;
  uint64_t *bm0, *bm1;
  BfsSlot *slots;  // 3 rotating frontier counters
  DirState *dir;   // 2, by level parity
  int64_t src;
  int alpha, beta;
  int32_t *min_dist;  // nullable: fused k-centers update
  KcResult *kc;
  int *done_dev, *host_flags;
  int32_t *dummy;  // one scratch word per thread of the grid
  int backoff;     // barrier poll back-off (units of s_sleep 2)
  int bstride;     // persistent kernel: only blocks with blockIdx % bstride == 0 work
  int vpw_min;     // narrow top-down levels: fewest queue entries per wave
  int epw;         // narrow top-down levels: edges per wave aimed at
  int hops;        // multi-hop traversal: hops a wave expands per grid barrier
  unsigned int *overflow;  // multi-hop traversal: set when the queue would overflow
  uint32_t qcap;           // multi-hop traversal: queue entries usable (0: n + 1)
  int hop_batch;           // multi-hop traversal: queue entries a wave starts a round with
};

// Block index and grid size of the working blocks (all of them unless the
// persistent kernel runs on every bstride-th block only).
__device__ __forceinline__ int64_t work_block(const BfsArgs &a) { return blockIdx.x / a.bstride; }
__device__ __forceinline__ int64_t work_grid(const BfsArgs &a) { return gridDim.x / a.bstride; }

// Initial state: every vertex unvisited, frontier = {src}.
__device__ void bfs_init(const BfsArgs &a) {
  const int64_t n = a.g.n;
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  const int64_t gsize = work_grid(a) * blockDim.x;
  for (int64_t i = gtid; i < n; i += gsize) st_dev(&a.dist[i], (i == a.src) ? 0 : kUnvisited);
  if (gtid == 0) {
    const int dsrc = a.g.deg32[a.src];
    st_dev(&a.slots[0].qs, (1ull << 32) | (unsigned long long)dsrc);
    st_dev(&a.slots[0].huge, dsrc >= kHugeDegree ? 1u : 0u);
    st_qent(&a.q0[0], QEnt{(uint32_t)a.src, a.g.rowptr[a.src], a.g.rowptr[a.src + 1], 0u});
    st_dev(&a.h0[0], (unsigned int)a.src);
    for (int s = 1; s < 3; s++) {
      st_dev(&a.slots[s].qs, 0ull);
      st_dev(&a.slots[s].huge, 0u);
    }
    DirState d0{};
    d0.edges_to_check = a.g.m;
    a.dir[0] = d0;
    st_dev(&a.kc->key, 0ull);
    st_dev(&a.kc->sumsq[0], 0ull);
    st_dev(&a.kc->sumsq[1], 0ull);
    st_dev(&a.kc->unreached, 0ull);
    st_dev(reinterpret_cast<unsigned long long *>(&a.kc->depth), 0ull);
    *a.done_dev = -1;
  }
}

__global__ __launch_bounds__(kBfsBlock) void bfs_init_kernel(BfsArgs a) { bfs_init(a); }

// Fused epilogue: k-centers update + argmax, sum of squares, unreached;
// with want_depth also the largest level (the source's eccentricity) into
// kc->depth, which must be 0 beforehand.
__device__ void bfs_epilogue(const BfsArgs &a, bool want_depth = false) {
  const int64_t n = a.g.n;
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  const int64_t gsize = work_grid(a) * blockDim.x;
  const int lane = __lane_id();
  unsigned long long key = 0, unreached = 0;
  U128 sumsq{0, 0};
  int32_t maxd = 0;
  for (int64_t i = gtid; i < n; i += gsize) {
    const int32_t d = ld_dev(&a.dist[i]);
    if (d == kUnvisited) {
      unreached++;
    } else {
      u128_add(sumsq, (unsigned long long)d * (unsigned long long)d);
      maxd = d > maxd ? d : maxd;
    }
    if (a.min_dist) {
      int32_t md = a.min_dist[i];
      if (d < md) {
        md = d;
        a.min_dist[i] = md;
      }
      const unsigned long long kk =
          ((unsigned long long)(uint32_t)md << 32) |
          (0xffffffffull - (unsigned long long)original_id(a.g, (uint32_t)i));
      key = kk > key ? kk : key;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ok = __shfl_xor(key, o);
    key = ok > key ? ok : key;
    unreached += __shfl_xor(unreached, o);
  }
  sumsq = wave_sum_u128(sumsq);
  if (want_depth) maxd = wave_max_i32(maxd);
  if (lane == 0) {
    if (a.min_dist && key) atomicMax(&a.kc->key, key);
    atomic_add_u128(a.kc->sumsq, sumsq);
    if (unreached) atomicAdd(&a.kc->unreached, unreached);
    if (want_depth && maxd > 0) atomicMax(reinterpret_cast<unsigned long long *>(&a.kc->depth), (unsigned long long)maxd);
  }
}

// Frontier counts of level L, read by one thread per block and broadcast
// through LDS: thousands of waves loading the same counter word at once
// right after a barrier congest its memory channel and stall every other
// access behind them.
struct LevelCounts {
  unsigned int nq, nh;
  unsigned long long sc;
};
__device__ __forceinline__ LevelCounts read_counts(const BfsSlot *cur) {
  // two waves load at once: (count, degree sum), and the huge count (one
  // wave's device-scope loads would each wait for the previous)
  __shared__ LevelCounts sh;
  if (threadIdx.x == 0) {
    const unsigned long long qs = ld_dev(&cur->qs);
    sh.nq = (unsigned int)(qs >> 32);
    sh.sc = qs & 0xffffffffull;
  } else if (threadIdx.x == kWave) {
    sh.nh = ld_dev(&cur->huge);
  }
  __syncthreads();
  const LevelCounts r = sh;
  __syncthreads();
  return r;
}

// Top-down expansion of the lanes' short adjacency lists (ds neighbours
// from adj[b], fewer than kBigDegree; maxd their wave maximum), G
// neighbours per lane per round: all loads, level checks and claims of a
// round in flight together, the winners appended with one atomic per wave.
template <int G>
__device__ __forceinline__ void small_lists(const BfsArgs &a, uint32_t b, int ds, int maxd, int32_t next_level,
                                            QEnt *q_next, uint32_t *h_next, BfsSlot *nxt) {
  const int lane = __lane_id();
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  const uint32_t *rowptr = a.g.rowptr;
  const uint32_t *adj = a.g.adj;
  for (int g0 = 0; g0 < maxd; g0 += G) {
    uint32_t vv[G], vb[G], ve[G];
    bool c[G];
#pragma unroll
    for (int j = 0; j < G; j++) {
      c[j] = g0 + j < ds;
      vv[j] = adj[c[j] ? b + g0 + j : 0u];  // unconditional: no branch per neighbour
    }
#pragma unroll
    for (int j = 0; j < G; j++) vv[j] = c[j] ? vv[j] : 0u;
    // Visited pre-check with plain (cacheable) loads, all issued before
    // any comparison: a stale line can only still read "unvisited"
    // (levels go from kUnvisited to a value once per traversal, and
    // each traversal starts in a new kernel), and the claim below decides.
    int32_t dd[G];
#pragma unroll
    for (int j = 0; j < G; j++) dd[j] = a.dist[c[j] ? vv[j] : 0u];
#pragma unroll
    for (int j = 0; j < G; j++) c[j] = c[j] && dd[j] == kUnvisited;
    // Claims, all issued before any result is used: lanes without a
    // candidate aim at a private dummy word, so the atomics need no
    // branch (a predicated atomic makes the compiler wait for each result
    // before issuing the next).  min(level) claims v exactly once: only
    // the first caller sees kUnvisited.
    int32_t won[G];
    int32_t *dummy = a.dummy + gtid;
#pragma unroll
    for (int j = 0; j < G; j++) won[j] = atomicMin(c[j] ? &a.dist[vv[j]] : dummy, next_level);
#pragma unroll
    for (int j = 0; j < G; j++) {
      const uint32_t r = c[j] ? vv[j] : 0u;
      vb[j] = rowptr[r];
      ve[j] = rowptr[r + 1];
    }
    // Keep the row-pointer loads in flight with the claims: the compiler
    // otherwise sinks them below the claims' results (waits on them, the
    // wave scan and the early exit first), one more round trip per group;
    // issued here the RGG 4.2M traversal went from 23.0 to 21.3 ms.  The
    // degree sum below is taken before the early exit for the same reason.
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < G; j++) c[j] = c[j] && won[j] == kUnvisited;
    int cnt = 0, hcnt = 0;
    unsigned long long dsum = 0;
#pragma unroll
    for (int j = 0; j < G; j++) {
      cnt += c[j] ? 1 : 0;
      const uint32_t dv = c[j] ? ve[j] - vb[j] : 0u;
      dsum += dv;
      hcnt += (c[j] && dv >= (uint32_t)kHugeDegree) ? 1 : 0;
    }
    int total = 0;
    const int off = wave_excl_scan(cnt, &total);
    const unsigned long long tsum = wave_sum_u64(dsum);
    if (total == 0) continue;
    unsigned int qb = 0;
    if (lane == 0) qb = (unsigned int)(atomicAdd(&nxt->qs, ((unsigned long long)total << 32) | tsum) >> 32);
    qb = __shfl(qb, 0);
    unsigned int pos = qb + (unsigned int)off;
#pragma unroll
    for (int j = 0; j < G; j++)
      if (c[j]) st_qent(&q_next[pos++], QEnt{vv[j], vb[j], ve[j], 0u});
    if (__ballot(hcnt > 0)) {
#pragma unroll
      for (int j = 0; j < G; j++)
        if (c[j] && ve[j] - vb[j] >= (uint32_t)kHugeDegree) st_dev(&h_next[atomicAdd(&nxt->huge, 1u)], vv[j]);
    }
  }
}

// What the previous level left for guessing this one's batches: its
// frontier count (0: no guess) and batch size.
struct LevelGuess {
  unsigned int nq;
  int vpw;
};

// One BFS level.  st holds the direction-rule state on entry (the same in
// every thread) and is advanced, as is gs.  Returns false when the frontier
// of level L is empty (the traversal is complete).
__device__ bool bfs_step(const BfsArgs &a, int L, DirState &st, LevelGuess &gs) {
  const int64_t n = a.g.n;
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  const int64_t gsize = work_grid(a) * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave;
  const int64_t nwaves = gsize / kWave;
  const uint32_t *rowptr = a.g.rowptr;
  const uint32_t *adj = a.g.adj;

  BfsSlot *cur = &a.slots[L % 3];
  BfsSlot *nxt = &a.slots[(L + 1) % 3];
  BfsSlot *old = &a.slots[(L + 2) % 3];
  const QEnt *q_cur = (L & 1) ? a.q1 : a.q0;
  // The wave's first batch of a top-down level's queue, loaded before the
  // level's counts arrive so that the two round trips overlap, on the guess
  // that the batch size stays 64 and the count below 1.5 x the previous
  // one (+ one batch).  Used only if the batch size is 64 after all;
  // entries below the count are final once the barrier has passed, the
  // others are ignored.  Every lane of a guessing wave loads: indices up to
  // n are in bounds (the queue holds n + 1 entries) and the count is at
  // most n.
  const int64_t wrank = (int64_t)(threadIdx.x / kWave) * work_grid(a) + work_block(a);
  const int64_t first = wrank * kWave;
  const bool spec = gs.nq != 0 && gs.vpw == kWave && !st.prev_bu &&
                    first < (int64_t)gs.nq + (gs.nq >> 1) + kWave;
  QEnt q_first{0u, 0u, 0u, 0u};
  if (spec && first + lane <= n) q_first = ld_qent(&q_cur[first + lane]);
  const LevelCounts cnt = read_counts(cur);
  const unsigned int nq = cnt.nq;
  gs.nq = nq;
  gs.vpw = 0;
  if (nq == 0) return false;
  const unsigned int nh = cnt.nh;
  const unsigned long long sc = cnt.sc;
  // ---- direction rule (bfs.cc:159-183), same result in every thread
  bool bu;
  if (st.in_bu) {
    const long long awake = nq;
    if (awake >= st.awake_old || awake > n / (long long)a.beta) {
      st.awake_old = awake;
      bu = true;
    } else {
      st.in_bu = 0;  // leave with scout_count = 1 (bfs.cc:174-175)
      st.after_bu = 1;
      bu = false;
    }
  } else {
    bu = false;
  }
  if (!st.in_bu) {
    const long long scout = st.after_bu ? 1 : (long long)sc;
    st.after_bu = 0;
    if (scout > st.edges_to_check / a.alpha) {
      st.in_bu = 1;
      st.awake_old = nq;
      bu = true;
    } else {
      st.edges_to_check -= scout;
      bu = false;
    }
  }
  const bool prev_bu = st.prev_bu != 0;
  st.prev_bu = bu ? 1 : 0;
  // Wide levels read the level's data through the caches.  Every write that
  // another block reads (levels, claims, queue entries, bitmap words) is a
  // device-scope store or atomic that reached memory before the barrier;
  // an agent-scope acquire here invalidates this CU's L1 and the XCD's L2
  // copies, so plain loads below see the state at the start of the level
  // and repeated reads (bitmap words, levels of hub neighbours) hit L2.  The
  // invalidation costs too much for the thousands of narrow levels of a
  // high-diameter graph, which keep their device-scope loads.
  const bool wide = bu || prev_bu || nq > (unsigned int)kWideFrontier;
  if (wide) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  // vertices found bottom-up: per-block total, published with one device
  // atomic per block at the end
  __shared__ unsigned int blk_found;
  if (bu) {
    if (threadIdx.x == 0) blk_found = 0;
    __syncthreads();
  }
  if (gtid == 0) {
    st_dev(&old->qs, 0ull);
    st_dev(&old->huge, 0u);
  }
  QEnt *q_next = (L & 1) ? a.q0 : a.q1;
  const uint32_t *h_cur = (L & 1) ? a.h1 : a.h0;
  uint32_t *h_next = (L & 1) ? a.h0 : a.h1;
  const uint64_t *bm_cur = (L & 1) ? a.bm1 : a.bm0;
  uint64_t *bm_next = (L & 1) ? a.bm0 : a.bm1;
  const int32_t next_level = L + 1;

  if (!bu) {
    // ---------------- top-down (bfs.cc:76-106)
    for (unsigned int hi = 0; hi < nh; hi++) {
      const uint32_t u = ld_dev(&h_cur[hi]);
      const uint32_t b = rowptr[u], e = rowptr[u + 1];
      for (int64_t k0 = (int64_t)b + gwave * kWave; k0 < (int64_t)e; k0 += nwaves * kWave) {
        const int64_t k = k0 + lane;
        const bool ok = k < (int64_t)e;
        const uint32_t v = ok ? adj[k] : 0u;
        uint32_t vb = 0, ve = 0;
        const bool f = ok && discover(a.dist, rowptr, v, next_level, &vb, &ve);
        wave_append(f, v, vb, ve, q_next, h_next, nxt);
      }
    }
    // The frontier: the queue, or after a bottom-up level its bitmap (a
    // bottom-up level builds no queue; bfs.cc:109-132 converts instead).
    // It is cut into batches of vpw entries, one wave each, dealt to the
    // blocks round-robin (consecutive batches on one CU queue up behind its
    // memory pipeline: each wave gathers 16 x 64 scattered lines per phase).
    // A narrow queue gets batches of about epw edges (the frontier's mean
    // degree is sc / nq): high-degree frontiers are spread over more waves,
    // while low-degree ones keep 64 entries per wave, since every wave that
    // appends costs an atomic on the shared queue counter.
    int vpw = kWave;
    if (!prev_bu && !wide) {
      const int64_t want = (int64_t)a.epw * (int64_t)nq / (int64_t)(sc > 0 ? sc : 1);
      vpw = want < a.vpw_min ? a.vpw_min : want > kWave ? kWave : (int)want;
    }
    gs.vpw = prev_bu ? 0 : vpw;
    const int64_t nitems = prev_bu ? ((n + 63) / 64) * 64 : (int64_t)nq;
    const int64_t nbatch = (nitems + vpw - 1) / vpw;
    for (int64_t k = wrank; k < nbatch; k += nwaves) {
      const int64_t base = k * vpw;
      const int64_t idx = base + lane;
      QEnt q{0u, 0u, 0u, 0u};
      if (prev_bu) {
        const uint64_t word = bm_cur[k];  // vpw == 64 here
        if (idx < n && ((word >> lane) & 1ull)) q = QEnt{(uint32_t)idx, rowptr[idx], rowptr[idx + 1], 0u};
      } else if (lane < vpw && idx < (int64_t)nq) {
        q = (spec && k == wrank && vpw == kWave) ? q_first : wide ? q_cur[idx] : ld_qent(&q_cur[idx]);
      }
      int d = (int)(q.e - q.b);
      if (d >= kHugeDegree && !prev_bu) d = 0;  // expanded above
      const bool big = d >= kBigDegree;
      unsigned long long bmask = __ballot(big);
      while (bmask) {
        const int l = __ffsll((long long)bmask) - 1;
        bmask &= bmask - 1;
        const uint32_t bb = __shfl(q.b, l), be = __shfl(q.e, l);
        for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
          const uint32_t k = k0 + lane;
          const bool ok = k < be;
          const uint32_t v = ok ? adj[k] : 0u;
          uint32_t vb = 0, ve = 0;
          const bool f = ok && discover(a.dist, rowptr, v, next_level, &vb, &ve);
          wave_append(f, v, vb, ve, q_next, h_next, nxt);
        }
      }
      // Small lists: each lane issues the loads, checks and claims of up to
      // G neighbours at once, then the wave appends all winners with one
      // atomic.  G is the wave's longest list rounded up to 8, 16 or 24 (one
      // round), else 16 per round.
      const int ds = big ? 0 : d;
      const int maxd = wave_max_i32(ds);
      if (maxd > 0) {
        if (maxd <= 8)
          small_lists<8>(a, q.b, ds, maxd, next_level, q_next, h_next, nxt);
        else if (maxd <= 16 || maxd > 24)
          small_lists<16>(a, q.b, ds, maxd, next_level, q_next, h_next, nxt);
        else
          small_lists<24>(a, q.b, ds, maxd, next_level, q_next, h_next, nxt);
      }
    }
  } else {
    // ---------------- bottom-up (bfs.cc:48-73)
    const int64_t nwords = (n + 63) / 64;
    for (int64_t w = gwave; w < nwords; w += nwaves) {
      const int64_t u = w * 64 + lane;
      bool f = false;
      uint32_t ub = 0, ue = 0;
      if (u < n && a.dist[u] == kUnvisited) {
        const uint32_t b = rowptr[u], e = rowptr[u + 1];
        // kBuGroup neighbours per round trip: all their ids, then all their
        // frontier tests, are in flight together; stop at the first group
        // that holds a frontier neighbour (the reference stops at the first
        // neighbour; the level written is the same)
        for (uint32_t k0 = b; k0 < e && !f; k0 += kBuGroup) {
          uint32_t vv[kBuGroup];
#pragma unroll
          for (int j = 0; j < kBuGroup; j++) vv[j] = adj[(k0 + j < e) ? k0 + j : b];
          bool hit[kBuGroup];
          if (prev_bu) {
#pragma unroll
            for (int j = 0; j < kBuGroup; j++) hit[j] = ((bm_cur[vv[j] >> 6] >> (vv[j] & 63)) & 1ull) != 0;
          } else {
#pragma unroll
            for (int j = 0; j < kBuGroup; j++) hit[j] = a.dist[vv[j]] == L;
          }
#pragma unroll
          for (int j = 0; j < kBuGroup; j++) f = f || (k0 + j < e && hit[j]);
        }
        if (f) st_dev(&a.dist[u], next_level);
        ub = b;
        ue = e;
      }
      const unsigned long long mask = __ballot(f);
      if (lane == 0) {
        st_dev((unsigned long long *)&bm_next[w], mask);
        if (mask) atomicAdd(&blk_found, (unsigned int)__popcll(mask));
      }
      (void)ub;
      (void)ue;
    }
  }
  if (bu) {
    __syncthreads();
    if (threadIdx.x == 0 && blk_found) atomicAdd(&nxt->qs, (unsigned long long)blk_found << 32);
  }
  return true;
}

// One launch per level (the host enqueues them ahead, levels.h); the
// direction state travels between launches through a parity slot.
__global__ __launch_bounds__(kBfsBlock) void bfs_level_kernel(BfsArgs a, int L) {
  __shared__ int finished;
  __shared__ DirState st_sh;
  if (threadIdx.x == 0) {
    finished = level_finished(a.done_dev, L);
    st_sh = a.dir[L & 1];
  }
  __syncthreads();
  if (finished) return;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  DirState st = st_sh;
  LevelGuess gs{0u, 0};  // no guess across launches
  if (!bfs_step(a, L, st, gs)) {
    if (gtid == 0) {
      level_done(a.done_dev, a.host_flags, L - 1);
      a.kc->depth = L - 1;
    }
    bfs_epilogue(a);
    return;
  }
  if (gtid == 0) {
    level_publish(a.host_flags, L);
    a.dir[(L + 1) & 1] = st;
  }
}

// The whole traversal in one persistent launch (every block resident): levels separated by the
// grid barrier above.  *arrive must be 0 at launch; *err receives 1 if a
// barrier wait timed out (a block whose wait timed out leaves the loop, the
// others then time out too, so the kernel always ends).
__global__ __launch_bounds__(kBfsBlock) void bfs_persistent_kernel(BfsArgs a, unsigned int *arrive,
                                                                   unsigned int *err) {
  if (blockIdx.x % a.bstride) return;
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  const unsigned int nb = (unsigned int)work_grid(a);
  unsigned int bar = 0;
  bfs_init(a);
  grid_barrier(arrive, nb * ++bar, err, a.backoff);
  DirState st{};
  st.edges_to_check = a.g.m;
  int L = 0;
  LevelGuess gs{1u, kWave};  // level 0: the source
  const int64_t maxL = a.g.n + 1;
  for (; L <= maxL; L++) {
    if (!bfs_step(a, L, st, gs)) break;
    if (grid_barrier(arrive, nb * ++bar, err, a.backoff)) break;
  }
  if (gtid == 0) a.kc->depth = L - 1;
  bfs_epilogue(a);
}


// ------------------------------------------------------------------ multi-hop traversal
// For deep graphs (thousands of narrow levels) the grid barrier between
// levels dominates.  This traversal lets each wave expand its batch of the
// frontier for `hops` hops between two grid barriers, keeping the vertices
// it discovers in a wave-local LDS queue (64 entries; the rest and the last
// hop's discoveries go to the global queue for the next round).  A vertex u
// reached from v (level l) is claimed with atomicMin(dist[u], l + 1) and
// expanded again whenever that lowers its level, so a vertex first reached
// by a longer path (another wave was slower to reach it) is corrected when
// the shorter path arrives: the levels at the end are the exact BFS levels,
// only the amount of work depends on timing (label-correcting search).
// Entries carry their level.  If a round would queue more entries than the
// queue holds, *overflow is set and the host repeats the traversal level by
// level.

// The multi-hop traversal in one persistent launch.  *arrive must be 0 at
// launch; *err receives 1 if a barrier wait timed out.
__global__ __launch_bounds__(kBfsBlock) void bfs_hops_kernel(BfsArgs a, unsigned int *arrive, unsigned int *err) {
  if (blockIdx.x % a.bstride) return;
  const unsigned int nb = (unsigned int)work_grid(a);
  const int64_t n = a.g.n;
  const uint32_t cap = (a.qcap > 0 && a.qcap < (uint64_t)(n + 1)) ? a.qcap : (uint32_t)(n + 1);  // queue entries
  const int64_t gtid = work_block(a) * blockDim.x + threadIdx.x;
  __shared__ QEnt locq[2][kBfsBlock / kWave][kWave];
  HopCtx h;
  h.g = a.g;
  h.lv = a.dist;
  h.col = 0;
  h.lvl_mask = 0xffffffffu;
  h.dummy = a.dummy + gtid;
  h.overflow = a.overflow;
  h.bound = kUnvisited - 1;
  h.cap = cap;
  h.hops = a.hops;
  h.batch = a.hop_batch;
  h.wrank = (int64_t)(threadIdx.x / kWave) * work_grid(a) + work_block(a);
  h.nwaves = work_grid(a) * (blockDim.x / kWave);
  unsigned int bar = 0;
  bfs_init(a);
  if (grid_barrier(arrive, nb * ++bar, err, a.backoff)) return;
  const int64_t maxR = n + 1;
  bool finished = false;
  for (int64_t R = 0; R <= maxR; R++) {
    BfsSlot *cur = &a.slots[R % 3];
    BfsSlot *nxt = &a.slots[(R + 1) % 3];
    BfsSlot *old = &a.slots[(R + 2) % 3];
    const LevelCounts cnt = read_counts(cur);
    if (cnt.nq == 0) {
      finished = true;
      break;
    }
    if (gtid == 0) {
      st_dev(&old->qs, 0ull);
      st_dev(&old->huge, 0u);
    }
    hop_round(h, (R & 1) ? a.q1 : a.q0, cnt.nq < cap ? cnt.nq : cap, (R & 1) ? a.q0 : a.q1, &nxt->qs,
              &locq[0][0][0]);
    if (grid_barrier(arrive, nb * ++bar, err, a.backoff)) return;
  }
  if (!finished && gtid == 0) atomicOr(a.overflow, 1u);  // round limit reached
  bfs_epilogue(a, true);
}

__global__ void dist_finalize_kernel(const int32_t *in, int32_t *out, int64_t n, int32_t from,
                                     int32_t to) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = in[i];
    out[i] = d == from ? to : d;
  }
}

int flags_setup(phde_ctx *c) {
  if (c->h_flags) return PHDE_OK;
  PHDE_HIP(hipHostMalloc(&c->h_flags, sizeof(int) * 16, hipHostMallocMapped));
  PHDE_HIP(hipHostGetDevicePointer((void **)&c->d_flags, c->h_flags, 0));
  PHDE_HIP(hipMalloc(&c->d_done, sizeof(int) * 4));
  return PHDE_OK;
}

int bfs_setup(phde_ctx *c) {
  if (c->bfs_grid > 0) return PHDE_OK;
  // one 256-thread block per CU: measured fastest for the persistent kernel
  // (fewer barrier participants); the per-level launches use two
  int per_cu = 1;
  if (const char *e = getenv("PHDE_BFS_BLOCKS_PER_CU")) per_cu = std::max(1, atoi(e));
  int resident = 0;
  PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, bfs_persistent_kernel, kBfsBlock, 0));
  c->bfs_persistent = resident >= 1;
  if (const char *e = getenv("PHDE_BFS_PERSISTENT")) c->bfs_persistent = atoi(e) != 0 && resident >= 1;
  if (c->bfs_persistent)
    per_cu = std::min(per_cu, resident);
  else if (!getenv("PHDE_BFS_BLOCKS_PER_CU"))
    per_cu = 2;
  c->bar_backoff = 1;
  if (const char *e = getenv("PHDE_BAR_BACKOFF")) c->bar_backoff = std::max(0, atoi(e));
  // work on every bstride-th block only (experiments with fewer XCDs:
  // blocks are dealt to the XCDs round-robin)
  c->bfs_stride = 1;
  if (const char *e = getenv("PHDE_BFS_STRIDE")) c->bfs_stride = std::max(1, std::min(64, atoi(e)));
  c->bfs_hops = 0;
  if (const char *e = getenv("PHDE_BFS_HOPS")) c->bfs_hops = std::max(0, std::min(64, atoi(e)));
  // a smaller multi-hop queue (tests of the level-by-level repeat)
  c->bfs_hop_batch = 16;
  if (const char *e = getenv("PHDE_BFS_HOP_BATCH")) c->bfs_hop_batch = std::max(1, std::min(kWave, atoi(e)));
  c->bfs_qcap = 0;
  if (const char *e = getenv("PHDE_BFS_QCAP")) c->bfs_qcap = (uint32_t)std::max(1, atoi(e));
  c->bfs_vpw = 1;
  if (const char *e = getenv("PHDE_BFS_VPW")) c->bfs_vpw = std::max(1, std::min(kWave, atoi(e)));
  c->bfs_epw = 1024;
  if (const char *e = getenv("PHDE_BFS_EPW")) c->bfs_epw = std::max(1, std::min(1 << 20, atoi(e)));
  c->bfs_grid = std::max(1, c->num_cus) * per_cu;
  const int64_t n = c->g.n;
  const int64_t nwords = (n + 63) / 64;
  const int64_t hcap = c->g.m / kHugeDegree + 2;
  PHDE_HIP(hipMalloc(&c->d_dist, sizeof(int32_t) * n));
  for (int i = 0; i < 2; i++) {
    PHDE_HIP(hipMalloc(&c->d_q[i], sizeof(QEnt) * (n + 1)));
    PHDE_HIP(hipMalloc(&c->d_hq[i], sizeof(uint32_t) * hcap));
    PHDE_HIP(hipMalloc(&c->d_bm[i], sizeof(uint64_t) * nwords));
  }
  PHDE_HIP(hipMalloc(&c->d_slots, sizeof(BfsSlot) * 3 + sizeof(DirState) * 2));
  PHDE_HIP(hipMalloc(&c->d_kc, sizeof(KcResult)));
  PHDE_HIP(hipHostMalloc(&c->h_kc, sizeof(KcResult), hipHostMallocDefault));
  PHDE_HIP(hipMalloc(&c->d_bar, sizeof(unsigned int) * 64));
  PHDE_HIP(hipMalloc(&c->d_dummy, sizeof(int32_t) * (size_t)c->bfs_grid * kBfsBlock));
  PHDE_HIP(hipHostMalloc(&c->h_bar, sizeof(unsigned int) * 4, hipHostMallocDefault));
  return flags_setup(c);
}

// A graph with a deep BFS (thousands of narrow levels, each bound by its
// chain of dependent round trips and the grid barrier) takes the multi-hop
// traversal: 12 hops per grid barrier from batches of 16 queue entries per
// wave.  Measured on RGG 4.2M: 15.5 ms per traversal level by level, 7.7 ms
// multi-hop (8.1 at 8 hops, 7.8 at 16, 8.4 at 24; 9.4 from batches of 8,
// 7.8 from 24, 11.0 from 48 at 3 hops); RGG 16.8M 37.9 -> 19.5 ms.  Level
// by level (PHDE_BFS_HOPS=0) such a graph runs faster on the even blocks
// only, one block per CU on 4 of the 8 XCDs (blocks are dealt to the XCDs
// round-robin): 16.5 -> 15.6 ms; the multi-hop traversal prefers every CU
// (7.7 vs 8.0 ms).  Low-diameter graphs keep the level-by-level traversal on
// every CU: their wide levels need the whole chip, and the multi-hop one
// re-expands too much (RMAT scale 22: 0.78 ms, 46 ms at 4 hops).
void bfs_deep_graph(phde_ctx *c) {
  if (!getenv("PHDE_BFS_HOPS")) c->bfs_hops = 12;
  if (!getenv("PHDE_BFS_STRIDE") && c->bfs_hops <= 1 && c->bfs_grid >= 16) c->bfs_stride = 2;
}

int bfs_run(phde_ctx *c, int64_t src, int alpha, int beta, int32_t *dist_out, int32_t *min_dist,
            bool sync_result) {
  int rc = bfs_setup(c);
  if (rc) return rc;
  if (src < 0 || src >= c->g.n) {
    set_error("bfs source out of range");
    return PHDE_ERR_ARG;
  }
  if (c->pg) {
    // deep graph cut into parts: the partitioned traversal (pbfs.hip)
    bool ran = false;
    hipEvent_t pev = nullptr;
    c->timer.begin(c->stream, PHDE_KERNEL_BFS_LEVEL, &pev);
    rc = pbfs_run(c, src, dist_out, min_dist, &ran);
    c->timer.end(c->stream, PHDE_KERNEL_BFS_LEVEL, pev);
    if (rc || ran) return rc;
  }
  BfsArgs a;
  a.g = c->g;
  a.dist = dist_out ? dist_out : c->d_dist;
  a.q0 = (QEnt *)c->d_q[0];
  a.q1 = (QEnt *)c->d_q[1];
  a.h0 = c->d_hq[0];
  a.h1 = c->d_hq[1];
  a.bm0 = c->d_bm[0];
  a.bm1 = c->d_bm[1];
  a.slots = c->d_slots;
  a.dir = (DirState *)(c->d_slots + 3);
  a.src = src;
  a.alpha = alpha > 0 ? alpha : 15;
  a.beta = beta > 0 ? beta : 18;
  a.min_dist = min_dist;
  a.kc = c->d_kc;
  a.done_dev = c->d_done;
  a.host_flags = c->d_flags;
  a.dummy = c->d_dummy;
  a.vpw_min = c->bfs_vpw;
  a.epw = c->bfs_epw;
  a.backoff = c->bar_backoff;
  a.bstride = c->bfs_persistent ? c->bfs_stride : 1;
  c->h_flags[kFlagDone] = -1;
  c->h_flags[kFlagProgress] = -1;
  hipEvent_t ev = nullptr;
  c->timer.begin(c->stream, PHDE_KERNEL_BFS_LEVEL, &ev);
  const int grid = c->bfs_grid;
  bool hops = c->bfs_persistent && c->bfs_hops > 1;
  a.hops = c->bfs_hops;
  a.qcap = c->bfs_qcap;
  a.hop_batch = c->bfs_hop_batch;
  a.overflow = c->d_bar + 2;
  if (c->bfs_persistent) {
    PHDE_TRY_RC(fill_u32(c, c->d_bar, 3, 0u));
    unsigned int *arrive = c->d_bar, *err = c->d_bar + 1;
    void *args[] = {&a, &arrive, &err};
    const hipError_t e = launch_resident(hops ? (const void *)bfs_hops_kernel : (const void *)bfs_persistent_kernel,
                                         grid, kBfsBlock, args, 0, c->stream);
    if (e != hipSuccess) {
      // this context then launches one kernel per level, which needs no
      // co-residency
      (void)hipGetLastError();
      c->bfs_persistent = false;
      a.bstride = 1;
      hops = false;
    }
  }
  if (!c->bfs_persistent) {
    bfs_init_kernel<<<grid, kBfsBlock, 0, c->stream>>>(a);
    PHDE_HIP(hipGetLastError());
    rc = run_levels(c, [&](int L) -> int {
      bfs_level_kernel<<<grid, kBfsBlock, 0, c->stream>>>(a, L);
      PHDE_HIP(hipGetLastError());
      return PHDE_OK;
    });
    if (rc) return rc;
  }
  c->timer.end(c->stream, PHDE_KERNEL_BFS_LEVEL, ev);
  if (sync_result || c->bfs_persistent) {
    PHDE_HIP(hipMemcpyAsync(c->h_kc, c->d_kc, sizeof(KcResult), hipMemcpyDeviceToHost, c->stream));
    if (c->bfs_persistent)
      PHDE_HIP(hipMemcpyAsync(c->h_bar, c->d_bar + 1, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                              c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
    if (c->bfs_persistent && c->h_bar[0]) {
      // the grid barrier timed out (blocks kept from becoming resident):
      // this context goes on level by level.  Every level the partial run
      // wrote is the length of a real path, so min_dist ends the same
      // after the exact pass.
      std::fprintf(stderr, "parhde: BFS grid barrier timed out; traversing level by level from now on\n");
      c->bfs_persistent = false;
      return bfs_run(c, src, alpha, beta, dist_out, min_dist, sync_result);
    }
    if (hops && c->h_bar[1]) {
      // the multi-hop traversal outgrew its queue: repeat it level by level.
      // Its epilogue has already folded its levels into min_dist, but every
      // level it wrote is the length of a real path, so at least the exact
      // level, and min_dist ends the same after the exact pass.
      const int saved = c->bfs_hops;
      c->bfs_hops = 0;
      rc = bfs_run(c, src, alpha, beta, dist_out, min_dist, sync_result);
      c->bfs_hops = saved;
      return rc;
    }
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
