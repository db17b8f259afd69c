// pbfs.hip — partitioned breadth-first search for graphs with a deep BFS.
//
// Reference behaviour followed: DOBFS (gapbs/bfs.cc:144-184) — exact hop
// levels from one source — and the k-centers step of ParHDE.cpp:124-140
// (min over columns, first maximiser, lowest index on ties) fused at the end.
//
// A deep graph (thousands of narrow BFS levels: meshes, geometric graphs)
// spends its traversal on the latency of each level's dependent global
// round trips.  Here the graph is cut once, at context creation, into
// compact parts whose local CSR, levels and work lists fit in one CU's LDS,
// and one persistent kernel (every block resident) runs the search as a work queue of
// parts: a workgroup takes a part some of whose vertices were lowered from
// outside, runs rounds inside LDS from those vertices (each listed vertex
// offers its level + 1 to its neighbours; a neighbour whose level drops is
// listed for the next round), writes the part's lowered levels back and
// lowers the levels of its neighbours in other parts (atomicMin), queueing
// their parts.  Every level written is the length of a real path and levels
// only decrease, so when no part is queued or running the levels are the
// exact BFS levels (label-correcting); only the amount of work depends on
// timing.  A list may hold vertices of several levels: a lowering that
// arrives while a part runs is taken up at once instead of after the run's
// current wave (measured: 3.9 ms per traversal of RGG n=4.2M, 4.4 ms with
// level-synchronous rounds that seed later arrivals level by level).
//
// Per-vertex state lives in part-major order (part p's vertices at
// [vstart[p], vstart[p+1]), in the part's local order): `plevel`, the
// part's own levels (written only by a run of the part), and `ilevel`, the
// levels other parts lowered its boundary vertices to (atomicMin).  A run
// reads ilevel for the part's boundary vertices only; a workgroup that
// runs the same part again keeps its LDS image when no other workgroup ran
// the part in between (a version counter, kept in the part's state word).
//
// Part states: 0 idle, 1 queued, 2 running, 3 running and lowered again
// (the runner then runs it once more).  Lowering a vertex from outside
// marks its part with atomicOr(1): 0 -> 1 queues it (and counts it in
// `pending`), 2 -> 3 flags the runner.  `pending` counts parts queued or
// running, so it reaches 0 only when the search is over.  Every wait is
// bounded: a timeout sets an error word, every workgroup stops, and the
// host repeats the traversal with the multi-hop kernel.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "device_util.h"
#include "pbfs.h"

namespace phde {

constexpr int kPartBytes = 144 * 1024;   // LDS image of one part (see part_bytes)
constexpr uint32_t kNoPart = 0xffffffffu;
// Control words: the ticket counter and the error flag share a line; the
// queue tail and the count of parts queued or running share one 64-bit word
// on a line of its own (tail << 32 | pending), so that queueing a part
// takes one atomic.
constexpr int kCtlHead = 0, kCtlErr = 1, kCtlTP = 64, kCtlWords = 96;
// Ring slots are 64-bit, (entry index + 1) << 32 | (part + 1), one per
// 128-byte line: the workgroup holding ticket h waits on slot h alone.
constexpr int kSlotStride = 16;
constexpr uint32_t kDoneSlot = 0xffffffffu;  // low word of a slot: the search is over
__device__ __forceinline__ unsigned long long *tp_word(uint32_t *ctl) {
  return reinterpret_cast<unsigned long long *>(ctl + kCtlTP);
}

// A part's static data is one blob, copied into LDS as it is: local row
// offsets (nv + 1), local adjacency (ne), boundary list (nb), external
// edges' local sources and target-part slots (nx each) as uint16, then the
// external edges' targets (part-major index, nx) and the distinct
// neighbouring parts (nn) as uint32; 16-byte aligned.
__host__ __device__ constexpr int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }
__host__ __device__ constexpr int64_t blob_bytes(int64_t nv, int64_t ne, int64_t nb, int64_t nx, int64_t nn) {
  return align_up(align_up(2 * (nv + 1 + ne + nb + 2 * nx), 4) + 4 * nx + 4 * nn, 16);
}
// LDS bytes of a part with nv vertices, ne internal and nx external
// (directed) edges, nb boundary vertices and nn neighbouring parts: the
// blob, levels with their list stamps (uint64), two vertex lists (uint16,
// nv + 1 entries), changed flags (nv) and lowered-neighbour flags (nn).
__host__ __device__ constexpr int64_t part_bytes(int64_t nv, int64_t ne, int64_t nb, int64_t nx, int64_t nn) {
  return blob_bytes(nv, ne, nb, nx, nn) + 8 * nv + 4 * (nv + 1) + nv + nn + 64;
}
// Rounds per run of a part (PHDE_PBFS_CAP): a run that reaches the cap
// writes back and pushes what it has, then goes on in place from its list,
// so that a part does not race ahead on levels that a neighbour is about to
// lower.  Measured on RGG n=4.2M: 3.93 ms per traversal uncapped, 4.3 at 32
// rounds, 4.5 at 16, 5.1 at 8 -- each extra run costs a write-back and push
// (about 6 us) that the work it saves does not repay -- so runs are
// uncapped by default.
constexpr int kRoundCap = 1 << 30;

struct PbfsArgs {
  DevGraph g;
  int32_t *level;   // output: the traversal's levels (internal ids)
  int32_t *plevel;  // part-major: the part's own levels
  int32_t *ilevel;  // part-major: levels lowered from outside
  const uint32_t *vstart, *pmidx, *pcount;  // pcount: (ne, nb, nx, nn) per part
  const uint4 *blob;
  const uint32_t *blob_off;  // per part, in 16-byte units
  uint32_t *st, *ctl;  // st: version << 2 | state
  unsigned long long *ring;  // ring_n slots, kSlotStride words apart
  uint32_t ring_n, src_part, src_loc;
  int64_t k, src;
  int32_t *min_dist;  // nullable: fused k-centers update
  KcResult *kc;
  unsigned int *arrive, *bar_err;
  int backoff;
  int cap;  // rounds per run of a part
  // with PHDE_PBFS_STATS: [0..3] per-phase wall time (100 MHz ticks) summed
  // over workgroups — waiting for a part, seeding (and loading images),
  // rounds, write-back and pushes; [4..6] activations, rounds, image loads;
  // [7, 8] first and last tick; [9, 10] expansion passes, runs going on
  // from the round cap; [13] listed vertices;
  // [16..19] queued-to-taken latency: sum, count, count above 10 us, max;
  // [32 + slot] the tick at which each ring slot was filled
  unsigned long long *stats;
};

__global__ void pbfs_init_kernel(PbfsArgs a) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const uint32_t src_pm = a.vstart[a.src_part] + a.src_loc;
  for (int64_t j = gtid; j < a.g.n; j += gsize) {
    a.plevel[j] = kUnvisited;
    a.ilevel[j] = j == src_pm ? 0 : kUnvisited;
  }
  for (int64_t p = gtid; p < a.k; p += gsize) a.st[p] = p == a.src_part ? 1u : 0u;
  for (int64_t r = gtid; r < a.ring_n; r += gsize)
    a.ring[r * kSlotStride] = r == 0 ? (1ull << 32) | (a.src_part + 1u) : 0ull;
  if (gtid < kCtlWords) a.ctl[gtid] = gtid == kCtlTP || gtid == kCtlTP + 1 ? 1u : 0u;  // tail 1, pending 1
  if (gtid == 0) {
    a.kc->key = 0;
    a.kc->sumsq[0] = a.kc->sumsq[1] = 0;
    a.kc->unreached = 0;
    a.kc->depth = 0;
    a.arrive[0] = 0;
    a.bar_err[0] = 0;
  }
}

// A vertex of part q was lowered from outside: queue q or flag its runner.
template <class A>
__device__ __forceinline__ void mark_lowered(const A &a, uint32_t q) {
  const uint32_t old = atomicOr(&a.st[q], 1u);
  if ((old & 3u) == 0u) {
    const uint32_t t = (uint32_t)(atomicAdd(tp_word(a.ctl), (1ull << 32) | 1ull) >> 32);
    if (a.stats) st_dev(&a.stats[32 + t % a.ring_n], __builtin_amdgcn_s_memrealtime());
    st_dev(&a.ring[(uint64_t)(t % a.ring_n) * kSlotStride], ((unsigned long long)(t + 1u) << 32) | (q + 1u));
  }
}

// After the last part is released (pending reached zero at tail `t`), the
// slots of tickets t .. t + gridDim.x - 1 -- the only ones still to be
// waited on, one per workgroup at most -- are marked done.
template <class A>
__device__ __forceinline__ void mark_done(const A &a, uint32_t t) {
  for (uint32_t j = 0; j < gridDim.x; j++) {
    const uint32_t u = t + j;
    st_dev(&a.ring[(uint64_t)(u % a.ring_n) * kSlotStride], ((unsigned long long)(u + 1u) << 32) | kDoneSlot);
  }
}

// The next queued part (thread 0 only), kNoPart when the search is over or
// an error stopped it.  Each call takes a ticket and waits on that ticket's
// slot alone, so idle workgroups poll separate lines instead of contending
// for one word.  Entry h of the queue lands in slot h % ring_n, tagged with
// h + 1; ring_n exceeds the parts plus the workgroups, so a slot is not
// reused while its ticket is still waiting (the bounded wait below turns a
// violation into an error, never into a wrong part).
template <class A>
__device__ uint32_t pop_part(const A &a) {
  const uint32_t h = atomicAdd(&a.ctl[kCtlHead], 1u);
  const unsigned long long *slot = &a.ring[(uint64_t)(h % a.ring_n) * kSlotStride];
  long spins = 0;
  for (;;) {
    const unsigned long long v = ld_dev(slot);
    if ((uint32_t)(v >> 32) == h + 1u) {
      const uint32_t q = (uint32_t)v;
      if (q == kDoneSlot) return kNoPart;
      if (a.stats && h > 0) {
        const unsigned long long tn = __builtin_amdgcn_s_memrealtime();
        const unsigned long long te = ld_dev(&a.stats[32 + h % a.ring_n]);
        const unsigned long long ql = tn > te ? tn - te : 0;
        atomicAdd(&a.stats[16], ql);
        atomicAdd(&a.stats[17], 1ull);
        if (ql > 1000) atomicAdd(&a.stats[18], 1ull);
        atomicMax(&a.stats[19], ql);
      }
      return q - 1u;
    }
    if ((spins & 63) == 63 && ld_dev(&a.ctl[kCtlErr])) return kNoPart;
    __builtin_amdgcn_s_sleep(1);
    if (++spins > (1L << 24)) {  // seconds with nothing queued: a lost wake-up
      atomicOr(&a.ctl[kCtlErr], 2u);
      return kNoPart;
    }
  }
}

// Wave-aggregated append of the lanes' `found` vertices to an LDS list;
// every lane of the wave must call it.
__device__ __forceinline__ void list_append(bool found, uint16_t v, uint16_t *list, int *count) {
  const unsigned long long m = __ballot(found);
  if (!m) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __builtin_amdgcn_readlane(base, leader);
  if (found) list[base + __popcll(m & lanemask_lt())] = v;
}

template <int Block, int Lanes, int Claims>
__global__ __launch_bounds__(Block) void pbfs_kernel(PbfsArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ uint32_t s_part, s_ver;
  __shared__ int s_cnt[3], s_rerun;
  __shared__ unsigned long long s_dummy[Block];  // targets of masked-off claims, one per thread
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  uint32_t cached = kNoPart, cached_ver = 0;  // the part whose image is in LDS, and its version
  // Rounds so far in this workgroup: each round's list is stamped with its
  // own epoch, so that a vertex joins it at most once.
  uint32_t epoch = 0;
  unsigned long long t_wait = 0, t_load = 0, t_round = 0, t_push = 0, n_act = 0, n_round = 0, n_img = 0;
  unsigned long long t_first = ~0ull, t_last = 0, t0 = 0, t1 = 0;
  unsigned long long n_pass = 0, n_carry = 0, n_front = 0;
  const bool stats = a.stats != nullptr;
  for (;;) {
    if (stats && tid == 0) t0 = __builtin_amdgcn_s_memrealtime();
    if (tid == 0) {
      const uint32_t p = pop_part(a);
      if (p != kNoPart) {
        // the exchange completes before the part's levels are read (every
        // datum shared between workgroups is a device-scope access, so
        // completing this wave's own accesses orders them; no cache
        // write-back is needed)
        s_ver = atomicAdd(&a.st[p], 1u) >> 2;  // queued -> running; the part's version
        __builtin_amdgcn_s_waitcnt(0);
      }
      s_part = p;
    }
    __syncthreads();
    const uint32_t p = s_part;
    if (stats && tid == 0) {
      t1 = __builtin_amdgcn_s_memrealtime();
      t_wait += t1 - t0;
      if (p != kNoPart && t1 < t_first) t_first = t1;
    }
    if (p == kNoPart) break;
    const uint32_t v0 = a.vstart[p];
    const int nv = (int)(a.vstart[p + 1] - v0);
    const int ne = (int)a.pcount[4 * p], nb = (int)a.pcount[4 * p + 1], nx = (int)a.pcount[4 * p + 2];
    const int nn = (int)a.pcount[4 * p + 3];
    const int nblob = (int)(a.blob_off[p + 1] - a.blob_off[p]);  // 16-byte words
    // LDS: the blob, then the per-run state
    uint16_t *lrow = reinterpret_cast<uint16_t *>(smem);
    uint16_t *ladj = lrow + nv + 1;
    uint16_t *blist = ladj + ne;
    uint16_t *xsrc = blist + nb;
    uint16_t *xnbr = xsrc + nx;
    const uint32_t *xdst = reinterpret_cast<const uint32_t *>(smem + align_up(2 * (nv + 1 + ne + nb + 2 * nx), 4));
    const uint32_t *nbr = xdst + nx;
    // level << 32 | epoch of the list the vertex last joined
    unsigned long long *lvl = reinterpret_cast<unsigned long long *>(smem + 16 * (int64_t)nblob);
    uint16_t *F0 = reinterpret_cast<uint16_t *>(lvl + nv);
    uint16_t *F1 = F0 + nv + 1;
    unsigned char *chg = reinterpret_cast<unsigned char *>(F1 + nv + 1);
    unsigned char *nflag = chg + nv;
    // the list being built (F0 / F1) and its counter (s_cnt[ci]); a run that
    // stops at the round cap leaves its list to the next run of the part
    int li = 0, ci = 0;
    bool carry = false;
    for (;;) {
      if (stats && tid == 0) {
        t0 = __builtin_amdgcn_s_memrealtime();
        n_act++;
        n_carry += carry ? 1 : 0;
      }
      if (cached != p || cached_ver != s_ver) {
        // the part's image: its blob (read-only, 16-byte loads, eight in
        // flight per thread) and its own levels (device-scope loads)
        const uint4 *gb = a.blob + a.blob_off[p];
        uint4 *sb = reinterpret_cast<uint4 *>(smem);
        for (int i = tid; i < nblob; i += 8 * Block) {
          uint4 t[8];
#pragma unroll
          for (int j = 0; j < 8; j++) t[j] = gb[i + j * Block < nblob ? i + j * Block : i];
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (i + j * Block < nblob) sb[i + j * Block] = t[j];
        }
        for (int i0 = 0; i0 < nv; i0 += 8 * Block) {
          int32_t t[8];
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int i = i0 + j * Block + tid;
            t[j] = i < nv ? ld_dev(&a.plevel[v0 + i]) : 0;
          }
#pragma unroll
          for (int j = 0; j < 8; j++) {
            const int i = i0 + j * Block + tid;
            if (i < nv) lvl[i] = (unsigned long long)(uint32_t)t[j] << 32;
          }
        }
        if (stats && tid == 0) n_img++;
      }
      for (int i = tid; i < nv; i += Block) chg[i] = 0;
      for (int i = tid; i < nn; i += Block) nflag[i] = 0;
      if (!carry) {
        // a new list: stamped with a new epoch
        epoch++;
        li = 0;
        ci = 0;
        if (tid == 0) s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
      }
      __syncthreads();
      // seeds: boundary vertices (and the source) lowered from outside join
      // the list, after any vertices carried over from the last run
      uint16_t *Fs = li ? F1 : F0;
      for (int bb = 0; bb <= nb; bb += Block) {
        const int b = bb + tid;
        int u = -1;
        if (b < nb)
          u = blist[b];
        else if (b == nb && p == a.src_part)
          u = (int)a.src_loc;
        bool add = false;
        if (u >= 0) {
          const int32_t l = ld_dev(&a.ilevel[v0 + u]);
          if ((uint32_t)l < (uint32_t)(lvl[u] >> 32)) {
            const unsigned long long old = atomicMin(&lvl[u], ((unsigned long long)(uint32_t)l << 32) | epoch);
            if ((uint32_t)(old >> 32) > (uint32_t)l) {
              chg[u] = 1;
              add = (uint32_t)old != epoch;
            }
          }
        }
        list_append(add, add ? (uint16_t)u : 0, Fs, &s_cnt[ci]);
      }
      __syncthreads();
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_load += t1 - t0;
        t0 = t1;
      }
      // Rounds: every listed vertex offers its own level + 1 to its
      // neighbours (label-correcting: the list may hold several levels);
      // a neighbour whose level drops joins the next list once.
      int nf = s_cnt[ci], r = 0;
      while (nf > 0 && r < a.cap) {
        const uint32_t en = epoch + 1u;  // the stamp of the list this round builds
        if (tid == 0) s_cnt[(ci + 2) % 3] = 0;
        const uint16_t *Fc = li ? F1 : F0;
        uint16_t *Fn = li ? F0 : F1;
        int *cn = &s_cnt[(ci + 1) % 3];
        if (stats && tid == 0) {
          n_pass += (nf + Block / Lanes - 1) / (Block / Lanes);
          n_front += nf;
        }
        // Lanes lanes per listed vertex; a lane claims up to Claims
        // neighbours (all reads and claims issued before any result is
        // used), then the wave appends its new entries with one atomic
        for (int f0 = 0; f0 < nf; f0 += Block / Lanes) {
          const int f = f0 + tid / Lanes;
          int xb = 0, e = 0;
          uint32_t nl = 0;
          if (f < nf) {
            const int u = Fc[f];
            xb = lrow[u];
            e = lrow[u + 1];
            nl = (uint32_t)(lvl[u] >> 32) + 1u;
          }
          xb += tid % Lanes;
          const unsigned long long offer = ((unsigned long long)nl << 32) | en;
          while (__ballot(xb < e)) {
            // unconditional reads and claims (clamped index, masked-off
            // lanes claim their own dummy word), all issued before any
            // result is used
            int w[Claims];
            unsigned long long old[Claims];
            bool ok[Claims];
#pragma unroll
            for (int q = 0; q < Claims; q++) {
              const int x = xb + q * Lanes;
              ok[q] = x < e;
              w[q] = ladj[ok[q] ? x : 0];
            }
#pragma unroll
            for (int q = 0; q < Claims; q++) old[q] = atomicMin(ok[q] ? &lvl[w[q]] : &s_dummy[tid], offer);
            // positions from ballots: new entry q of a lane goes after every
            // new entry q' < q of the wave and after entries q of lower lanes
            unsigned long long mq[Claims];
            int pre[Claims], total = 0;
#pragma unroll
            for (int q = 0; q < Claims; q++) {
              const bool low = ok[q] && (uint32_t)(old[q] >> 32) > nl;  // the level dropped
              if (low) chg[w[q]] = 1;
              mq[q] = __ballot(low && (uint32_t)old[q] != en);
              pre[q] = total;
              total += __popcll(mq[q]);
            }
            if (total) {
              int base = 0;
              if (lane == 0) base = atomicAdd(cn, total);
              base = __builtin_amdgcn_readfirstlane(base);
              const unsigned long long lt = lanemask_lt();
#pragma unroll
              for (int q = 0; q < Claims; q++)
                if ((mq[q] >> lane) & 1ull) Fn[base + pre[q] + __popcll(mq[q] & lt)] = (uint16_t)w[q];
            }
            xb += Claims * Lanes;
          }
        }
        __syncthreads();
        epoch = en;
        li ^= 1;
        ci = (ci + 1) % 3;
        nf = s_cnt[ci];
        r++;
      }
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_round += t1 - t0;
        t0 = t1;
        n_round += r;
      }
      // write back the lowered levels, then lower the neighbours in other
      // parts (up to four external edges per thread, their claims in flight
      // together) and flag their parts; one thread per flagged part then
      // queues it (or flags its runner)
      for (int i = tid; i < nv; i += Block)
        if (chg[i]) st_dev(&a.plevel[v0 + i], (int32_t)(lvl[i] >> 32));
      for (int x0 = 0; x0 < nx; x0 += 4 * Block) {
        int32_t nl[4], old[4];
        bool on[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int x = x0 + j * Block + tid;
          on[j] = false;
          nl[j] = 0;
          if (x < nx) {
            const int u = xsrc[x];
            on[j] = chg[u] != 0;
            nl[j] = (int32_t)(lvl[u] >> 32) + 1;
          }
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
          old[j] = on[j] ? atomicMin(&a.ilevel[xdst[x0 + j * Block + tid]], nl[j]) : 0;
#pragma unroll
        for (int j = 0; j < 4; j++)
          if (on[j] && old[j] > nl[j]) nflag[xnbr[x0 + j * Block + tid]] = 1;
      }
      __syncthreads();
      for (int i = tid; i < nn; i += Block)
        if (nflag[i]) mark_lowered(a, nbr[i]);
      // this run's stores and claims complete before the part is released
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_push += t1 - t0;
        if (t1 > t_last) t_last = t1;
      }
      if (tid == 0) {
        // this run's image becomes the part's next version, published with
        // the state (only the part's runner changes the version)
        const uint32_t v = s_ver;
        s_ver = v + 1u;
        if (nf > 0) {
          // stopped at the round cap: go on in place with the list; a
          // lowering from now on flags the part again
          atomicExch(&a.st[p], (v + 1u) << 2 | 2u);
          __builtin_amdgcn_s_waitcnt(0);
          s_rerun = 2;
        } else {
          const uint32_t old = atomicCAS(&a.st[p], v << 2 | 2u, (v + 1u) << 2);
          if (old == (v << 2 | 2u)) {
            s_rerun = 0;
            const unsigned long long tp = atomicSub(tp_word(a.ctl), 1ull);
            if ((uint32_t)tp == 1u) mark_done(a, (uint32_t)(tp >> 32));  // the last part: the search is over
          } else {
            // lowered again while it ran: run it once more
            atomicExch(&a.st[p], (v + 1u) << 2 | 2u);
            __builtin_amdgcn_s_waitcnt(0);
            s_rerun = 1;
          }
        }
      }
      __syncthreads();
      cached = p;
      cached_ver = s_ver;
      if (!s_rerun) break;
      carry = s_rerun == 2;
    }
  }
  if (stats && tid == 0) {
    atomicAdd(&a.stats[0], t_wait);
    atomicAdd(&a.stats[1], t_load);
    atomicAdd(&a.stats[2], t_round);
    atomicAdd(&a.stats[3], t_push);
    atomicAdd(&a.stats[4], n_act);
    atomicAdd(&a.stats[5], n_round);
    atomicAdd(&a.stats[6], n_img);
    atomicMin(&a.stats[7], t_first);
    atomicMax(&a.stats[8], t_last);
    atomicAdd(&a.stats[9], n_pass);
    atomicAdd(&a.stats[10], n_carry);
    atomicAdd(&a.stats[13], n_front);
  }
  // every workgroup has stopped taking parts; the levels are final unless
  // an error stopped the search
  if (grid_barrier(a.arrive, gridDim.x, a.bar_err, a.backoff)) return;
  if (ld_dev(&a.ctl[kCtlErr])) return;
  // levels to the output (internal ids) and the fused k-centers step:
  // min_dist = min(min_dist, level), the first maximiser (caller's id on
  // ties), exact sum of squares, unreached count, largest level
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + tid;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  unsigned long long key = 0, unreached = 0;
  U128 sumsq{0, 0};
  int32_t maxd = 0;
  for (int64_t i = gtid; i < a.g.n; i += gsize) {
    const int32_t d = ld_dev(&a.plevel[a.pmidx[i]]);
    a.level[i] = d;
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
      const unsigned long long kk = ((unsigned long long)(uint32_t)md << 32) |
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
  maxd = wave_max_i32(maxd);
  if (lane == 0) {
    if (a.min_dist && key) atomicMax(&a.kc->key, key);
    atomic_add_u128(a.kc->sumsq, sumsq);
    if (unreached) atomicAdd(&a.kc->unreached, unreached);
    if (maxd > 0) atomicMax(reinterpret_cast<unsigned long long *>(&a.kc->depth), (unsigned long long)maxd);
  }
}

// ------------------------------------------------- boundary-distance traversal
//
// The same parts, searched at two scales.  At context creation every part's
// boundary-to-boundary distances inside the part (paths that stay in the
// part) are computed once, as a byte matrix Dp (255: no such path).  A
// shortest path from the source to a boundary vertex is a chain of such
// in-part segments joined by edges between parts, so the boundary levels of
// the whole graph are the shortest-path distances of the small graph of
// boundary vertices whose edges are the parts' Dp entries and the edges
// between parts.  A traversal then has three steps:
//   1. the source's part: a BFS inside it from the source gives its
//      boundary vertices their first levels (lowered "from outside");
//   2. the queue of parts as in pbfs_kernel, but a run of part p takes only
//      its lowered boundary vertices (the seeds) and sets every boundary
//      level to min(own, seed level + Dp[seed][b]) — one min-plus pass over
//      rows of Dp instead of rounds of a BFS — then lowers the neighbours
//      across its external edges.  Dp obeys the triangle inequality, so a
//      vertex lowered by the pass needs no pass of its own: what it would
//      offer inside p its seed already offered;
//   3. when the queue is empty every boundary level is exact, and each part
//      computes its interior levels in one BFS inside LDS that starts each
//      boundary vertex at its exact level (and the source at 0): level by
//      level, a seed joining the frontier at its own level, so every vertex
//      is reached once and at its exact level.  The k-centers step is fused
//      into the write-out.
// Step 2 is still label-correcting (a part can run before all its seeds have
// their final levels), but a run costs one pass over |seeds| rows of Dp
// instead of tens of dependent rounds, and the interior of a part is
// traversed once per traversal instead of at every run.

constexpr unsigned char kDInf = 255;  // Dp entry: no path inside the part

// The part's compact image for in-part BFS: row offsets (nv + 1), adjacency
// (ne) and boundary list (nb), uint16, 16-byte aligned.
__host__ __device__ constexpr int64_t cblob_bytes(int64_t nv, int64_t ne, int64_t nb) {
  return align_up(2 * (nv + 1 + ne + nb), 16);
}
// LDS of part_bfs: the image, levels (int32 nv), two frontier lists (uint16,
// nv + 1 each), the seed list (uint16 vertices and int32 levels, nb + 2).
__host__ __device__ constexpr int64_t final_bytes(int64_t nv, int64_t ne, int64_t nb) {
  return cblob_bytes(nv, ne, nb) + 4 * nv + align_up(4 * (nv + 1), 4) + align_up(2 * (nb + 2), 4) + 4 * (nb + 2);
}
// LDS of part_table: levels (int32, nv rounded up to 16), the boundary
// levels, the entries' levels and rows (uint16).
__host__ __device__ constexpr int64_t table_bytes(int64_t nv, int64_t nb) {
  return 4 * align_up(nv, 16) + 12 * nb + 2 * align_up(2 * nb, 4) + 16;
}
// LDS of a multi-source interior pass (part_bfs_multi): the image, three
// 32-bit source masks per vertex, the seeds' relative levels (bytes, 32 per
// boundary vertex) and the sources' lowest levels in the part.
constexpr int kMultiW = 32;
__host__ __device__ constexpr int64_t mfinal_bytes(int64_t nv, int64_t ne, int64_t nb) {
  return cblob_bytes(nv, ne, nb) + 12 * nv + align_up(kMultiW * nb, 16) + 4 * kMultiW + 16;
}
// LDS of a run in step 2: boundary levels and seed levels (int32 nb each),
// seed indices (uint16 nb), changed flags (nb) and neighbouring-part flags (nn).
__host__ __device__ constexpr int64_t orun_bytes(int64_t nb, int64_t nn) {
  return 8 * nb + align_up(2 * nb, 4) + align_up(nb, 4) + nn + 16;
}
// Row stride of Dp: whole 16-byte groups.
__host__ __device__ constexpr int64_t drow(int64_t nb) { return align_up(nb, 16); }
// Bytes of a part's edge blob (nx external edges, nn neighbouring parts).
__host__ __device__ constexpr int64_t oedge_bytes(int64_t nx, int64_t nn) {
  return 2 * align_up(4 * nx, 16) + align_up(4 * nn, 16);
}

constexpr int kMaxSources = 64;  // sources of one multi-source search

struct ObfsArgs {
  DevGraph g;
  int32_t *level;     // output: levels (internal ids), n per source
  int32_t *min_dist;  // nullable: fused k-centers update
  KcResult *kc;
  int64_t k;
  const uint32_t *vstart, *pverts;  // part-major vertex ranges; part-major index -> internal id
  const uint32_t *bstart;           // boundary-major ranges of the parts' boundary vertices
  const uint32_t *xoff, *noff;      // external edges and neighbouring parts of each part
  // each part's edge blob (16-byte aligned sections): external edges'
  // sources (boundary index in the part) and neighbour slots (uint16), their
  // targets (boundary-major, uint32), the neighbouring parts (uint32)
  const uint4 *oblob;
  const uint32_t *oblob_off;  // per part, in 16-byte units
  uint32_t olds;              // dynamic LDS bytes of the kernel
  const uint4 *cblob;
  const uint32_t *cblob_off;  // per part, in 16-byte units
  const unsigned char *dmat;
  const unsigned long long *doff;  // per part, bytes
  // with din: every part's in-part distances from each boundary vertex to
  // each vertex (row per boundary vertex, nv bytes, kDInf: no path inside
  // the part), so that step 1 reads the source's column instead of a BFS
  const unsigned char *din;
  const unsigned long long *dinoff;  // per part, bytes
  // single-source search with tables: per boundary vertex, 1 when its level
  // came from outside the part (a seed) and no other boundary vertex of the
  // part gave it through the part (0: its min-plus did) -- written with the
  // level at every change, read by step 3 as the entry candidates
  int32_t *entf;
  int32_t *blevel;   // boundary-major: a part's boundary levels (written by its runs)
  int32_t *bilevel;  // boundary-major: levels lowered from outside (atomicMin)
  uint32_t *st, *ctl;
  unsigned long long *ring;
  uint32_t ring_n;
  // sources searched together (1 for a k-centers traversal): work items
  // are (source, part) pairs, item = source * k + part; per-source state
  // (blevel, bilevel: nbtot entries each; st: k entries; level column: n)
  // at the source's stride
  int nsrc;
  int64_t nbtot;
  uint32_t src_part[kMaxSources], src_loc[kMaxSources];
  // seeds at level 0: nseed == nsrc, seed s the source s; or, with one
  // search (nsrc == 1), several seeds at level 0 (the distance to the
  // nearest of them), at most two in a part.  The items queued first:
  // qitem[0 .. nq)
  int nseed, nq;
  uint32_t qitem[kMaxSources];
  // levels above bound are not pushed to other parts, and parts with no
  // entry are not written out: levels up to bound are exact, the others
  // above it or unwritten (kUnvisited: no bound)
  int32_t bound;
  // multi-source interiors: each vertex's arrival round per source of the
  // current group of 32 (bytes, part-major rows of 32; 255: not reached)
  unsigned char *mscr;
  // with PHDE_PBFS_STATS: [0..3] wait, seed, min-plus, push (ticks summed
  // over workgroups); [4] runs, [5] seeds, [6] level loads, [7, 8] first
  // and last tick, [9] interior time, [10] interior rounds, [11] boundary
  // levels lowered by the min-plus passes, [12] end of step 2;
  // [16..19] queued-to-taken latency; [32 + slot] fill ticks
  unsigned long long *stats;
};
constexpr int kCtlFin = 32;  // ctl word: next part of step 3

// lv[b] = min(lv[b], sl[i] + D[sj[i]][b]) over the ns seeds (D: nb rows of
// nbp bytes, kDInf = no path; in LDS or global memory); chg[b] = 1 where lv
// drops.  A thread takes one 16-column group of the rows of every R-th seed.
template <int Block>
__device__ __forceinline__ void minplus(const unsigned char *D, int nb, int nbp, int ns, const uint16_t *sj,
                                        const int32_t *sl, int32_t *lv, unsigned char *chg) {
  const int tid = threadIdx.x;
  const int G = nbp / 16;
  if (G <= Block) {
    const int R = Block / G;
    const int g = tid % G, r = tid / G;
    if (r >= R) return;
    int32_t mn[16];
#pragma unroll
    for (int c = 0; c < 16; c++) mn[c] = kUnvisited;
    for (int i0 = r; i0 < ns; i0 += 4 * R) {
      uint4 w[4];
      int32_t L[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i0 + q * R;
        const int ii = i < ns ? i : i0;
        w[q] = *reinterpret_cast<const uint4 *>(D + (size_t)sj[ii] * nbp + 16 * g);
        L[q] = sl[ii];
        if (i >= ns) w[q] = make_uint4(~0u, ~0u, ~0u, ~0u);  // past the last seed: no path
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t wq[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
#pragma unroll
        for (int c = 0; c < 16; c++) {
          const int32_t b = (int32_t)((wq[c / 4] >> (8 * (c % 4))) & 0xffu);
          const int32_t cand = b == kDInf ? kUnvisited : L[q] + b;
          mn[c] = cand < mn[c] ? cand : mn[c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const int col = 16 * g + c;
      if (col < nb && mn[c] < lv[col]) {
        const int32_t old = atomicMin(&lv[col], mn[c]);
        if (old > mn[c]) chg[col] = 1;
      }
    }
  } else {
    for (int64_t it = tid; it < (int64_t)ns * G; it += Block) {
      const int i = (int)(it / G), g = (int)(it % G);
      const uint4 w = *reinterpret_cast<const uint4 *>(D + (size_t)sj[i] * nbp + 16 * g);
      const uint32_t wq[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const int col = 16 * g + c;
        const int32_t b = (int32_t)((wq[c / 4] >> (8 * (c % 4))) & 0xffu);
        if (col < nb && b != kDInf) {
          const int32_t cand = sl[i] + b;
          if (cand < lv[col]) {
            const int32_t old = atomicMin(&lv[col], cand);
            if (old > cand) chg[col] = 1;
          }
        }
      }
    }
  }
}

__global__ void obfs_init_kernel(ObfsArgs a) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = gtid; j < a.nsrc * a.nbtot; j += gsize) {
    a.blevel[j] = kUnvisited;
    a.bilevel[j] = kUnvisited;
  }
  // the seeds' parts queued (items qitem[r] in slots r)
  for (int64_t t = gtid; t < a.nsrc * a.k; t += gsize) {
    bool q = false;
    for (int r = 0; r < a.nq; r++) q = q || a.qitem[r] == (uint32_t)t;
    a.st[t] = q ? 1u : 0u;
  }
  for (int64_t r = gtid; r < a.ring_n; r += gsize)
    a.ring[r * kSlotStride] = r < a.nq ? ((unsigned long long)(r + 1) << 32) | (a.qitem[r] + 1u) : 0ull;
  if (gtid < kCtlWords) a.ctl[gtid] = gtid == kCtlTP || gtid == kCtlTP + 1 ? (uint32_t)a.nq : 0u;  // tail, pending
  if (gtid == 0) {
    a.kc->key = 0;
    a.kc->sumsq[0] = a.kc->sumsq[1] = 0;
    a.kc->unreached = 0;
    a.kc->depth = 0;
  }
}

// Copies part p's compact image into LDS (16-byte loads, eight in flight
// per thread); returns its size in 16-byte words.
template <int Block>
__device__ __forceinline__ void load_cblob(const uint4 *cblob, const uint32_t *cblob_off, uint32_t p,
                                           unsigned char *smem) {
  const uint4 *gb = cblob + cblob_off[p];
  const int nblob = (int)(cblob_off[p + 1] - cblob_off[p]);
  uint4 *sb = reinterpret_cast<uint4 *>(smem);
  for (int i = threadIdx.x; i < nblob; i += 8 * Block) {
    uint4 t[8];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = gb[i + j * Block < nblob ? i + j * Block : i];
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (i + j * Block < nblob) sb[i + j * Block] = t[j];
  }
}

// Running totals of the fused k-centers step over the vertices one thread
// wrote out.
struct KcAcc {
  unsigned long long key = 0, unreached = 0;
  U128 sumsq{0, 0};
  int32_t maxd = 0;
};

// Write-out of part p's levels (LDS, nv entries) with the fused k-centers
// step: min_dist = min(min_dist, level), the first maximiser (caller's id
// on ties), exact sum of squares, unreached count, largest level.
template <int Block>
__device__ __forceinline__ void part_write_out(const ObfsArgs &a, uint32_t v0, int nv, const int32_t *lvl, int src,
                                               KcAcc &acc) {
  for (int v = threadIdx.x; v < nv; v += Block) {
    const int32_t d = lvl[v];
    const uint32_t i = a.pverts[v0 + v];
    a.level[src * a.g.n + i] = d;
    if (d == kUnvisited) {
      acc.unreached++;
    } else {
      u128_add(acc.sumsq, (unsigned long long)d * (unsigned long long)d);
      acc.maxd = d > acc.maxd ? d : acc.maxd;
    }
    if (a.min_dist) {
      int32_t md = a.min_dist[i];
      if (d < md) {
        md = d;
        a.min_dist[i] = md;
      }
      const unsigned long long kk = ((unsigned long long)(uint32_t)md << 32) |
                                    (0xffffffffull - (unsigned long long)original_id(a.g, i));
      acc.key = kk > acc.key ? kk : acc.key;
    }
  }
}

// Step 3 for a part without the source, from the in-part distance tables
// (single-source search): the last stretch of a shortest path to v inside
// the part starts at one of its boundary vertices, so level(v) = min over
// the part's boundary b of L_b + Din[b][v], with L_b the exact boundary
// levels of step 2.  Only entries count: a boundary vertex j with
// L_i + Dp[i][j] <= L_j for another i adds nothing, since
// Din[i][v] <= Dp[i][j] + Din[j][v].  The entries' rows are read 16
// vertices per load, each thread taking one 16-vertex group and every G-th
// entry, the groups combined with LDS atomicMin; no rounds, no barrier per
// level.
template <int Block>
__device__ void part_table(const ObfsArgs &a, uint32_t p, unsigned char *smem, int *s_cnt, KcAcc &acc) {
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  // with stats: [20] boundary levels and entries, [21] the entries' rows,
  // [23] write-out, [24] entries
  const bool stats = a.stats != nullptr;
  unsigned long long tp0 = 0, tp1 = 0;
  if (stats && tid == 0) tp0 = __builtin_amdgcn_s_memrealtime();
  auto phase = [&](int slot) {
    if (stats && tid == 0) {
      tp1 = __builtin_amdgcn_s_memrealtime();
      atomicAdd(&a.stats[slot], tp1 - tp0);
      tp0 = tp1;
    }
  };
  const uint32_t v0 = a.vstart[p];
  const int nv = (int)(a.vstart[p + 1] - v0);
  const int nvp = (int)align_up(nv, 16);
  const uint32_t b0 = a.bstart[p];
  const int nb = (int)(a.bstart[p + 1] - b0);
  const int nbp = (int)drow(nb);
  int32_t *lvl = reinterpret_cast<int32_t *>(smem);
  int32_t *L = lvl + nvp;  // boundary levels
  int32_t *Lc = L + nb;    // the candidates' levels (kUnvisited elsewhere)
  int32_t *el = Lc + nb;   // flags, then the entries' levels
  uint16_t *ej = reinterpret_cast<uint16_t *>(el + nb);                          // the entries
  uint16_t *cj = reinterpret_cast<uint16_t *>(reinterpret_cast<unsigned char *>(ej) + align_up(2 * nb, 4));
  if (tid == 0) s_cnt[0] = s_cnt[1] = 0;
  for (int v = tid; v < nvp; v += Block) lvl[v] = kUnvisited;
  // candidates: the boundary vertices step 2 left flagged as reached from
  // outside (a superset of the entries: ties are not flagged off)
  for (int jj = 0; jj < nb; jj += Block) {
    const int j = jj + tid;
    int32_t Lj = kUnvisited;
    bool cand = false;
    if (j < nb) {
      Lj = ld_dev(&a.blevel[b0 + j]);
      // (a candidate above the bound gives no level within it, nor does
      // any candidate it dominates)
      cand = Lj != kUnvisited && Lj <= a.bound && ld_dev(&a.entf[b0 + j]) != 0;
      L[j] = Lj;
      Lc[j] = cand ? Lj : kUnvisited;
      el[j] = cand ? Lj : kUnvisited;
    }
    const unsigned long long m = __ballot(cand);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&s_cnt[1], __popcll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (cand) cj[base + __popcll(m & lanemask_lt())] = (uint16_t)j;
    }
  }
  __syncthreads();
  const int nc = s_cnt[1];
  // the candidates dominated by another candidate (enough: a vertex
  // dominated at all is dominated by an entry, by transitivity): each
  // (candidate, 16-column group of its Dp row) an item of its own, the
  // loads issued together (Dp is symmetric: row j holds the distances to j)
  const unsigned char *D = a.dmat + a.doff[p];
  const int ng = nbp / 16;
  constexpr int kItems = 8;
  if (stats && tid == 0) atomicAdd(&a.stats[25], (unsigned long long)nc);
  for (int it0 = tid; it0 < nc * ng; it0 += kItems * Block) {
    uint4 w[kItems];
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      const int it = it0 + q * Block;
      const int j = it < nc * ng ? cj[it / ng] : 0, g = it < nc * ng ? it % ng : 0;
      w[q] = reinterpret_cast<const uint4 *>(D + (size_t)j * nbp)[g];
    }
#pragma unroll
    for (int q = 0; q < kItems; q++) {
      const int it = it0 + q * Block;
      if (it >= nc * ng) break;
      const int j = cj[it / ng], g = it % ng;
      const int32_t Lj = L[j];
      const uint32_t wq[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
      bool dom = false;
#pragma unroll
      for (int c = 0; c < 16; c++) {
        const int i = 16 * g + c;
        const int32_t d = (int32_t)((wq[c / 4] >> (8 * (c % 4))) & 0xffu);
        if (i < nb && i != j && d != kDInf) {
          const int32_t Li = Lc[i];
          dom = dom || (Li != kUnvisited && Li + d <= Lj);
        }
      }
      if (dom) el[j] = kUnvisited;
    }
  }
  __syncthreads();
  // the entries, listed
  for (int jj = 0; jj < nb; jj += Block) {
    const int j = jj + tid;
    const int32_t Lj = j < nb ? el[j] : kUnvisited;
    const bool entry = Lj != kUnvisited;
    __syncthreads();  // every flag read before the list overwrites el
    const unsigned long long m = __ballot(entry);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&s_cnt[0], __popcll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (entry) {
        const int at = base + __popcll(m & lanemask_lt());
        ej[at] = (uint16_t)j;
        el[at] = Lj;
      }
    }
  }
  __syncthreads();
  const int ne = s_cnt[0];
  if (ne == 0 && a.bound != kUnvisited) return;  // nothing within the bound
  phase(20);
  if (stats && tid == 0) atomicAdd(&a.stats[24], (unsigned long long)ne);
  const unsigned char *Din = a.din + a.dinoff[p];
  const int nw = nvp / 16;
  const int G = nw < Block ? Block / nw : 1;
  for (int t = tid; t < nw * G; t += Block) {
    const int w = t % nw, g = t / nw;
    int32_t mn[16];
#pragma unroll
    for (int c = 0; c < 16; c++) mn[c] = kUnvisited;
    // kRowsAt entries' rows in flight together
    constexpr int kRowsAt = 8;
    for (int e0 = g; e0 < ne; e0 += kRowsAt * G) {
      uint4 q[kRowsAt];
      int32_t Le[kRowsAt];
#pragma unroll
      for (int r = 0; r < kRowsAt; r++) {
        const int e = e0 + r * G;
        const int ee = e < ne ? e : e0;
        q[r] = *reinterpret_cast<const uint4 *>(Din + (size_t)ej[ee] * nvp + 16 * w);
        Le[r] = e < ne ? el[ee] : kUnvisited;
      }
#pragma unroll
      for (int r = 0; r < kRowsAt; r++) {
        const uint32_t qq[4] = {q[r].x, q[r].y, q[r].z, q[r].w};
#pragma unroll
        for (int c = 0; c < 16; c++) {
          const int32_t d = (int32_t)((qq[c / 4] >> (8 * (c % 4))) & 0xffu);
          const int32_t cand = d == kDInf || Le[r] == kUnvisited ? kUnvisited : Le[r] + d;
          mn[c] = cand < mn[c] ? cand : mn[c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (mn[c] != kUnvisited) atomicMin(&lvl[16 * w + c], mn[c]);
  }
  __syncthreads();
  phase(21);
  part_write_out<Block>(a, v0, nv, lvl, 0, acc);
  if (stats) {
    __syncthreads();
    phase(23);
  }
}

// BFS inside part p (steps 1 and 3).  Source mode (final == false): from
// the source alone; the levels of the part's boundary vertices go to
// bilevel.  Final mode: every boundary vertex starts at its exact level
// (blevel) and the source, when in p, at 0; the levels are written out
// (internal ids) with the fused k-centers step.  Level by level from the
// lowest seed level: a seed joins the frontier at its own level, and a
// level with an empty frontier jumps to the next seed level, so every vertex
// is listed once.  Returns the rounds taken.
template <int Block, int Lanes, int Claims>
__device__ int part_bfs(const ObfsArgs &a, uint32_t p, int src, bool final_mode, unsigned char *smem, int *s_cnt,
                        int *s_min, int32_t *s_dummy, KcAcc &acc) {
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const uint32_t v0 = a.vstart[p];
  const int nv = (int)(a.vstart[p + 1] - v0);
  const uint32_t b0 = a.bstart[p];
  const int nb = (int)(a.bstart[p + 1] - b0);
  const bool stats = a.stats != nullptr && final_mode;
  unsigned long long tp0 = 0, tp1 = 0;
  if (stats && tid == 0) tp0 = __builtin_amdgcn_s_memrealtime();
  load_cblob<Block>(a.cblob, a.cblob_off, p, smem);
  if (tid == 0) {
    s_cnt[0] = s_cnt[1] = s_cnt[2] = 0;
    *s_min = kUnvisited;
  }
  __syncthreads();
  // with stats: [20] image load, [21] seeds, [22] rounds, [23] write-out
  auto phase = [&](int slot) {
    if (stats && tid == 0) {
      tp1 = __builtin_amdgcn_s_memrealtime();
      atomicAdd(&a.stats[slot], tp1 - tp0);
      tp0 = tp1;
    }
  };
  phase(20);
  const uint16_t *lrow = reinterpret_cast<const uint16_t *>(smem);
  const int ne = lrow[nv];
  const uint16_t *ladj = lrow + nv + 1;
  const uint16_t *blist = ladj + ne;
  int32_t *lvl = reinterpret_cast<int32_t *>(smem + cblob_bytes(nv, ne, nb));
  uint16_t *F0 = reinterpret_cast<uint16_t *>(lvl + nv);
  uint16_t *F1 = F0 + nv + 1;
  uint16_t *sv = reinterpret_cast<uint16_t *>(smem + cblob_bytes(nv, ne, nb) + 4 * nv + align_up(4 * (nv + 1), 4));
  int32_t *sl = reinterpret_cast<int32_t *>(reinterpret_cast<unsigned char *>(sv) + align_up(2 * (nb + 2), 4));
  for (int v = tid; v < nv; v += Block) lvl[v] = kUnvisited;
  __syncthreads();
  // seeds: the boundary at its exact levels (final mode), the source (one
  // search with several seeds: every seed in p)
  const int nsd = final_mode && a.nsrc == 1 ? a.nseed : 1;
  for (int jj = 0; jj < nb + nsd; jj += Block) {
    const int j = jj + tid;
    int u = 0;
    int32_t L = kUnvisited;
    if (j < nb) {
      if (final_mode) {
        u = blist[j];
        L = ld_dev(&a.blevel[src * a.nbtot + b0 + j]);
      }
    } else if (j < nb + nsd) {
      const int t = nsd > 1 ? j - nb : src;
      if (p == a.src_part[t]) {
        u = (int)a.src_loc[t];
        L = 0;
      }
    }
    const bool add = L != kUnvisited;
    if (add) {
      atomicMin(&lvl[u], L);
      atomicMin(s_min, L);
    }
    const unsigned long long m = __ballot(add);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&s_cnt[2], __popcll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (add) {
        const int at = base + __popcll(m & lanemask_lt());
        sv[at] = (uint16_t)u;
        sl[at] = L;
      }
    }
  }
  __syncthreads();
  const int ns = s_cnt[2];
  int32_t Lc = *s_min;
  __syncthreads();  // every thread has read s_min before it is reused
  int rounds = 0;
  // lists F0 / F1 alternate; their counts rotate over s_cnt[0..2] so that
  // the count the next round appends to is cleared in this one
  int li = 0, ci = 0;
  if (ns > 0) {
    phase(21);
    if (tid == 0) s_cnt[0] = s_cnt[1] = 0;
    __syncthreads();
    // the seeds at level Lc into list li (count ci); a seed that a shorter
    // path reached (impossible for exact seeds) is not listed again
    auto join = [&](int32_t level, uint16_t *F, int *cnt) {
      for (int ii = 0; ii < ns; ii += Block) {
        const int i = ii + tid;
        bool add = false;
        int u = 0;
        if (i < ns && sl[i] == level) {
          u = sv[i];
          add = lvl[u] == level;
        }
        const unsigned long long m = __ballot(add);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          int base = 0;
          if (lane == leader) base = atomicAdd(cnt, __popcll(m));
          base = __builtin_amdgcn_readlane(base, leader);
          if (add) F[base + __popcll(m & lanemask_lt())] = (uint16_t)u;
        }
      }
    };
    join(Lc, F0, &s_cnt[0]);
    __syncthreads();
    for (;;) {
      uint16_t *Fc = li ? F1 : F0;
      uint16_t *Fn = li ? F0 : F1;
      const int nf = s_cnt[ci];
      const int cn_i = (ci + 1) % 3;
      if (nf == 0) {
        // nothing at this level: jump to the next seed level, if any
        if (tid == 0) *s_min = kUnvisited;
        __syncthreads();
        for (int i = tid; i < ns; i += Block)
          if (sl[i] > Lc) atomicMin(s_min, sl[i]);
        __syncthreads();
        const int32_t nL = *s_min;
        if (nL == kUnvisited) break;
        Lc = nL;
        join(Lc, Fc, &s_cnt[ci]);
        __syncthreads();
        continue;
      }
      rounds++;
      const int32_t nl = Lc + 1;
      int *cn = &s_cnt[cn_i];
      if (tid == 0) s_cnt[(ci + 2) % 3] = 0;
      // the seeds of the next level join the next list first; the claims
      // below never list an exact seed (it already holds nl)
      join(nl, Fn, cn);
      for (int f0 = 0; f0 < nf; f0 += Block / Lanes) {
        const int f = f0 + tid / Lanes;
        int xb = 0, e = 0;
        if (f < nf) {
          const int u = Fc[f];
          xb = lrow[u];
          e = lrow[u + 1];
        }
        xb += tid % Lanes;
        while (__ballot(xb < e)) {
          int w[Claims];
          int32_t old[Claims];
          bool ok[Claims];
#pragma unroll
          for (int q = 0; q < Claims; q++) {
            const int x = xb + q * Lanes;
            ok[q] = x < e;
            w[q] = ladj[ok[q] ? x : 0];
          }
#pragma unroll
          for (int q = 0; q < Claims; q++) old[q] = atomicMin(ok[q] ? &lvl[w[q]] : &s_dummy[tid], nl);
          unsigned long long mq[Claims];
          int pre[Claims], total = 0;
#pragma unroll
          for (int q = 0; q < Claims; q++) {
            mq[q] = __ballot(ok[q] && old[q] > nl);
            pre[q] = total;
            total += __popcll(mq[q]);
          }
          if (total) {
            int base = 0;
            if (lane == 0) base = atomicAdd(cn, total);
            base = __builtin_amdgcn_readfirstlane(base);
            const unsigned long long lt = lanemask_lt();
#pragma unroll
            for (int q = 0; q < Claims; q++)
              if ((mq[q] >> lane) & 1ull) Fn[base + pre[q] + __popcll(mq[q] & lt)] = (uint16_t)w[q];
          }
          xb += Claims * Lanes;
        }
      }
      __syncthreads();
      li ^= 1;
      ci = cn_i;
      Lc = nl;
    }
  }
  phase(22);
  if (!final_mode) {
    for (int j = tid; j < nb; j += Block) {
      const int32_t L = lvl[blist[j]];
      if (L != kUnvisited) st_dev(&a.bilevel[src * a.nbtot + b0 + j], L);
    }
    return rounds;
  }
  part_write_out<Block>(a, v0, nv, lvl, src, acc);
  if (stats) {
    __syncthreads();
    phase(23);
  }
  return rounds;
}

// Step 3 of a multi-source search for sources s0 .. s0 + w - 1 (w <= 32) of
// part p, all at once: a bit-parallel BFS inside the part (bit i of a
// vertex's mask: reached from source s0 + i).  Levels are relative to each
// source's lowest level in the part (lmin[i]: its lowest boundary level, or
// 0 in the source's own part); round r holds the vertices at level
// lmin[i] + r for source i, and a boundary vertex joins source i's frontier
// at its own relative level (exact, so it is never reached earlier).  Each
// vertex's arrival round per source goes to a byte row of mscr written and
// read back by the thread that owns the vertex; the write-out fills the w
// level columns.  The part's image must already be in LDS.  Returns false
// (nothing written) when a relative level or a round reaches 255: the
// caller then takes the sources one at a time.
template <int Block>
__device__ bool part_bfs_multi(const ObfsArgs &a, uint32_t p, int s0, int w, unsigned char *smem, int *s_mf) {
  const int tid = threadIdx.x;
  const uint32_t v0 = a.vstart[p];
  const int nv = (int)(a.vstart[p + 1] - v0);
  const uint32_t b0 = a.bstart[p];
  const int nb = (int)(a.bstart[p + 1] - b0);
  const uint16_t *lrow = reinterpret_cast<const uint16_t *>(smem);
  const int ne = lrow[nv];
  const uint16_t *ladj = lrow + nv + 1;
  const uint16_t *blist = ladj + ne;
  uint32_t *vis = reinterpret_cast<uint32_t *>(smem + cblob_bytes(nv, ne, nb));
  uint32_t *fr = vis + nv;
  uint32_t *nx = fr + nv;
  unsigned char *rel = reinterpret_cast<unsigned char *>(nx + nv);
  int32_t *lmin = reinterpret_cast<int32_t *>(rel + align_up(kMultiW * nb, 16));
  unsigned char *r8 = a.mscr + (size_t)v0 * kMultiW;
  for (int v = tid; v < nv; v += Block) {
    vis[v] = fr[v] = nx[v] = 0;
    uint4 *row = reinterpret_cast<uint4 *>(r8 + (size_t)v * kMultiW);
    row[0] = row[1] = make_uint4(~0u, ~0u, ~0u, ~0u);
  }
  if (tid < kMultiW) lmin[tid] = tid < w && a.src_part[s0 + tid] == p ? 0 : kUnvisited;
  if (tid == 0) s_mf[0] = s_mf[1] = s_mf[2] = s_mf[3] = 0;
  __syncthreads();
  // each source's lowest boundary level in the part
  for (int idx = tid; idx < nb * w; idx += Block) {
    const int i = idx / nb, b = idx % nb;
    const int32_t L = ld_dev(&a.blevel[(s0 + i) * a.nbtot + b0 + b]);
    if (L != kUnvisited) atomicMin(&lmin[i], L);
  }
  __syncthreads();
  // the seeds' relative levels (255: not a seed of that source)
  for (int idx = tid; idx < nb * kMultiW; idx += Block) {
    const int i = idx / nb, b = idx % nb;
    int r = 255;
    if (i < w) {
      const int32_t L = ld_dev(&a.blevel[(s0 + i) * a.nbtot + b0 + b]);
      if (L != kUnvisited) {
        const int32_t d = L - lmin[i];
        if (d >= 255)
          s_mf[0] = 1;
        else {
          r = d;
          atomicMax(&s_mf[1], d);
        }
      }
    }
    rel[b * kMultiW + i] = (unsigned char)r;
  }
  __syncthreads();
  if (s_mf[0]) return false;
  const int rmax = s_mf[1];
  const uint32_t *rel32 = reinterpret_cast<const uint32_t *>(rel);
  for (int r = 0;; r++) {
    // the seeds of round r join (and the sources in their own part at 0)
    for (int idx = tid; idx < nb * (kMultiW / 4); idx += Block) {
      const int b = idx / (kMultiW / 4), q = idx % (kMultiW / 4);
      const uint32_t word = rel32[idx];
      uint32_t bits = 0;
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (((word >> (8 * c)) & 0xffu) == (uint32_t)r) bits |= 1u << (4 * q + c);
      if (bits) {
        const int v = blist[b];
        atomicOr(&fr[v], bits);
        atomicOr(&vis[v], bits);
      }
    }
    if (r == 0 && tid < w && a.src_part[s0 + tid] == p) {
      const int v = (int)a.src_loc[s0 + tid];
      atomicOr(&fr[v], 1u << tid);
      atomicOr(&vis[v], 1u << tid);
    }
    __syncthreads();
    // the round's frontier: arrival rounds recorded, masks offered to the
    // neighbours
    for (int v = tid; v < nv; v += Block) {
      const uint32_t f = fr[v];
      if (!f) continue;
      uint32_t bits = f;
      while (bits) {
        const int i = __ffs(bits) - 1;
        bits &= bits - 1u;
        r8[(size_t)v * kMultiW + i] = (unsigned char)r;
      }
      const int e = lrow[v + 1];
      for (int x = lrow[v]; x < e; x++) atomicOr(&nx[ladj[x]], f);
    }
    __syncthreads();
    for (int v = tid; v < nv; v += Block) {
      const uint32_t nw = nx[v] & ~vis[v];
      fr[v] = nw;
      vis[v] |= nw;
      nx[v] = 0;
      if (nw) s_mf[2 + (r & 1)] = 1;
    }
    if (tid == 0) s_mf[2 + ((r + 1) & 1)] = 0;
    __syncthreads();
    if (!s_mf[2 + (r & 1)] && r >= rmax) break;
    // a frontier that would reach round 255 has no byte for its rounds
    // (nothing is written yet: the caller takes the sources one at a time)
    if (r + 1 >= 255) return false;
  }
  // write-out: each thread reads back its own vertices' rows
  for (int v = tid; v < nv; v += Block) {
    const uint32_t i_int = a.pverts[v0 + v];
    const uint4 *row = reinterpret_cast<const uint4 *>(r8 + (size_t)v * kMultiW);
    const uint4 q0 = row[0], q1 = row[1];
    const uint32_t wd[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
#pragma unroll
    for (int i = 0; i < kMultiW; i++)
      if (i < w) {
        const uint32_t d8 = (wd[i / 4] >> (8 * (i % 4))) & 0xffu;
        a.level[(s0 + i) * a.g.n + i_int] = d8 == 255u ? kUnvisited : lmin[i] + (int32_t)d8;
      }
  }
  __syncthreads();
  return true;
}

// Step 1: one workgroup, the source's part.
template <int Block, int Lanes, int Claims>
__global__ __launch_bounds__(Block) void obfs_source_kernel(ObfsArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ int s_cnt[3], s_min;
  __shared__ int32_t s_dummy[Block];
  const int t = blockIdx.x;  // the seed
  const int src = a.nsrc > 1 ? t : 0;
  const uint32_t p = a.src_part[t];
  if (a.din) {
    // the seed's column of the part's distance table: its in-part
    // distance to every boundary vertex (symmetric graph); seeds sharing
    // a part keep the least
    const uint32_t nvp = (uint32_t)align_up(a.vstart[p + 1] - a.vstart[p], 16);
    const uint32_t b0 = a.bstart[p];
    const int nb = (int)(a.bstart[p + 1] - b0);
    const unsigned char *col = a.din + a.dinoff[p] + a.src_loc[t];
    for (int j = threadIdx.x; j < nb; j += Block) {
      const unsigned char d = col[(size_t)j * nvp];
      if (d != kDInf) atomicMin(&a.bilevel[src * a.nbtot + b0 + j], (int32_t)d);
    }
    return;
  }
  KcAcc acc;
  part_bfs<Block, Lanes, Claims>(a, p, src, false, smem, s_cnt, &s_min, s_dummy, acc);
}

// Steps 2 and 3: every workgroup resident; the queue of parts, then the
// parts' interiors.
template <int Block, int Lanes, int Claims>
__global__ __launch_bounds__(Block) void obfs_kernel(ObfsArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ uint32_t s_part, s_ver;
  __shared__ int s_ns, s_rerun, s_cnt[3], s_min;
  __shared__ int32_t s_dummy[Block];
  const int tid = threadIdx.x;
  const int lane = tid % kWave;
  const bool stats = a.stats != nullptr;
  // the item whose levels are in LDS (and its version), and the part whose
  // edge blob is
  uint32_t cached = kNoPart, cached_ver = 0, cached_part = kNoPart;
  // thread 0's timings with stats, kept in LDS rather than in registers
  __shared__ unsigned long long s_tm[11];
  unsigned long long &t_wait = s_tm[0], &t_seed = s_tm[1], &t_mp = s_tm[2], &t_push = s_tm[3], &n_act = s_tm[4],
                     &n_seed = s_tm[5], &n_img = s_tm[6], &t_first = s_tm[7], &t_last = s_tm[8], &t0 = s_tm[9],
                     &t1 = s_tm[10];
  if (tid == 0) {
    for (int i = 0; i < 11; i++) s_tm[i] = 0;
    t_first = ~0ull;
  }
  __syncthreads();
  for (;;) {
    if (stats && tid == 0) t0 = __builtin_amdgcn_s_memrealtime();
    if (tid == 0) {
      const uint32_t it = pop_part(a);
      if (it != kNoPart) {
        s_ver = atomicAdd(&a.st[it], 1u) >> 2;  // queued -> running; the item's version
        __builtin_amdgcn_s_waitcnt(0);
      }
      s_part = it;
    }
    __syncthreads();
    const uint32_t item = s_part;
    if (stats && tid == 0) {
      t1 = __builtin_amdgcn_s_memrealtime();
      t_wait += t1 - t0;
      if (item != kNoPart && t1 < t_first) t_first = t1;
    }
    if (item == kNoPart) break;
    const uint32_t src = item / (uint32_t)a.k, p = item % (uint32_t)a.k;
    int32_t *blev = a.blevel + src * a.nbtot;
    int32_t *bilev = a.bilevel + src * a.nbtot;
    const uint32_t ibase = src * (uint32_t)a.k;  // this source's first item
    const uint32_t b0 = a.bstart[p];
    const int nb = (int)(a.bstart[p + 1] - b0);
    const int nbp = (int)drow(nb);
    const int nx = (int)(a.xoff[p + 1] - a.xoff[p]);
    const int nn = (int)(a.noff[p + 1] - a.noff[p]);
    // LDS: the run's state, then the part's image (its edge blob); Dp is
    // read from global memory, only the seeds' rows (measured: copying the
    // whole Dp into LDS when a part is taken, to serve its re-runs, made
    // the first run -- the one on the wave's path -- slower, 1.15 -> 1.19 ms
    // per traversal of RGG n=4.2M)
    const int ob = (int)align_up(orun_bytes(nb, nn), 16);
    const int eb = (int)oedge_bytes(nx, nn);
    int32_t *lv = reinterpret_cast<int32_t *>(smem);
    int32_t *sl = lv + nb;
    uint16_t *sj = reinterpret_cast<uint16_t *>(sl + nb);
    unsigned char *chg = reinterpret_cast<unsigned char *>(smem + 8 * nb + align_up(2 * nb, 4));
    unsigned char *nflag = chg + align_up(nb, 4);
    const uint16_t *xs = reinterpret_cast<const uint16_t *>(smem + ob);
    const uint16_t *xn = xs + nx;
    const uint32_t *xd = reinterpret_cast<const uint32_t *>(smem + ob + align_up(4 * nx, 16));
    const uint32_t *nbl = reinterpret_cast<const uint32_t *>(smem + ob + 2 * align_up(4 * nx, 16));
    for (;;) {
      if (stats && tid == 0) {
        t0 = __builtin_amdgcn_s_memrealtime();
        n_act++;
      }
      const bool fresh = cached != item || cached_ver != s_ver;
      const bool fresh_blob = cached_part != p;
      // every load of the run in flight together: the levels lowered from
      // outside, and for a fresh image the part's levels and edge blob
      constexpr int kR = 4;
      const bool few = nb <= kR * Block;
      int32_t il[kR], ol[kR];
      if (few) {
#pragma unroll
        for (int q = 0; q < kR; q++) {
          const int j = tid + q * Block;
          il[q] = j < nb ? ld_dev(&bilev[b0 + j]) : kUnvisited;
          ol[q] = fresh && j < nb ? ld_dev(&blev[b0 + j]) : 0;
        }
      }
      if (fresh_blob) {
        const int we = eb / 16;
        const uint4 *ge = a.oblob + a.oblob_off[p];
        uint4 *se = reinterpret_cast<uint4 *>(smem + ob);
        for (int i = tid; i < we; i += 8 * Block) {
          uint4 t[8];
#pragma unroll
          for (int j = 0; j < 8; j++) t[j] = ge[i + j * Block < we ? i + j * Block : i];
#pragma unroll
          for (int j = 0; j < 8; j++)
            if (i + j * Block < we) se[i + j * Block] = t[j];
        }
      }
      if (fresh) {
        if (few) {
#pragma unroll
          for (int q = 0; q < kR; q++)
            if (tid + q * Block < nb) lv[tid + q * Block] = ol[q];
        } else {
          for (int j = tid; j < nb; j += Block) lv[j] = ld_dev(&blev[b0 + j]);
        }
        if (stats && tid == 0) n_img++;
      }
      for (int j = tid; j < nb; j += Block) chg[j] = 0;
      for (int i = tid; i < nn; i += Block) nflag[i] = 0;
      if (tid == 0) s_ns = 0;
      __syncthreads();
      // seeds: boundary vertices lowered from outside
      auto seed = [&](int j, int32_t v, bool in) {
        bool add = false;
        if (in && v < lv[j]) {
          lv[j] = v;
          chg[j] = 2;  // from outside (the min-plus below sets 1 if it lowers j further)
          add = true;
        }
        const unsigned long long m = __ballot(add);
        if (m) {
          const int leader = __ffsll((long long)m) - 1;
          int base = 0;
          if (lane == leader) base = atomicAdd(&s_ns, __popcll(m));
          base = __builtin_amdgcn_readlane(base, leader);
          if (add) {
            const int at = base + __popcll(m & lanemask_lt());
            sj[at] = (uint16_t)j;
            sl[at] = v;
          }
        }
      };
      if (few) {
#pragma unroll
        for (int q = 0; q < kR; q++) {
          const int j = tid + q * Block;
          seed(j < nb ? j : 0, il[q], j < nb);
        }
      } else {
        for (int jj = 0; jj < nb; jj += Block) {
          const int j = jj + tid;
          seed(j < nb ? j : 0, j < nb ? ld_dev(&bilev[b0 + j]) : kUnvisited, j < nb);
        }
      }
      __syncthreads();
      const int ns = s_ns;
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_seed += t1 - t0;
        t0 = t1;
        n_seed += ns;
      }
      // min-plus: lv[b] = min(lv[b], sl[i] + Dp[sj[i]][b]) over the seeds
      if (ns > 0) minplus<Block>(a.dmat + a.doff[p], nb, nbp, ns, sj, sl, lv, chg);
      __syncthreads();
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_mp += t1 - t0;
        t0 = t1;
      }
      // write back the lowered levels, lower the neighbours across the
      // external edges (four in flight per thread) and flag their parts
      for (int j = tid; j < nb; j += Block)
        if (chg[j]) {
          st_dev(&blev[b0 + j], lv[j]);
          if (a.entf && a.nsrc == 1) st_dev(&a.entf[b0 + j], chg[j] == 2 ? 1 : 0);
          if (stats) atomicAdd(&a.stats[11], 1ull);
        }
      for (int xx = 0; xx < nx; xx += 4 * Block) {
        int32_t nl[4], old[4];
        bool on[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int x = xx + q * Block + tid;
          on[q] = false;
          nl[q] = 0;
          if (x < nx) {
            const int j = xs[x];
            nl[q] = lv[j] + 1;
            on[q] = chg[j] != 0 && nl[q] <= a.bound;
          }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) old[q] = on[q] ? atomicMin(&bilev[xd[xx + q * Block + tid]], nl[q]) : 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (on[q] && old[q] > nl[q]) nflag[xn[xx + q * Block + tid]] = 1;
      }
      __syncthreads();
      // queue the flagged parts (measured: a run that kept one newly
      // queued neighbour for its own workgroup instead of the ring ran the
      // wave's front out of order and re-ran parts more, 1.15 -> 1.42 ms)
      for (int i = tid; i < nn; i += Block)
        if (nflag[i]) mark_lowered(a, ibase + nbl[i]);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (stats && tid == 0) {
        t1 = __builtin_amdgcn_s_memrealtime();
        t_push += t1 - t0;
        if (t1 > t_last) t_last = t1;
      }
      if (tid == 0) {
        const uint32_t v = s_ver;
        s_ver = v + 1u;
        const uint32_t old = atomicCAS(&a.st[item], v << 2 | 2u, (v + 1u) << 2);
        if (old == (v << 2 | 2u)) {
          s_rerun = 0;
          const unsigned long long tp = atomicSub(tp_word(a.ctl), 1ull);
          if ((uint32_t)tp == 1u) mark_done(a, (uint32_t)(tp >> 32));  // the last part: step 2 is over
        } else {
          // lowered again while it ran: run it once more
          atomicExch(&a.st[item], (v + 1u) << 2 | 2u);
          __builtin_amdgcn_s_waitcnt(0);
          s_rerun = 1;
        }
      }
      __syncthreads();
      cached = item;
      cached_ver = s_ver;
      cached_part = p;
      if (!s_rerun) break;
    }
  }
  if (stats && tid == 0) {
    atomicAdd(&a.stats[0], t_wait);
    atomicAdd(&a.stats[1], t_seed);
    atomicAdd(&a.stats[2], t_mp);
    atomicAdd(&a.stats[3], t_push);
    atomicAdd(&a.stats[4], n_act);
    atomicAdd(&a.stats[5], n_seed);
    atomicAdd(&a.stats[6], n_img);
    atomicMin(&a.stats[7], t_first);
    atomicMax(&a.stats[8], t_last);
    atomicMax(&a.stats[12], __builtin_amdgcn_s_memrealtime());
  }
  // the queue is empty and every run has completed: the boundary levels
  // are exact (unless an error stopped the search; the host then repeats
  // the traversal another way)
  // a multi-source search fills its interiors in obfs_multi_final_kernel
  if (a.nsrc > 1 || ld_dev(&a.ctl[kCtlErr])) return;
  KcAcc acc;
  unsigned long long t_fin = 0, n_fr = 0;
  if (stats && tid == 0) t0 = __builtin_amdgcn_s_memrealtime();
  {
    // each part's successor is claimed while the part runs
    uint32_t next = 0;
    if (tid == 0) next = atomicAdd(&a.ctl[kCtlFin], 1u);
    for (;;) {
      if (tid == 0) {
        s_part = next;
        if ((int64_t)next < a.k) next = atomicAdd(&a.ctl[kCtlFin], 1u);
      }
      __syncthreads();
      const uint32_t p = s_part;
      if ((int64_t)p >= a.k) break;
      bool seeded = false;
      for (int t = 0; t < a.nseed; t++) seeded = seeded || p == a.src_part[t];
      if (a.din && !seeded)
        part_table<Block>(a, p, smem, s_cnt, acc);
      else
        n_fr += part_bfs<Block, Lanes, Claims>(a, p, 0, true, smem, s_cnt, &s_min, s_dummy, acc);
      __syncthreads();
    }
  }
  if (stats && tid == 0) {
    t_fin = __builtin_amdgcn_s_memrealtime() - t0;
    atomicAdd(&a.stats[9], t_fin);
    atomicAdd(&a.stats[10], n_fr);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long ok = __shfl_xor(acc.key, o);
    acc.key = ok > acc.key ? ok : acc.key;
    acc.unreached += __shfl_xor(acc.unreached, o);
  }
  const U128 sumsq = wave_sum_u128(acc.sumsq);
  const int32_t maxd = wave_max_i32(acc.maxd);
  if (lane == 0) {
    if (a.min_dist && acc.key) atomicMax(&a.kc->key, acc.key);
    atomic_add_u128(a.kc->sumsq, sumsq);
    if (acc.unreached) atomicAdd(&a.kc->unreached, acc.unreached);
    if (maxd > 0) atomicMax(reinterpret_cast<unsigned long long *>(&a.kc->depth), (unsigned long long)maxd);
  }
}

// Step 3 of a multi-source search (after obfs_kernel's queue): part by part,
// the sources in groups of 32 by part_bfs_multi where the part fits, else
// one at a time.  Every block takes parts until none is left.
template <int Block, int Lanes, int Claims>
__global__ __launch_bounds__(Block) void obfs_multi_final_kernel(ObfsArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ uint32_t s_part;
  __shared__ int s_cnt[3], s_min, s_mf[4];
  __shared__ int32_t s_dummy[Block];
  const int tid = threadIdx.x;
  KcAcc acc;
  unsigned long long n_fr = 0;
    // part by part: groups of 32 sources at once where the part fits; each
    // part's successor claimed while the part runs
    uint32_t next = 0;
    if (tid == 0) next = atomicAdd(&a.ctl[kCtlFin], 1u);
    for (;;) {
      if (tid == 0) {
        s_part = next;
        if ((int64_t)next < a.k) next = atomicAdd(&a.ctl[kCtlFin], 1u);
      }
      __syncthreads();
      const uint32_t p = s_part;
      if ((int64_t)p >= a.k) break;
      const int nv = (int)(a.vstart[p + 1] - a.vstart[p]);
      const int nb = (int)(a.bstart[p + 1] - a.bstart[p]);
      const int nblob = (int)(a.cblob_off[p + 1] - a.cblob_off[p]);  // the image, in 16-byte units
      const bool fits = (int64_t)16 * nblob + 12 * nv + align_up(kMultiW * nb, 16) + 4 * kMultiW + 16 <= (int64_t)a.olds;
      bool image = false;
      for (int s0 = 0; s0 < a.nsrc; s0 += kMultiW) {
        const int w = a.nsrc - s0 < kMultiW ? a.nsrc - s0 : kMultiW;
        bool done = false;
        if (fits) {
          if (!image) {
            load_cblob<Block>(a.cblob, a.cblob_off, p, smem);
            __syncthreads();
            image = true;
          }
          done = part_bfs_multi<Block>(a, p, s0, w, smem, s_mf);
        }
        if (!done) {
          for (int i = 0; i < w; i++) {
            n_fr += part_bfs<Block, Lanes, Claims>(a, p, s0 + i, true, smem, s_cnt, &s_min, s_dummy, acc);
            __syncthreads();
          }
          image = false;
        }
      }
      __syncthreads();
    }
  (void)n_fr;
}

// Dp of every part (context creation): bit-parallel BFS inside the part
// from W = 8 sizeof(M) of its boundary vertices at a time (bit i of a
// vertex's mask: reached from boundary vertex s0 + i), level-synchronous
// and pulled (a vertex ORs its neighbours' newly reached bits).  A boundary
// vertex newly reached at level L gets L in the rows of those sources.
struct DmatArgs {
  int64_t k;
  const uint32_t *vstart, *bstart, *cblob_off;
  const uint4 *cblob;
  unsigned char *dmat;
  const unsigned long long *doff;
  unsigned char *din;  // nullable: also every boundary vertex's distance to every vertex
  const unsigned long long *dinoff;
  unsigned int *err;  // set when an in-part distance reaches kDInf
};

template <typename M, int Block>
__global__ __launch_bounds__(Block) void dmat_kernel(DmatArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ int s_any[2];
  constexpr int W = 8 * (int)sizeof(M);
  const int tid = threadIdx.x;
  for (int64_t pp = blockIdx.x; pp < a.k; pp += gridDim.x) {
    const uint32_t p = (uint32_t)pp;
    const uint32_t v0 = a.vstart[p];
    const int nv = (int)(a.vstart[p + 1] - v0);
    const int nb = (int)(a.bstart[p + 1] - a.bstart[p]);
    const int nbp = (int)drow(nb);
    load_cblob<Block>(a.cblob, a.cblob_off, p, smem);
    __syncthreads();
    const uint16_t *lrow = reinterpret_cast<const uint16_t *>(smem);
    const int ne = lrow[nv];
    const uint16_t *ladj = lrow + nv + 1;
    const uint16_t *blist = ladj + ne;
    uint16_t *bpos = reinterpret_cast<uint16_t *>(smem + cblob_bytes(nv, ne, nb));
    M *vis = reinterpret_cast<M *>(smem + cblob_bytes(nv, ne, nb) + align_up(2 * nv, 16));
    M *cur = vis + nv;
    M *nxt = cur + nv;
    for (int v = tid; v < nv; v += Block) bpos[v] = 0xffff;
    __syncthreads();
    for (int j = tid; j < nb; j += Block) bpos[blist[j]] = (uint16_t)j;
    unsigned char *D = a.dmat + a.doff[p];
    unsigned char *Din = a.din ? a.din + a.dinoff[p] : nullptr;
    const int nvp = (int)align_up(nv, 16);  // Din row stride
    for (int s0 = 0; s0 < nb; s0 += W) {
      const int sw = nb - s0 < W ? nb - s0 : W;
      const M full = (M)(sw == W ? ~0ull : ((1ull << sw) - 1ull));
      for (int v = tid; v < nv; v += Block) vis[v] = cur[v] = 0;
      __syncthreads();
      for (int i = tid; i < sw; i += Block) {
        const int v = blist[s0 + i];
        vis[v] = cur[v] = (M)(1ull << i);
        D[(size_t)(s0 + i) * nbp + s0 + i] = 0;
        if (Din) Din[(size_t)(s0 + i) * nvp + v] = 0;
      }
      if (tid == 0) s_any[0] = s_any[1] = 0;
      __syncthreads();
      for (int L = 1;; L++) {
        for (int v = tid; v < nv; v += Block) {
          M nw = 0;
          const M vv = vis[v];
          if (vv != full) {
            M acc = 0;
            for (int x = lrow[v]; x < lrow[v + 1]; x++) acc |= cur[ladj[x]];
            nw = (M)(acc & (M)~vv);
          }
          nxt[v] = nw;
        }
        __syncthreads();
        for (int v = tid; v < nv; v += Block) {
          M nw = nxt[v];
          cur[v] = nw;
          if (Din && nw) {
            if (L >= (int)kDInf) {
              atomicOr(a.err, 2u);  // a distance the byte table cannot hold: no table
            } else {
              // source by source, so that a wave's byte stores for one
              // source land in consecutive bytes of that source's row
              for (int i = 0; i < sw; i++)
                if ((nw >> i) & 1) Din[(size_t)(s0 + i) * nvp + v] = (unsigned char)L;
            }
          }
          if (nw) {
            vis[v] |= nw;
            s_any[L & 1] = 1;
            const int j = bpos[v];
            if (j != 0xffff) {
              if (L >= (int)kDInf) {
                atomicOr(a.err, 1u);
              } else {
                unsigned long long bits = (unsigned long long)nw;
                while (bits) {
                  const int i = __ffsll((long long)bits) - 1;
                  bits &= bits - 1ull;
                  D[(size_t)(s0 + i) * nbp + j] = (unsigned char)L;
                }
              }
            }
          }
        }
        if (tid == 0) s_any[(L + 1) & 1] = 0;
        __syncthreads();
        if (!s_any[L & 1]) break;
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ host side

// Workgroup shapes: threads, lanes per frontier vertex, claims per lane.
struct PbfsShape {
  int block, lanes, claims;
  void (*fn)(PbfsArgs);
};
static const PbfsShape kShapes[] = {
    {1024, 16, 1, pbfs_kernel<1024, 16, 1>}, {512, 8, 2, pbfs_kernel<512, 8, 2>},
    {512, 16, 1, pbfs_kernel<512, 16, 1>},   {256, 4, 3, pbfs_kernel<256, 4, 3>},
    {256, 8, 2, pbfs_kernel<256, 8, 2>},     {512, 4, 3, pbfs_kernel<512, 4, 3>},
    {256, 16, 1, pbfs_kernel<256, 16, 1>},
};

// Shapes of the boundary-distance kernels: threads, lanes per frontier
// vertex and claims per lane of the interior passes.
struct ObfsShape {
  int block;
  void (*fn)(ObfsArgs);
  void (*src_fn)(ObfsArgs);
  void (*mfin_fn)(ObfsArgs);
};
#define OBFS_SHAPE(B, L, C) {B, obfs_kernel<B, L, C>, obfs_source_kernel<B, L, C>, obfs_multi_final_kernel<B, L, C>}
static const ObfsShape kObfsShapes[] = {
    OBFS_SHAPE(1024, 16, 1), OBFS_SHAPE(1024, 8, 2), OBFS_SHAPE(512, 8, 2), OBFS_SHAPE(512, 4, 3),
    OBFS_SHAPE(256, 4, 3),
};
#undef OBFS_SHAPE

// The parts of a deep graph (device arrays) and the traversal's state.
struct PartGraph {
  int64_t k = 0;
  uint32_t *vstart = nullptr, *pmidx = nullptr, *pcount = nullptr, *blob_off = nullptr;
  uint4 *blob = nullptr;
  std::vector<uint32_t> h_part_of, h_loc;
  int32_t *plevel = nullptr, *ilevel = nullptr;
  uint32_t *st = nullptr, *ctl = nullptr;
  unsigned long long *ring = nullptr;
  uint32_t ring_n = 0;
  bool disabled = false;
  PbfsShape shape;
  unsigned long long *stats = nullptr;  // PHDE_PBFS_STATS
  // the boundary-distance traversal (ovl false: not built, or PHDE_OBFS=0)
  bool ovl = false;
  int64_t nbtot = 0;
  uint32_t *pverts = nullptr, *bstart = nullptr, *xoff = nullptr, *noff = nullptr, *cblob_off = nullptr,
           *oblob_off = nullptr;
  uint4 *cblob = nullptr, *oblob = nullptr;
  unsigned char *dmat = nullptr;
  unsigned long long *doff = nullptr;
  unsigned char *din = nullptr;  // in-part distance tables for step 1 (nullptr: step 1 runs a BFS)
  unsigned long long *dinoff = nullptr;
  int32_t *entf = nullptr;  // step 3's entry candidates (with din)
  int32_t *blevel = nullptr, *bilevel = nullptr;
  ObfsShape oshape{};  // one search (k-centers, the selection's rounds)
  ObfsShape mshape{};  // several sources searched together
  size_t olds = 0;  // dynamic LDS of the boundary-distance kernels
  int ocap = 1;     // sources the item state, ring and boundary levels have room for
  unsigned char *mscr = nullptr;  // multi-source interiors: arrival rounds (n x 32 bytes)
};

// Frees the boundary-distance traversal's arrays (the partitioned traversal
// stays usable).
static void obfs_release(PartGraph *pg) {
  void *ptr[] = {pg->pverts, pg->bstart, pg->xoff, pg->noff,   pg->cblob_off, pg->oblob_off, pg->cblob,
                 pg->oblob,  pg->dmat,   pg->doff, pg->blevel, pg->bilevel,   pg->mscr,      pg->din,
                 pg->dinoff, pg->entf};
  for (void *x : ptr)
    if (x) (void)hipFree(x);
  pg->pverts = pg->bstart = pg->xoff = pg->noff = pg->cblob_off = pg->oblob_off = nullptr;
  pg->cblob = pg->oblob = nullptr;
  pg->dmat = nullptr;
  pg->doff = nullptr;
  pg->din = nullptr;
  pg->dinoff = nullptr;
  pg->entf = nullptr;
  pg->blevel = pg->bilevel = nullptr;
  pg->mscr = nullptr;
  pg->ovl = false;
}

void pbfs_release(phde_ctx *c) {
  PartGraph *pg = c->pg;
  if (!pg) return;
  obfs_release(pg);
  void *ptr[] = {pg->vstart, pg->pmidx, pg->pcount, pg->blob_off, pg->blob, pg->plevel, pg->ilevel,
                 pg->st,     pg->ring,  pg->ctl,    pg->stats};
  for (void *x : ptr)
    if (x) (void)hipFree(x);
  delete pg;
  c->pg = nullptr;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src) {
  PHDE_HIP(hipMalloc(dst, sizeof(T) * std::max<size_t>(src.size(), 1)));
  if (!src.empty()) PHDE_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return PHDE_OK;
}

// Host-side arrays of the boundary-distance traversal, built part by part
// with the partition.
struct ObfsHost {
  std::vector<uint32_t> bmidx, bstart, xoff{0}, noff{0}, coff{0}, eoff{0};
  std::vector<uint4> cblob, oblob;
  std::vector<unsigned long long> doff{0}, dinoff{0};
  int64_t nx_tot = 0, nn_tot = 0;
  // largest LDS of step 3 (and 1), of a run of step 2 (with its edge
  // blob), and of dmat_kernel with 32-, 16- and 8-bit masks
  int64_t max_final = 0, max_orun = 0, max_mf = 0, max_dm[3] = {0, 0, 0}, max_nb = 0, max_table = 0;
  std::vector<int64_t> fin;  // every part's step-3 LDS (statistics)
  bool ok = true;
  void add_part(int64_t p, uint32_t v0, uint32_t nv, const std::vector<uint32_t> &verts,
                const std::vector<uint16_t> &lrow, const std::vector<uint16_t> &ladj,
                const std::vector<uint16_t> &blist, const std::vector<uint16_t> &pxsrc,
                const std::vector<uint32_t> &pxdst, const std::vector<uint16_t> &pxnbr,
                const std::vector<uint32_t> &pnbr) {
    if (!ok) return;
    const uint32_t b0 = bstart[p];
    const int64_t ne = ladj.size(), nb = blist.size(), nn = pnbr.size(), nx = pxsrc.size();
    // the edge blob: external edges come only from boundary vertices (their
    // boundary index), pxdst holds the target's part-major index
    {
      const int64_t eb = oedge_bytes(nx, nn);
      std::vector<unsigned char> raw(eb, 0);
      uint16_t *xs = reinterpret_cast<uint16_t *>(raw.data());
      uint32_t *xd = reinterpret_cast<uint32_t *>(raw.data() + align_up(4 * nx, 16));
      uint32_t *nl = reinterpret_cast<uint32_t *>(raw.data() + 2 * align_up(4 * nx, 16));
      for (int64_t x = 0; x < nx; x++) {
        xs[x] = (uint16_t)(bmidx[verts[v0 + pxsrc[x]]] - b0);
        xs[nx + x] = pxnbr[x];
        xd[x] = bmidx[verts[pxdst[x]]];
      }
      for (int64_t i = 0; i < nn; i++) nl[i] = pnbr[i];
      const size_t at = oblob.size();
      oblob.resize(at + eb / 16);
      if (eb) std::memcpy(&oblob[at], raw.data(), eb);
      eoff.push_back((uint32_t)oblob.size());
      nx_tot += nx;
      nn_tot += nn;
      xoff.push_back((uint32_t)nx_tot);
      noff.push_back((uint32_t)nn_tot);
      max_orun = std::max(max_orun, align_up(orun_bytes(nb, nn), 16) + eb);
    }
    const int64_t cb = cblob_bytes(nv, ne, nb);
    std::vector<uint16_t> b16(cb / 2, 0);
    std::copy(lrow.begin(), lrow.end(), b16.begin());
    std::copy(ladj.begin(), ladj.end(), b16.begin() + lrow.size());
    std::copy(blist.begin(), blist.end(), b16.begin() + lrow.size() + ladj.size());
    const size_t at = cblob.size();
    cblob.resize(at + cb / 16);
    std::memcpy(&cblob[at], b16.data(), cb);
    coff.push_back((uint32_t)cblob.size());
    doff.push_back(doff.back() + (unsigned long long)(nb * drow(nb)));
    dinoff.push_back(dinoff.back() + (unsigned long long)nb * (unsigned long long)align_up(nv, 16));
    max_table = std::max(max_table, table_bytes(nv, nb));
    max_final = std::max(max_final, final_bytes(nv, ne, nb));
    fin.push_back(final_bytes(nv, ne, nb));
    max_mf = std::max(max_mf, mfinal_bytes(nv, ne, nb));
    max_nb = std::max(max_nb, nb);
    const int64_t w[3] = {4, 2, 1};
    for (int i = 0; i < 3; i++) max_dm[i] = std::max(max_dm[i], cb + align_up(2 * nv, 16) + 3 * w[i] * nv);
    if (nx_tot >= UINT32_MAX || nn_tot >= UINT32_MAX || cblob.size() >= UINT32_MAX || oblob.size() >= UINT32_MAX)
      ok = false;
  }
};

// Builds the boundary-distance traversal over the parts just cut: uploads
// its arrays and computes every part's Dp on the GPU.  Leaves pg->ovl false
// (the partitioned traversal alone) when PHDE_OBFS=0, when a part's LDS
// would not fit, or when an in-part distance reaches 255.
static int obfs_setup(phde_ctx *c, PartGraph *pg, ObfsHost &oh, const std::vector<uint32_t> &verts) {
  if (const char *e = getenv("PHDE_OBFS"))
    if (atoi(e) == 0) return PHDE_OK;
  // a level plus an in-part distance (< 255) must stay below kUnvisited
  if (c->g.n >= (int64_t)kUnvisited - 1024) return PHDE_OK;
  if (!oh.ok || std::max(oh.max_final, oh.max_orun) > kPartBytes) return PHDE_OK;
  if (oh.oblob.empty()) oh.oblob.push_back(make_uint4(0, 0, 0, 0));
  int mask = -1;
  for (int i = 0; i < 3 && mask < 0; i++)
    if (oh.max_dm[i] <= kPartBytes) mask = i;
  if (mask < 0) return PHDE_OK;
  PHDE_TRY_RC(upload(&pg->pverts, verts));
  PHDE_TRY_RC(upload(&pg->bstart, oh.bstart));
  PHDE_TRY_RC(upload(&pg->xoff, oh.xoff));
  PHDE_TRY_RC(upload(&pg->noff, oh.noff));
  PHDE_TRY_RC(upload(&pg->cblob_off, oh.coff));
  PHDE_TRY_RC(upload(&pg->oblob_off, oh.eoff));
  PHDE_TRY_RC(upload(&pg->cblob, oh.cblob));
  PHDE_TRY_RC(upload(&pg->oblob, oh.oblob));
  PHDE_TRY_RC(upload(&pg->doff, oh.doff));
  pg->nbtot = oh.bstart.back();
  const size_t dbytes = std::max<size_t>(oh.doff.back(), 16);
  PHDE_HIP(hipMalloc(&pg->dmat, dbytes));
  PHDE_HIP(hipMalloc(&pg->blevel, sizeof(int32_t) * std::max<int64_t>(pg->nbtot, 1)));
  PHDE_HIP(hipMalloc(&pg->bilevel, sizeof(int32_t) * std::max<int64_t>(pg->nbtot, 1)));
  PHDE_HIP(hipMemsetAsync(pg->dmat, kDInf, dbytes, c->stream));
  // step 1's tables (every boundary vertex's in-part distance to every
  // vertex of its part), up to kDinCap bytes; PHDE_OBFS_DIN=0: step 1 runs a
  // BFS inside the source's part instead
  constexpr unsigned long long kDinCap = 16ull << 30;
  bool din = oh.dinoff.back() > 0 && oh.dinoff.back() <= kDinCap;
  if (const char *e = getenv("PHDE_OBFS_DIN")) din = din && atoi(e) != 0;
  if (din) {
    PHDE_TRY_RC(upload(&pg->dinoff, oh.dinoff));
    if (hipMalloc(&pg->din, oh.dinoff.back()) != hipSuccess) {
      (void)hipGetLastError();
      pg->din = nullptr;
    } else {
      PHDE_HIP(hipMemsetAsync(pg->din, kDInf, oh.dinoff.back(), c->stream));
    }
  }
  DBuf<unsigned int> err;
  PHDE_TRY_RC(err.alloc(1));
  PHDE_HIP(hipMemsetAsync(err.p, 0, sizeof(unsigned int), c->stream));
  DmatArgs da;
  da.k = pg->k;
  da.vstart = pg->vstart;
  da.bstart = pg->bstart;
  da.cblob_off = pg->cblob_off;
  da.cblob = pg->cblob;
  da.dmat = pg->dmat;
  da.doff = pg->doff;
  da.din = pg->din;
  da.dinoff = pg->dinoff;
  da.err = err.p;
  void (*dk)(DmatArgs) = mask == 0 ? dmat_kernel<uint32_t, 1024> : mask == 1 ? dmat_kernel<uint16_t, 1024>
                                                                             : dmat_kernel<uint8_t, 1024>;
  const size_t dl = (size_t)oh.max_dm[mask];
  PHDE_HIP(hipFuncSetAttribute((const void *)dk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dl));
  const auto t0 = std::chrono::steady_clock::now();
  const int dgrid = (int)std::min<int64_t>(pg->k, (int64_t)1 << 20);
  dk<<<dgrid, 1024, dl, c->stream>>>(da);
  PHDE_HIP(hipGetLastError());
  unsigned int h_err = 0;
  PHDE_HIP(hipMemcpyAsync(&h_err, err.p, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  // one search: 512 threads -- one workgroup per CU all the same (LDS),
  // two waves per SIMD with room for 256 VGPRs each and eight waves per
  // barrier (RGG 4.2M: 0.860 ms per traversal; 1,024 threads 0.941, 256
  // threads 1.056); several sources together: 1,024 threads (50 sources
  // at RGG 4.2M 20.1 ms, at 512 threads 20.7; 16.8M 85.3 and 88.0)
  const int nshapes = (int)(sizeof(kObfsShapes) / sizeof(kObfsShapes[0]));
  pg->oshape = kObfsShapes[2];
  pg->mshape = kObfsShapes[1];
  if (const char *e = getenv("PHDE_OBFS_SHAPE"))
    pg->oshape = pg->mshape = kObfsShapes[std::max(0, std::min(nshapes - 1, atoi(e)))];
  // PHDE_OBFS_MSHAPE: the several-source shape alone (random-pivot columns
  // at RGG 4.2M: 24.1 ms at 1,024 threads × 8 lanes, 24.2 at 16 lanes,
  // 24.7 and 25.1 at 512 threads)
  if (const char *e = getenv("PHDE_OBFS_MSHAPE")) pg->mshape = kObfsShapes[std::max(0, std::min(nshapes - 1, atoi(e)))];
  // LDS: step 3's largest part, a run's state with its edge blob and, up to
  // 152 KB (the kernel's static LDS takes 4 KB more), the multi-source
  // interior pass (parts beyond it take their sources one at a time)
  pg->olds = (size_t)align_up(
      std::max({oh.max_final, oh.max_orun, oh.max_table, std::min<int64_t>(oh.max_mf, 152 * 1024)}), 256);
  for (const void *f : {(const void *)pg->oshape.fn, (const void *)pg->oshape.src_fn, (const void *)pg->oshape.mfin_fn,
                        (const void *)pg->mshape.fn, (const void *)pg->mshape.src_fn, (const void *)pg->mshape.mfin_fn})
    PHDE_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pg->olds));
  if ((h_err & 2u) && pg->din) {
    // some vertex lies 255 or more in-part hops from a boundary vertex: no
    // tables, step 1 runs its BFS
    (void)hipFree(pg->din);
    pg->din = nullptr;
  }
  if (pg->din) PHDE_HIP(hipMalloc(&pg->entf, sizeof(int32_t) * (size_t)std::max<int64_t>(pg->nbtot, 1)));
  if (pg->stats) {
    std::fprintf(stderr, "[obfs] boundary %lld (largest part %lld), Dp %.1f MB in %.1f ms (%d-bit masks)%s; "
                 "step-1 tables %.1f MB%s\n",
                 (long long)pg->nbtot, (long long)oh.max_nb, oh.doff.back() / 1e6, ms, 32 >> mask,
                 (h_err & 1u) ? ": an in-part distance reached 255, not used" : "",
                 pg->din ? oh.dinoff.back() / 1e6 : 0.0, pg->din ? "" : " (off)");
    std::vector<int64_t> f = oh.fin;
    std::sort(f.begin(), f.end());
    if (!f.empty())
      std::fprintf(stderr, "[obfs] LDS %zu B; step-3 LDS per part: median %lld, 90%% %lld, 99%% %lld, max %lld\n",
                   pg->olds, (long long)f[f.size() / 2], (long long)f[f.size() * 9 / 10],
                   (long long)f[f.size() * 99 / 100], (long long)f.back());
  }
  if (h_err & 1u) {
    obfs_release(pg);
    return PHDE_OK;
  }
  pg->ovl = true;
  return PHDE_OK;
}

// Greedy graph growing: the lowest-numbered unassigned vertex seeds a part,
// which grows breadth-first over unassigned vertices while its LDS image
// fits the budget (internal and external edges counted exactly as the part
// grows, every vertex counted as a boundary vertex).  Each part's vertices
// are numbered locally in that growth order; an edge to another part
// becomes an external edge of the part (local source, part-major index of
// the target, target part) and makes its target a boundary vertex of the
// target's part.
int pbfs_setup(phde_ctx *c) {
  if (c->pg) return PHDE_OK;
  if (const char *e = getenv("PHDE_PBFS"))
    if (atoi(e) == 0) return PHDE_OK;
  const int64_t n = c->g.n, m = c->g.m;
  if (n >= (int64_t)UINT32_MAX - 2) return PHDE_OK;
  // part budget in LDS bytes (smaller budgets: more parts, for tests)
  int64_t budget = kPartBytes;
  if (const char *e = getenv("PHDE_PBFS_BYTES")) budget = std::max<int64_t>(128, std::min<int64_t>(kPartBytes, atoll(e)));
  PbfsShape shape = kShapes[0];
  if (const char *e = getenv("PHDE_PBFS_SHAPE"))
    shape = kShapes[std::max(0, std::min((int)(sizeof(kShapes) / sizeof(kShapes[0])) - 1, atoi(e)))];
  int resident = 0;
  PHDE_HIP(hipFuncSetAttribute((const void *)shape.fn, hipFuncAttributeMaxDynamicSharedMemorySize, kPartBytes));
  PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, shape.fn, shape.block, kPartBytes));
  if (resident < 1) return PHDE_OK;
  std::vector<uint32_t> rp(n + 1), adj(std::max<int64_t>(m, 1));
  PHDE_HIP(hipMemcpy(rp.data(), c->g.rowptr, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost));
  if (m) PHDE_HIP(hipMemcpy(adj.data(), c->g.adj, sizeof(uint32_t) * m, hipMemcpyDeviceToHost));
  std::vector<uint32_t> part_of(n, kNoPart), vstart{0}, verts;
  verts.reserve(n);
  std::vector<uint32_t> q;
  int64_t k = 0;
  for (int64_t s = 0; s < n; s++) {
    if (part_of[s] != kNoPart) continue;
    const uint32_t p = (uint32_t)k++;
    q.clear();
    q.push_back((uint32_t)s);
    part_of[s] = p;
    int64_t nv = 1, ne = 0, nx = 0;
    for (uint32_t x = rp[s]; x < rp[s + 1]; x++) {
      if (adj[x] == (uint32_t)s) ne++;  // a self loop stays inside
      else nx++;
    }
    if (part_bytes(nv, ne, nv, nx, nx) > budget || ne + nx > 65535) return PHDE_OK;  // a vertex too wide for one part
    for (size_t h = 0; h < q.size(); h++) {
      const uint32_t u = q[h];
      for (uint32_t x = rp[u]; x < rp[u + 1]; x++) {
        const uint32_t w = adj[x];
        if (part_of[w] != kNoPart) continue;
        // w joins: its edges into the part turn internal from both sides
        int64_t in = 0, out = 0;
        for (uint32_t y = rp[w]; y < rp[w + 1]; y++) {
          const uint32_t z = adj[y];
          if (part_of[z] == p || z == w) in++;
          else out++;
        }
        const int64_t ne2 = ne + 2 * in, nx2 = nx - in + out;
        // 16-bit local ids and row offsets
        if (part_bytes(nv + 1, ne2, nv + 1, nx2, nx2) > budget || nv + 1 > 65535 || ne2 + nv + 1 > 65535) continue;
        nv++;
        ne = ne2;
        nx = nx2;
        part_of[w] = p;
        q.push_back(w);
      }
    }
    verts.insert(verts.end(), q.begin(), q.end());
    vstart.push_back((uint32_t)verts.size());
  }
  std::vector<uint16_t> loc(n);
  std::vector<uint32_t> pmidx(n);
  for (int64_t p = 0; p < k; p++)
    for (uint32_t i = vstart[p]; i < vstart[p + 1]; i++) {
      loc[verts[i]] = (uint16_t)(i - vstart[p]);
      pmidx[verts[i]] = i;
    }
  std::vector<unsigned char> boundary(n, 0);
  for (int64_t u = 0; u < n; u++)
    for (uint32_t x = rp[u]; x < rp[u + 1]; x++)
      if (part_of[adj[x]] != part_of[u]) boundary[adj[x]] = 1;
  std::vector<uint32_t> pcount(4 * k), blob_off{0};
  std::vector<int32_t> slot_of(k, -1);  // a part's slot in the current part's neighbour list
  std::vector<uint16_t> b16;   // one part's uint16 section
  std::vector<uint32_t> b32;   // and its uint32 section
  std::vector<uint4> blob;
  blob.reserve((size_t)(2 * m + 8 * n) / 16 + k);
  int64_t mn = n, mx = 0, worst = 0, next = 0, nbtot = 0;
  // the boundary-distance traversal's arrays: boundary-major numbering (part
  // by part, each part's boundary vertices in local order), external edges
  // from boundary index to boundary-major index, compact images, and the
  // byte offsets of the parts' Dp
  ObfsHost oh;
  oh.bmidx.assign(n, 0xffffffffu);
  oh.bstart.push_back(0);
  {
    uint32_t nbt = 0;
    for (int64_t p = 0; p < k; p++) {
      for (uint32_t i = vstart[p]; i < vstart[p + 1]; i++)
        if (boundary[verts[i]]) oh.bmidx[verts[i]] = nbt++;
      oh.bstart.push_back(nbt);
    }
  }
  for (int64_t p = 0; p < k; p++) {
    const uint32_t v0 = vstart[p], nv = vstart[p + 1] - v0;
    std::vector<uint16_t> lrow, ladj, blist, xsrc, xnbr;
    std::vector<uint32_t> xdst, nbr;
    for (uint32_t i = 0; i < nv; i++) {
      const uint32_t u = verts[v0 + i];
      lrow.push_back((uint16_t)ladj.size());
      for (uint32_t x = rp[u]; x < rp[u + 1]; x++) {
        const uint32_t w = adj[x];
        if (part_of[w] == (uint32_t)p) {
          ladj.push_back(loc[w]);
        } else {
          xsrc.push_back((uint16_t)i);
          xdst.push_back(pmidx[w]);
          int32_t &slot = slot_of[part_of[w]];
          if (slot < 0) {
            slot = (int32_t)nbr.size();
            nbr.push_back(part_of[w]);
          }
          xnbr.push_back((uint16_t)slot);
        }
      }
      if (boundary[u]) blist.push_back((uint16_t)i);
    }
    lrow.push_back((uint16_t)ladj.size());
    for (uint32_t q : nbr) slot_of[q] = -1;
    const int64_t ne = ladj.size(), nb = blist.size(), nx = xsrc.size(), nn = nbr.size();
    pcount[4 * p] = (uint32_t)ne;
    pcount[4 * p + 1] = (uint32_t)nb;
    pcount[4 * p + 2] = (uint32_t)nx;
    pcount[4 * p + 3] = (uint32_t)nn;
    b16.assign(lrow.begin(), lrow.end());
    b16.insert(b16.end(), ladj.begin(), ladj.end());
    b16.insert(b16.end(), blist.begin(), blist.end());
    b16.insert(b16.end(), xsrc.begin(), xsrc.end());
    b16.insert(b16.end(), xnbr.begin(), xnbr.end());
    const int64_t bytes = blob_bytes(nv, ne, nb, nx, nn);
    std::vector<unsigned char> raw(bytes, 0);
    std::memcpy(raw.data(), b16.data(), 2 * b16.size());
    const int64_t o32 = align_up(2 * (int64_t)b16.size(), 4);
    std::memcpy(raw.data() + o32, xdst.data(), 4 * nx);
    std::memcpy(raw.data() + o32 + 4 * nx, nbr.data(), 4 * nn);
    const size_t at = blob.size();
    blob.resize(at + bytes / 16);
    std::memcpy(&blob[at], raw.data(), bytes);
    next += bytes / 16;
    if (next > (int64_t)UINT32_MAX) return PHDE_OK;
    blob_off.push_back((uint32_t)next);
    mn = std::min<int64_t>(mn, nv);
    mx = std::max<int64_t>(mx, nv);
    worst = std::max(worst, part_bytes(nv, ne, nb, nx, nn));
    nbtot += nb;
    oh.add_part(p, v0, nv, verts, lrow, ladj, blist, xsrc, xdst, xnbr, nbr);
  }
  if (worst > kPartBytes) return PHDE_OK;  // cannot happen: the growth counted every vertex as boundary
  PartGraph *pg = new PartGraph();
  c->pg = pg;
  pg->k = k;
  pg->shape = shape;
  pg->ring_n = (uint32_t)(k + 4 * std::max(1, c->num_cus) + 64);
  PHDE_TRY_RC(upload(&pg->vstart, vstart));
  PHDE_TRY_RC(upload(&pg->pmidx, pmidx));
  PHDE_TRY_RC(upload(&pg->pcount, pcount));
  PHDE_TRY_RC(upload(&pg->blob_off, blob_off));
  PHDE_TRY_RC(upload(&pg->blob, blob));
  PHDE_HIP(hipMalloc(&pg->plevel, sizeof(int32_t) * n));
  PHDE_HIP(hipMalloc(&pg->ilevel, sizeof(int32_t) * n));
  PHDE_HIP(hipMalloc(&pg->st, sizeof(uint32_t) * k));
  PHDE_HIP(hipMalloc(&pg->ring, sizeof(unsigned long long) * kSlotStride * pg->ring_n));
  PHDE_HIP(hipMalloc(&pg->ctl, sizeof(uint32_t) * kCtlWords));
  if (getenv("PHDE_PBFS_STATS")) {
    PHDE_HIP(hipMalloc(&pg->stats, sizeof(unsigned long long) * (32 + pg->ring_n)));
    std::fprintf(stderr,
                 "[pbfs] %lld parts, vertices per part %lld..%lld, largest LDS image %lld bytes, %lld boundary "
                 "vertices, blobs %.1f MB\n",
                 (long long)k, (long long)mn, (long long)mx, (long long)worst, (long long)nbtot, 16.0 * next / 1e6);
  }
  pg->h_part_of.swap(part_of);
  pg->h_loc.resize(n);
  for (int64_t v = 0; v < n; v++) pg->h_loc[v] = loc[v];
  return obfs_setup(c, pg, oh, verts);
}

int64_t pbfs_num_parts(const phde_ctx *c) { return c->pg && !c->pg->disabled ? c->pg->k : 0; }
int64_t pbfs_num_boundary(const phde_ctx *c) {
  return c->pg && !c->pg->disabled && c->pg->ovl ? c->pg->nbtot : -1;
}

// One traversal by the boundary-distance method (steps 1-3 above).
// Room for nsrc sources searched together: item state (nsrc * k), the
// ring, the per-source boundary levels.
static int obfs_reserve(phde_ctx *c, PartGraph *pg, int nsrc) {
  if (pg->ocap >= nsrc) return PHDE_OK;
  const uint32_t ring_n = (uint32_t)(nsrc * pg->k + 4 * std::max(1, c->num_cus) + 64);
  PHDE_HIP(hipStreamSynchronize(c->stream));
  for (void *x : {(void *)pg->st, (void *)pg->ring, (void *)pg->blevel, (void *)pg->bilevel, (void *)pg->stats})
    if (x) (void)hipFree(x);
  pg->st = nullptr;
  pg->ring = nullptr;
  pg->blevel = pg->bilevel = nullptr;
  const bool stats = pg->stats != nullptr;
  pg->stats = nullptr;
  PHDE_HIP(hipMalloc(&pg->st, sizeof(uint32_t) * nsrc * pg->k));
  PHDE_HIP(hipMalloc(&pg->ring, sizeof(unsigned long long) * kSlotStride * ring_n));
  PHDE_HIP(hipMalloc(&pg->blevel, sizeof(int32_t) * std::max<int64_t>(nsrc * pg->nbtot, 1)));
  PHDE_HIP(hipMalloc(&pg->bilevel, sizeof(int32_t) * std::max<int64_t>(nsrc * pg->nbtot, 1)));
  if (stats) PHDE_HIP(hipMalloc(&pg->stats, sizeof(unsigned long long) * (32 + ring_n)));
  if (nsrc > 1 && !pg->mscr) PHDE_HIP(hipMalloc(&pg->mscr, (size_t)kMultiW * std::max<int64_t>(c->g.n, 1)));
  pg->ring_n = ring_n;
  pg->ocap = nsrc;
  return PHDE_OK;
}

// One boundary-distance search (steps 1-3 above) from nsrc <= kMaxSources
// sources (internal ids) together; levels to out (n per source, internal
// ids).  min_dist (one source only): the fused k-centers step.
static int obfs_run(phde_ctx *c, PartGraph *pg, const int64_t *srcs, int nsrc, int32_t *out, int32_t *min_dist,
                    bool *ran, bool seeds = false, int32_t bound = kUnvisited) {
  // seeds: one search from the nsrc sources together at level 0
  const int nseed = nsrc;
  if (seeds) nsrc = 1;
  PHDE_TRY_RC(obfs_reserve(c, pg, nsrc));
  const int64_t src = srcs[0];
  ObfsArgs a;
  a.g = c->g;
  a.level = out;
  a.min_dist = nsrc == 1 ? min_dist : nullptr;
  a.kc = c->d_kc;
  a.k = pg->k;
  a.vstart = pg->vstart;
  a.pverts = pg->pverts;
  a.bstart = pg->bstart;
  a.xoff = pg->xoff;
  a.noff = pg->noff;
  a.oblob = pg->oblob;
  a.oblob_off = pg->oblob_off;
  a.olds = (uint32_t)pg->olds;
  a.cblob = pg->cblob;
  a.cblob_off = pg->cblob_off;
  a.dmat = pg->dmat;
  a.doff = pg->doff;
  a.din = pg->din;
  a.dinoff = pg->dinoff;
  a.entf = pg->entf;
  a.blevel = pg->blevel;
  a.bilevel = pg->bilevel;
  a.st = pg->st;
  a.ctl = pg->ctl;
  a.ring = pg->ring;
  a.ring_n = pg->ring_n;
  a.nsrc = nsrc;
  a.nbtot = pg->nbtot;
  a.mscr = pg->mscr;
  a.nseed = nseed;
  a.bound = bound;
  a.nq = 0;
  for (int i = 0; i < nseed; i++) {
    a.src_part[i] = pg->h_part_of[srcs[i]];
    a.src_loc[i] = pg->h_loc[srcs[i]];
    const uint32_t item = seeds ? a.src_part[i] : (uint32_t)(i * pg->k) + a.src_part[i];
    bool dup = false;
    for (int r = 0; r < a.nq; r++) dup = dup || a.qitem[r] == item;
    if (!dup) a.qitem[a.nq++] = item;
  }
  a.stats = pg->stats;
  if (pg->stats) {
    const unsigned long long init[32] = {0, 0, 0, 0, 0, 0, 0, ~0ull, 0};
    PHDE_HIP(hipMemcpyAsync(pg->stats, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  }
  const ObfsShape &sh = nsrc > 1 ? pg->mshape : pg->oshape;
  const auto fn = sh.fn;
  const int block = sh.block;
  int resident = 0;
  PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, fn, block, pg->olds));
  // at most 4 workgroups per CU (the queue's ring holds the items + 4 per CU)
  const int grid = std::max(1, c->num_cus) * std::max(1, std::min(4, resident));
  if (resident < 1 || (int64_t)pg->ring_n < nsrc * pg->k + grid) {
    obfs_release(pg);
    return PHDE_OK;
  }
  const int ib = (int)std::max<int64_t>(
      1, std::min<int64_t>((std::max<int64_t>(nsrc * pg->nbtot, pg->ring_n) + 255) / 256, (int64_t)c->num_cus * 8));
  obfs_init_kernel<<<ib, 256, 0, c->stream>>>(a);
  PHDE_HIP(hipGetLastError());
  sh.src_fn<<<nseed, block, pg->olds, c->stream>>>(a);
  PHDE_HIP(hipGetLastError());
  void *args[] = {&a};
  const hipError_t e = launch_resident((const void *)fn, grid, block, args, pg->olds, c->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    obfs_release(pg);
    return PHDE_OK;
  }
  if (nsrc > 1) {
    sh.mfin_fn<<<grid, block, pg->olds, c->stream>>>(a);
    PHDE_HIP(hipGetLastError());
  }
  unsigned int flag = 0;
  PHDE_HIP(hipMemcpyAsync(c->h_kc, c->d_kc, sizeof(KcResult), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(&flag, pg->ctl + kCtlErr, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (flag) {
    // a wait timed out: this context goes back to the partitioned traversal
    std::fprintf(stderr, "parhde: boundary-distance BFS stopped (queue %u); using the partitioned traversal\n", flag);
    obfs_release(pg);
    return PHDE_OK;
  }
  if (pg->stats) {
    unsigned long long s[32];
    PHDE_HIP(hipMemcpy(s, pg->stats, sizeof(s), hipMemcpyDeviceToHost));
    const double us = 0.01;  // ticks of 10 ns
    const double na = (double)std::max(1ull, s[4]);
    std::fprintf(stderr,
                 "[obfs] src %lld: step 2 %.1f us (to last release %.1f us); summed over workgroups: wait %.0f seed "
                 "%.0f min-plus %.0f push %.0f us; %llu runs (%.2f per part), %.1f seeds and %.1f lowered per run, "
                 "%llu level loads; per run: seed %.2f min-plus %.2f push %.2f us; step 3 %.0f us summed, %llu "
                 "rounds (%.1f per part)\n",
                 (long long)src, (s[12] - s[7]) * us, (s[8] - s[7]) * us, s[0] * us, s[1] * us, s[2] * us, s[3] * us,
                 s[4], s[4] / (double)pg->k, s[5] / na, s[11] / na, s[6], s[1] * us / na, s[2] * us / na,
                 s[3] * us / na, s[9] * us, s[10], s[10] / (double)pg->k);
    std::fprintf(stderr, "[obfs]   queue latency (queued to taken): %llu parts, mean %.2f us, %llu above 10 us, max %.1f us\n",
                 s[17], s[16] * 0.01 / (double)std::max(1ull, s[17]), s[18], s[19] * 0.01);
    std::fprintf(stderr,
                 "[obfs]   step 3 per part: image or entries %.2f seeds or table rows %.2f rounds %.2f write-out %.2f "
                 "us; %.3f us per round; %.1f entries of %.1f candidates per part\n",
                 s[20] * us / pg->k, s[21] * us / pg->k, s[22] * us / pg->k, s[23] * us / pg->k,
                 s[22] * us / (double)std::max(1ull, s[10]), s[24] / (double)pg->k, s[25] / (double)pg->k);
  }
  *ran = true;
  return PHDE_OK;
}

int pbfs_multi_run(phde_ctx *c, const int64_t *src, int k, int32_t *D, bool *ran) {
  *ran = false;
  PartGraph *pg = c->pg;
  if (!pg || pg->disabled || !c->bfs_persistent || !pg->ovl || k < 1 || k > kMaxSources) return PHDE_OK;
  return obfs_run(c, pg, src, k, D, nullptr, ran);
}

bool pbfs_seeds_ok(const phde_ctx *c) {
  const PartGraph *pg = c->pg;
  return pg && !pg->disabled && c->bfs_persistent && pg->ovl && pg->din;
}

int pbfs_seeds_run(phde_ctx *c, const int64_t *seeds, int k, int32_t *col, int32_t *min_dist, int32_t bound,
                   bool *ran) {
  *ran = false;
  PartGraph *pg = c->pg;
  if (!pbfs_seeds_ok(c) || k < 1) return PHDE_OK;
  // groups of at most kMaxSources seeds and two in a part (the interior
  // pass of a part lists its boundary and two more vertices); the distance
  // to the nearest seed is the least over the groups
  std::vector<int64_t> rest(seeds, seeds + k);
  while (!rest.empty()) {
    std::vector<int64_t> grp, later;
    std::map<uint32_t, int> in_part;
    for (const int64_t v : rest) {
      int &cnt = in_part[pg->h_part_of[v]];
      if (cnt < 2 && (int)grp.size() < kMaxSources) {
        cnt++;
        grp.push_back(v);
      } else {
        later.push_back(v);
      }
    }
    bool r = false;
    PHDE_TRY_RC(obfs_run(c, pg, grp.data(), (int)grp.size(), col, min_dist, &r, true, bound));
    if (!r) return PHDE_OK;
    rest.swap(later);
  }
  *ran = true;
  return PHDE_OK;
}

int pbfs_run(phde_ctx *c, int64_t src, int32_t *dist_out, int32_t *min_dist, bool *ran) {
  *ran = false;
  PartGraph *pg = c->pg;
  if (!pg || pg->disabled || !c->bfs_persistent) return PHDE_OK;
  if (pg->ovl) {
    PHDE_TRY_RC(obfs_run(c, pg, &src, 1, dist_out ? dist_out : c->d_dist, min_dist, ran));
    if (*ran) return PHDE_OK;
  }
  PbfsArgs a;
  a.g = c->g;
  a.level = dist_out ? dist_out : c->d_dist;
  a.plevel = pg->plevel;
  a.ilevel = pg->ilevel;
  a.vstart = pg->vstart;
  a.pmidx = pg->pmidx;
  a.pcount = pg->pcount;
  a.blob = pg->blob;
  a.blob_off = pg->blob_off;
  a.st = pg->st;
  a.ring = pg->ring;
  a.ctl = pg->ctl;
  a.ring_n = pg->ring_n;
  a.src_part = pg->h_part_of[src];
  a.src_loc = pg->h_loc[src];
  a.k = pg->k;
  a.src = src;
  a.min_dist = min_dist;
  a.kc = c->d_kc;
  a.arrive = c->d_bar;
  a.bar_err = c->d_bar + 1;
  a.backoff = c->bar_backoff;
  a.cap = kRoundCap;
  if (const char *e = getenv("PHDE_PBFS_CAP")) a.cap = std::max(1, atoi(e));
  a.stats = pg->stats;
  if (pg->stats) {
    const unsigned long long init[32] = {0, 0, 0, 0, 0, 0, 0, ~0ull, 0};
    PHDE_HIP(hipMemcpyAsync(pg->stats, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
  }
  const int ib = (int)std::max<int64_t>(1, std::min<int64_t>((c->g.n + 255) / 256, (int64_t)c->num_cus * 8));
  pbfs_init_kernel<<<ib, 256, 0, c->stream>>>(a);
  PHDE_HIP(hipGetLastError());
  const int grid = std::max(1, c->num_cus);
  if ((int64_t)pg->ring_n < pg->k + grid) {  // the ticket slots would be reused while waited on
    pg->disabled = true;
    return PHDE_OK;
  }
  void *args[] = {&a};
  const hipError_t e = launch_resident((const void *)pg->shape.fn, grid, pg->shape.block, args, kPartBytes,
                                       c->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pg->disabled = true;
    return PHDE_OK;
  }
  unsigned int flags[2] = {0, 0};
  PHDE_HIP(hipMemcpyAsync(c->h_kc, c->d_kc, sizeof(KcResult), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(flags, c->d_bar + 1, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(flags + 1, pg->ctl + kCtlErr, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (flags[0] || flags[1]) {
    // a wait timed out: this context goes back to the multi-hop traversal
    std::fprintf(stderr, "parhde: partitioned BFS stopped (barrier %u, queue %u); using the multi-hop traversal\n",
                 flags[0], flags[1]);
    pg->disabled = true;
    return PHDE_OK;
  }
  if (pg->stats) {
    unsigned long long s[32];
    PHDE_HIP(hipMemcpy(s, pg->stats, sizeof(s), hipMemcpyDeviceToHost));
    std::fprintf(stderr, "[pbfs]   queue latency (queued to taken): %llu parts, mean %.2f us, %llu above 10 us, max %.1f us\n",
                 s[17], s[16] * 0.01 / (double)std::max(1ull, s[17]), s[18], s[19] * 0.01);
    const double us = 0.01;  // ticks of 10 ns
    const double na = (double)std::max(1ull, s[4]);
    std::fprintf(stderr,
                 "[pbfs] src %lld: span %.1f us; summed over workgroups: wait %.0f seed %.0f rounds %.0f push %.0f us; "
                 "%llu activations, %llu rounds (%.1f per activation), %llu image loads; per activation: seed %.2f "
                 "rounds %.2f push %.2f us, %.3f us per round\n",
                 (long long)src, (s[8] - s[7]) * us, s[0] * us, s[1] * us, s[2] * us, s[3] * us, s[4], s[5],
                 s[5] / na, s[6], s[1] * us / na, s[2] * us / na, s[3] * us / na,
                 s[2] * us / (double)std::max(1ull, s[5]));
    std::fprintf(stderr, "[pbfs]   %llu passes, %llu runs going on from the round cap, %.1f listed vertices per round\n",
                 s[9], s[10], (double)s[13] / (double)std::max(1ull, s[5]));
  }
  *ran = true;
  return PHDE_OK;
}

}  // namespace phde
