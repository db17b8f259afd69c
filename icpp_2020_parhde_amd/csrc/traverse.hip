// traverse.hip — weighted shortest paths (delta-stepping) and 64-wide
// bit-parallel multi-source BFS for gfx950, one kernel launch per round /
// level (levels.h).
//
// SSSP follows DeltaStep (gapbs/sssp.cc:53-140) in what it computes: exact
// int32 distances, 2^30-1 for unreached vertices.  The bucket structure is
// the two-pile (near / far) form of delta-stepping: the near queue holds
// vertices whose tentative distance is below the current bucket bound and is
// relaxed round after round; improvements beyond the bound go to the far
// pile, which is filtered into the next bucket when the near queue empties.
// Relaxation is an atomicMin on the distance, so the result is the exact
// shortest-path distance for any delta (label correcting).  Vertices enter a
// queue at most once per round thanks to per-vertex round stamps.  The next
// bucket bound comes from the smallest distance pushed to the far pile
// (tracked with atomicMin), so no extra pass is needed to find it.
//
// Multi-source BFS carries up to 64 searches in one 64-bit mask per vertex
// (bit k = source k): a vertex's `fresh` mask holds the searches that first
// reached it at the current level, pushed to neighbours with atomicOr; the
// level of search k at v is written when bit k first appears.  No reference
// counterpart; it equals k single-source BFS.

#include <algorithm>
#include <climits>
#include <cstdio>

#include "device_util.h"
#include "hops.h"
#include "phde_common.h"
#include "levels.h"
#include "pbfs.h"
#include "traverse.h"


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

struct SsspState {
  long long bound;  // current bucket: tentative distances < bound
  int epoch, nsel, fsel, pad;
};

struct SsspArgs {
  DevGraph g;
  int32_t *dist;
  uint32_t *near0, *near1, *far0, *far1;
  int32_t *nstamp, *fstamp;  // round / epoch in which a vertex was last queued
  unsigned int *ncnt;        // [3] near-queue sizes, rotating by round
  unsigned int *fcnt;        // [2] far-pile sizes, by epoch parity
  int *fmin;                 // [2] smallest distance pushed to each far pile
  SsspState *st;             // [2] by round parity
  int64_t src;
  int32_t delta;
  int32_t *min_dist;
  KcResult *kc;
  int *done_dev, *host_flags;
  int32_t *dummy;  // one scratch word per thread: target of masked-off atomics
  int hops;        // > 1: a wave relaxes what its batch improves within the bucket for up to this many hops
  int hop_batch;   // near-queue entries a wave starts with (hops > 1)
};

constexpr int kSsspGroup = 8;  // edges a lane relaxes per round trip

__global__ __launch_bounds__(kTBlock) void sssp_init_kernel(SsspArgs a) {
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gtid; i < n; i += gsize) {
    a.dist[i] = (i == a.src) ? 0 : kDistInf;
    a.nstamp[i] = -1;
    a.fstamp[i] = -1;
  }
  if (gtid == 0) {
    a.ncnt[0] = 1;
    a.ncnt[1] = a.ncnt[2] = 0;
    a.fcnt[0] = a.fcnt[1] = 0;
    a.fmin[0] = a.fmin[1] = INT32_MAX;
    a.near0[0] = (uint32_t)a.src;
    SsspState s0{};
    s0.bound = a.delta;
    a.st[0] = s0;
    a.kc->key = 0;
    a.kc->sumsq[0] = a.kc->sumsq[1] = 0;
    a.kc->unreached = 0;
    a.kc->depth = 0;
    *a.done_dev = -1;
  }
}

// Relax one edge (u -> v, tentative nd) of the near queue.
__device__ __forceinline__ void relax(const SsspArgs &a, uint32_t v, int32_t nd, long long bound,
                                      int round, int epoch, bool *pn, bool *pf) {
  if (nd < a.dist[v]) {
    const int32_t old = atomicMin(&a.dist[v], nd);
    if (nd < old) {
      if (nd < bound) {
        *pn = atomicExch(&a.nstamp[v], round + 1) != round + 1;
      } else {
        atomicMin(&a.fmin[epoch & 1], nd);
        *pf = atomicExch(&a.fstamp[v], epoch) != epoch;
      }
    }
  }
}

__device__ void sssp_epilogue(const SsspArgs &a) {
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  unsigned long long key = 0, unreached = 0;
  U128 sumsq{0, 0};
  for (int64_t i = gtid; i < n; i += gsize) {
    const int32_t d = a.dist[i];
    if (d >= kDistInf)
      unreached++;
    else
      u128_add(sumsq, (unsigned long long)d * (unsigned long long)d);
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
  if (__lane_id() == 0) {
    if (a.min_dist && key) atomicMax(&a.kc->key, key);
    atomic_add_u128(a.kc->sumsq, sumsq);
    if (unreached) atomicAdd(&a.kc->unreached, unreached);
  }
}

// Near-queue relaxation of one round with several hops per launch (graphs
// with a deep BFS): a wave takes a batch of the near queue and relaxes the
// vertices its relaxations improve below the bucket bound itself, up to
// a.hops hops, through a 64-entry LDS queue that carries each vertex's
// distance and row range; what does not fit and the last hop's
// improvements join the next round's near queue (near stamp), improvements
// beyond the bound the far pile (far stamp), as in the one-hop round.
// Relaxation stays atomicMin on the distance, so the distances are exact;
// a vertex may be relaxed more often than in the one-hop rounds.
__device__ void sssp_relax_hops(const SsspArgs &a, int R, const SsspState &s, unsigned int nq, const uint32_t *ncur,
                                uint32_t *nnext, unsigned int *nnext_cnt, uint32_t *fcur, unsigned int *fcnt,
                                QEnt *locq) {
  constexpr int G = kSsspGroup;
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  const int wpb = blockDim.x / kWave;
  const int64_t wrank = (int64_t)w * gridDim.x + blockIdx.x;
  const int64_t nwaves = (int64_t)gridDim.x * wpb;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t *dum = a.dummy + gtid;
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;
  const int32_t *wt = a.g.w;
  const int hb = a.hop_batch;
  const int64_t nbatch = ((int64_t)nq + hb - 1) / hb;
  for (int64_t k = wrank; k < nbatch; k += nwaves) {
    const int64_t idx = k * hb + lane;
    bool valid = lane < hb && idx < (int64_t)nq;
    uint32_t u = valid ? ncur[idx] : 0u;
    int32_t du = valid ? a.dist[u] : 0;
    uint32_t b = valid ? rowptr[u] : 0u, e = valid ? rowptr[u + 1] : 0u;
    int sel = 0;
    for (int hop = 0; hop < a.hops; hop++) {
      const bool local = hop + 1 < a.hops;
      QEnt *loc = locq + ((sel ^ 1) * wpb + w) * kWave;
      int loc_n = 0;
      const int d = valid ? (int)(e - b) : 0;
      const bool big = d >= 32;
      unsigned long long bmask = __ballot(big);
      while (bmask) {  // long rows by the whole wave, improvements to the global piles
        const int l = __ffsll((long long)bmask) - 1;
        bmask &= bmask - 1;
        const uint32_t bb = __shfl(b, l), be = __shfl(e, l);
        const int32_t bd = __shfl(du, l);
        for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
          const uint32_t kk = k0 + lane;
          bool pn = false, pf = false;
          uint32_t v = 0;
          if (kk < be) {
            v = adj[kk];
            relax(a, v, bd + wt[kk], s.bound, R, s.epoch, &pn, &pf);
          }
          wave_push(pn, v, nnext, nnext_cnt);
          wave_push(pf, v, fcur, fcnt);
        }
      }
      const int ds = big ? 0 : d;
      int maxd = ds;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
      for (int t0 = 0; t0 < maxd; t0 += G) {
        uint32_t vv[G], vb[G], ve[G];
        int32_t nd[G], cur[G], wv[G];
        bool ok[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
          ok[j] = t0 + j < ds;
          const uint32_t kk = ok[j] ? b + t0 + j : 0u;
          vv[j] = adj[kk];
          wv[j] = wt[kk];
        }
#pragma unroll
        for (int j = 0; j < G; j++) {
          vv[j] = ok[j] ? vv[j] : 0u;
          nd[j] = ok[j] ? du + wv[j] : INT32_MAX;
        }
#pragma unroll
        for (int j = 0; j < G; j++) cur[j] = a.dist[vv[j]];
#pragma unroll
        for (int j = 0; j < G; j++) ok[j] = ok[j] && nd[j] < cur[j];
#pragma unroll
        for (int j = 0; j < G; j++) cur[j] = atomicMin(ok[j] ? &a.dist[vv[j]] : dum, nd[j]);
#pragma unroll
        for (int j = 0; j < G; j++) {
          const uint32_t r = ok[j] ? vv[j] : 0u;
          vb[j] = rowptr[r];
          ve[j] = rowptr[r + 1];
        }
        __builtin_amdgcn_sched_barrier(0);
        bool nr[G], fr[G];
        int fm = INT32_MAX, cn = 0;
#pragma unroll
        for (int j = 0; j < G; j++) {
          const bool imp = ok[j] && nd[j] < cur[j];
          nr[j] = imp && nd[j] < s.bound;
          fr[j] = imp && !(nd[j] < s.bound);
          if (fr[j]) fm = min(fm, nd[j]);
          cn += nr[j] ? 1 : 0;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) fm = min(fm, __shfl_xor(fm, o));
        if (lane == 0 && fm != INT32_MAX) atomicMin(&a.fmin[s.epoch & 1], fm);
        // near improvements: the wave's local queue while it has room
        int totn = 0;
        const int offn = wave_excl_scan(cn, &totn);
        const int avail = local ? kWave - loc_n : 0;
        bool gn[G];
        int r = offn;
#pragma unroll
        for (int j = 0; j < G; j++) {
          gn[j] = false;
          if (nr[j]) {
            if (r < avail)
              loc[loc_n + r] = QEnt{vv[j], vb[j], ve[j], (uint32_t)nd[j]};
            else
              gn[j] = true;
            r++;
          }
        }
        loc_n = local ? (loc_n + totn < kWave ? loc_n + totn : kWave) : 0;
        // the others: near / far stamps in one batch of exchanges, then the
        // global appends
#pragma unroll
        for (int j = 0; j < G; j++)
          cur[j] = atomicExch(gn[j] ? &a.nstamp[vv[j]] : fr[j] ? &a.fstamp[vv[j]] : dum, gn[j] ? R + 1 : s.epoch);
        int cg = 0, cf = 0;
#pragma unroll
        for (int j = 0; j < G; j++) {
          gn[j] = gn[j] && cur[j] != R + 1;
          fr[j] = fr[j] && cur[j] != s.epoch;
          cg += gn[j] ? 1 : 0;
          cf += fr[j] ? 1 : 0;
        }
        int totg = 0, totf = 0;
        const int offg = wave_excl_scan(cg, &totg);
        const int offf = wave_excl_scan(cf, &totf);
        unsigned int bg = 0, bf = 0;
        if (lane == 0) {
          if (totg) bg = atomicAdd(nnext_cnt, (unsigned int)totg);
          if (totf) bf = atomicAdd(fcnt, (unsigned int)totf);
        }
        bg = __shfl(bg, 0) + (unsigned int)offg;
        bf = __shfl(bf, 0) + (unsigned int)offf;
#pragma unroll
        for (int j = 0; j < G; j++) {
          if (gn[j]) nnext[bg++] = vv[j];
          if (fr[j]) fcur[bf++] = vv[j];
        }
      }
      if (!local || loc_n == 0) break;
      // the wave's own LDS writes, read back by other lanes of the wave
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      valid = lane < loc_n;
      const QEnt en = valid ? loc[lane] : QEnt{0u, 0u, 0u, 0u};
      __builtin_amdgcn_wave_barrier();
      u = en.v;
      b = en.b;
      e = en.e;
      du = (int32_t)en.lvl;
      sel ^= 1;
    }
  }
}

__global__ __launch_bounds__(kTBlock) void sssp_round_kernel(SsspArgs a, int R) {
  // the round's control words are independent: load them in one batch
  // before the early exit instead of one dependent load after another
  // (relaxed atomic loads, so that the compiler keeps them ahead of the
  // branch without waiting on each)
  SsspState *sp = &a.st[R & 1];
  const int done = __hip_atomic_load(a.done_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  SsspState s;
  s.bound = __hip_atomic_load(&sp->bound, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s.epoch = __hip_atomic_load(&sp->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s.nsel = __hip_atomic_load(&sp->nsel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s.fsel = __hip_atomic_load(&sp->fsel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  s.pad = 0;
  const unsigned int nq = aload(&a.ncnt[R % 3]);
  const unsigned int fc0 = aload(&a.fcnt[0]), fc1 = aload(&a.fcnt[1]);
  if (done >= 0 && done + 1 < R) return;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave, nwaves = gsize / kWave;
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;
  const int32_t *wt = a.g.w;

  const unsigned int nf = (s.epoch & 1) ? fc1 : fc0;
  if (nq == 0 && nf == 0) {
    if (gtid == 0) {
      level_done(a.done_dev, a.host_flags, R - 1);
      a.kc->depth = R;
    }
    sssp_epilogue(a);
    return;
  }
  if (gtid == 0) {
    level_publish(a.host_flags, R);
    a.ncnt[(R + 2) % 3] = 0;
  }
  const uint32_t *ncur = s.nsel ? a.near1 : a.near0;
  uint32_t *nnext = s.nsel ? a.near0 : a.near1;
  uint32_t *fcur = s.fsel ? a.far1 : a.far0;
  unsigned int *nnext_cnt = &a.ncnt[(R + 1) % 3];
  SsspState nx = s;
  nx.nsel ^= 1;
  if (nq > 0) {
    // ---- relax the near queue (sssp.cc:74-108)
    if (gtid == 0) {
      // the pile consumed by the last filter is free: ready it for the next
      a.fcnt[(s.epoch + 1) & 1] = 0;
      a.fmin[(s.epoch + 1) & 1] = INT32_MAX;
    }
    unsigned int *fcnt = &a.fcnt[s.epoch & 1];
    __shared__ QEnt locq[2][kTBlock / kWave][kWave];
    if (a.hops > 1) {
      sssp_relax_hops(a, R, s, nq, ncur, nnext, nnext_cnt, fcur, fcnt, &locq[0][0][0]);
      if (gtid == 0) a.st[(R + 1) & 1] = nx;
      return;
    }
    // Small rows: kSsspGroup edges per lane per round trip, every load and
    // atomic of the group issued before any result is used (masked-off
    // atomics go to the lane's scratch word); near / far appends staged in
    // LDS and published with one global atomic per block and chunk; the far
    // pile's minimum reduced per wave.  Rows of 32 or more edges are relaxed
    // by the whole wave.
    // One stage holds both: near appends grow up from 0, far appends down
    // from the top (a lane appends at most 31 of either kind per chunk).
    constexpr unsigned int kStage = kTBlock * 31;
    __shared__ uint32_t stage[kStage];
    __shared__ unsigned int nstn, nstf, gbn, gbf;
    if (threadIdx.x == 0) nstn = nstf = 0;
    int32_t *dum = a.dummy + gtid;
    const int w = threadIdx.x / kWave;
    for (int64_t cb = (int64_t)blockIdx.x * kTBlock; cb < (int64_t)nq; cb += (int64_t)gridDim.x * kTBlock) {
      __syncthreads();
      const int64_t idx = cb + w * kWave + lane;
      const bool act = idx < (int64_t)nq;
      const uint32_t u = act ? ncur[idx] : 0u;
      const uint32_t b = act ? rowptr[u] : 0u, e = act ? rowptr[u + 1] : 0u;
      const int32_t du = act ? a.dist[u] : 0;
      const int d = (int)(e - b);
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
            relax(a, v, bd + wt[k], s.bound, R, s.epoch, &pn, &pf);
          }
          wave_push(pn, v, nnext, nnext_cnt);
          wave_push(pf, v, fcur, fcnt);
        }
      }
      const int ds = big ? 0 : d;
      int maxd = ds;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
      for (int t0 = 0; t0 < maxd; t0 += kSsspGroup) {
        uint32_t vv[kSsspGroup];
        int32_t nd[kSsspGroup], cur[kSsspGroup];
        bool ok[kSsspGroup];
        // unconditional loads (index 0 for lanes past their row): a load
        // under a lane condition whose value is used in the same branch
        // makes the compiler wait for it there, one round trip per edge
        int32_t wv[kSsspGroup];
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          ok[j] = t0 + j < ds;
          const uint32_t k = ok[j] ? b + t0 + j : 0u;
          vv[j] = adj[k];
          wv[j] = wt[k];
        }
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          vv[j] = ok[j] ? vv[j] : 0u;
          nd[j] = ok[j] ? du + wv[j] : INT32_MAX;
        }
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) cur[j] = a.dist[vv[j]];
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) ok[j] = ok[j] && nd[j] < cur[j];
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) cur[j] = atomicMin(ok[j] ? &a.dist[vv[j]] : dum, nd[j]);
        bool nr[kSsspGroup], fr[kSsspGroup];
        int fm = INT32_MAX;
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          const bool imp = ok[j] && nd[j] < cur[j];
          nr[j] = imp && nd[j] < s.bound;
          fr[j] = imp && !(nd[j] < s.bound);
          if (fr[j]) fm = min(fm, nd[j]);
        }
        // One stamp exchange per edge: an improvement goes to the near
        // queue or to the far pile, never both, so the near and far stamps
        // are swapped in one batch (one round trip instead of two).
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++)
          cur[j] = atomicExch(nr[j] ? &a.nstamp[vv[j]] : fr[j] ? &a.fstamp[vv[j]] : dum, nr[j] ? R + 1 : s.epoch);
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          nr[j] = nr[j] && cur[j] != R + 1;
          fr[j] = fr[j] && cur[j] != s.epoch;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) fm = min(fm, __shfl_xor(fm, o));
        if (lane == 0 && fm != INT32_MAX) atomicMin(&a.fmin[s.epoch & 1], fm);
        int cn = 0, cf = 0;
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          cn += nr[j] ? 1 : 0;
          cf += fr[j] ? 1 : 0;
        }
        int totn = 0, totf = 0;
        const int offn = wave_excl_scan(cn, &totn);
        const int offf = wave_excl_scan(cf, &totf);
        unsigned int bn = 0, bf = 0;
        if (lane == 0) {
          if (totn) bn = atomicAdd(&nstn, (unsigned int)totn);
          if (totf) bf = atomicAdd(&nstf, (unsigned int)totf);
        }
        bn = __shfl(bn, 0) + (unsigned int)offn;
        bf = __shfl(bf, 0) + (unsigned int)offf;
#pragma unroll
        for (int j = 0; j < kSsspGroup; j++) {
          if (nr[j]) stage[bn++] = vv[j];
          if (fr[j]) stage[kStage - 1 - bf++] = vv[j];
        }
      }
      __syncthreads();
      const unsigned int nn = nstn, nf2 = nstf;
      if (nn | nf2) {
        if (threadIdx.x == 0) {
          gbn = nn ? atomicAdd(nnext_cnt, nn) : 0u;
          gbf = nf2 ? atomicAdd(fcnt, nf2) : 0u;
        }
        __syncthreads();
        for (unsigned int i = threadIdx.x; i < nn; i += kTBlock) nnext[gbn + i] = stage[i];
        for (unsigned int i = threadIdx.x; i < nf2; i += kTBlock) fcur[gbf + i] = stage[kStage - 1 - i];
        __syncthreads();
        if (threadIdx.x == 0) nstn = nstf = 0;
      }
    }
  } else {
    // ---- near queue empty: the next bucket starts at the far pile's
    // smallest pushed distance; split the pile (sssp.cc:109-131)
    const int fm = a.fmin[s.epoch & 1];
    const long long bound = ((long long)fm / a.delta + 1) * (long long)a.delta;
    uint32_t *fnext = s.fsel ? a.far0 : a.far1;
    const int e1 = s.epoch + 1;
    unsigned int *fnext_cnt = &a.fcnt[e1 & 1];
    for (int64_t base = gwave * kWave; base < (int64_t)nf; base += nwaves * kWave) {
      const int64_t idx = base + lane;
      bool pn = false, pf = false;
      uint32_t v = 0;
      if (idx < (int64_t)nf) {
        v = fcur[idx];
        const int32_t dv = a.dist[v];
        if (dv < bound) {
          pn = atomicExch(&a.nstamp[v], R + 1) != R + 1;
        } else {
          atomicMin(&a.fmin[e1 & 1], dv);
          pf = atomicExch(&a.fstamp[v], e1) != e1;
        }
      }
      wave_push(pn, v, nnext, nnext_cnt);
      wave_push(pf, v, fnext, fnext_cnt);
    }
    nx.bound = bound;
    nx.epoch = e1;
    nx.fsel ^= 1;
  }
  if (gtid == 0) a.st[(R + 1) & 1] = nx;
}

// ------------------------------------------------------------------ MS-BFS

struct MsArgs {
  DevGraph g;
  int k;
  unsigned long long *visit, *fr[3];
  uint32_t *q[3];
  unsigned int *cnt;  // [4] queue sizes, rotating by level
  const int64_t *src;
  int32_t *D;  // n*k column-major
  int *done_dev, *host_flags;
  unsigned long long *dummy;  // one scratch word per thread: target of masked-off atomics
};

constexpr int kMsGroup = 8;  // neighbours a lane handles per round trip

__global__ __launch_bounds__(kTBlock) void msbfs_fill_kernel(MsArgs a) {
  const int64_t n = a.g.n;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gtid; i < n; i += gsize) {
    a.visit[i] = 0;
    a.fr[0][i] = 0;
    a.fr[1][i] = 0;
    a.fr[2][i] = 0;
  }
  for (int64_t e = gtid; e < n * a.k; e += gsize) a.D[e] = -1;
  if (gtid == 0) {
    for (int i = 0; i < 4; i++) a.cnt[i] = 0;
    *a.done_dev = -1;
  }
}

// Level 0 frontier: each source's bit in its pushed mask (visit and the
// level-0 distance are written when level 0 processes the queue).
__global__ void msbfs_source_kernel(MsArgs a) {
  const int t = threadIdx.x;
  if (t < a.k) {
    const int64_t s = a.src[t];
    const unsigned long long old = atomicOr(&a.fr[0][s], 1ull << t);
    if (old == 0) a.q[0][atomicAdd(&a.cnt[0], 1u)] = (uint32_t)s;
  }
}

// Offer `bits` to v for the next level: one atomic, and true for the offer
// that made v's pushed mask non-empty (v joins the next queue).  Bits v
// already holds may still be offered (visit is read without ordering); the
// next level filters them.
__device__ __forceinline__ bool ms_push(const MsArgs &a, unsigned long long bits, uint32_t v,
                                        unsigned long long *fn) {
  const unsigned long long nb = bits & ~a.visit[v];
  if (!nb) return false;
  return atomicOr(&fn[v], nb) == 0ull;
}

__global__ __launch_bounds__(kTBlock) void msbfs_level_kernel(MsArgs a, int L) {
  // control words in one batch of relaxed atomic loads ahead of the exit
  const int done = __hip_atomic_load(a.done_dev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned int nq = aload(&a.cnt[L % 4]);
  const unsigned int np = aload(&a.cnt[(L + 3) % 4]);
  if (done >= 0 && done + 1 < L) return;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;
  if (nq == 0) {
    if (gtid == 0) level_done(a.done_dev, a.host_flags, L - 1);
    return;
  }
  if (gtid == 0) {
    level_publish(a.host_flags, L);
    a.cnt[(L + 2) % 4] = 0;
  }
  // retire the previous level's fresh masks (read by the previous kernel)
  if (L > 0) {
    const uint32_t *qp = a.q[(L + 2) % 3];
    unsigned long long *fp = a.fr[(L + 2) % 3];
    for (int64_t i = gtid; i < (int64_t)np; i += gsize) fp[qp[i]] = 0;
  }
  const uint32_t *qc = a.q[L % 3];
  uint32_t *qn = a.q[(L + 1) % 3];
  const unsigned long long *fc = a.fr[L % 3];
  unsigned long long *fn = a.fr[(L + 1) % 3];
  unsigned int *cnext = &a.cnt[(L + 1) % 4];
  // Next-frontier appends of the small rows are staged in LDS and published
  // with one global atomic per block and chunk (a chunk of kTBlock queue
  // entries with degree < 32 appends at most kTBlock * 31 vertices); the
  // wave-cooperative big rows append to the global queue directly.
  __shared__ uint32_t stage[kTBlock * 31];
  __shared__ unsigned int nstage, gbase;
  if (threadIdx.x == 0) nstage = 0;
  const int w = threadIdx.x / kWave;
  for (int64_t cb = (int64_t)blockIdx.x * kTBlock; cb < (int64_t)nq; cb += (int64_t)gridDim.x * kTBlock) {
    __syncthreads();
    const int64_t idx = cb + w * kWave + lane;
    const bool act = idx < (int64_t)nq;
    const uint32_t u = act ? qc[idx] : 0u;
    // the bits that reach u first at this level: its distance for them is L
    const unsigned long long pushed = act ? fc[u] : 0ull;
    const unsigned long long vis = act ? a.visit[u] : ~0ull;
    const unsigned long long bits = pushed & ~vis;
    if (bits) {
      a.visit[u] = vis | bits;  // u occurs once in the queue: sole writer
      unsigned long long t = bits;
      while (t) {
        const int s = __ffsll((long long)t) - 1;
        t &= t - 1;
        a.D[(int64_t)s * a.g.n + u] = L;
      }
    }
    const uint32_t b = bits ? rowptr[u] : 0u, e = bits ? rowptr[u + 1] : 0u;
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
          push = ms_push(a, bbits, v, fn);
        }
        wave_push(push, v, qn, cnext);
      }
    }
    // Small rows: kMsGroup neighbours per lane per round trip, every load
    // and atomic of the group issued before any result is used; masked-off
    // atomics go to the lane's own scratch word instead of being predicated.
    const int ds = big ? 0 : d;
    int maxd = ds;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) maxd = max(maxd, __shfl_xor(maxd, o));
    unsigned long long *dummy = a.dummy + gtid;
    for (int t0 = 0; t0 < maxd; t0 += kMsGroup) {
      uint32_t vv[kMsGroup];
      unsigned long long nb[kMsGroup], old[kMsGroup];
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) vv[j] = adj[(t0 + j < ds) ? b + t0 + j : 0u];
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) vv[j] = (t0 + j < ds) ? vv[j] : 0u;
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) nb[j] = a.visit[vv[j]];
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) nb[j] = (t0 + j < ds) ? (bits & ~nb[j]) : 0ull;
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) old[j] = atomicOr(nb[j] ? &fn[vv[j]] : dummy, nb[j]);
      int npush = 0;
#pragma unroll
      for (int j = 0; j < kMsGroup; j++) npush += (nb[j] != 0ull && old[j] == 0ull) ? 1 : 0;
      int total = 0;
      const int off = wave_excl_scan(npush, &total);
      if (total) {
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(&nstage, (unsigned int)total);
        base = __shfl(base, 0) + (unsigned int)off;
#pragma unroll
        for (int j = 0; j < kMsGroup; j++)
          if (nb[j] != 0ull && old[j] == 0ull) stage[base++] = vv[j];
      }
    }
    __syncthreads();
    const unsigned int ns = nstage;
    if (ns) {
      if (threadIdx.x == 0) gbase = atomicAdd(cnext, ns);
      __syncthreads();
      for (unsigned int i = threadIdx.x; i < ns; i += kTBlock) qn[gbase + i] = stage[i];
      __syncthreads();
      if (threadIdx.x == 0) nstage = 0;
    }
  }
}

// ---- multi-source BFS of a graph with a deep BFS: the k searches run as
// concurrent multi-hop searches (hops.h) in one persistent launch, search
// j's levels in column j of D; the same levels as the bit-parallel sweep.
struct MsHopArgs {
  DevGraph g;
  int k;
  const int64_t *src;
  int32_t *D;
  QEnt *q0, *q1;
  unsigned long long *cnt;  // [3] queue counts (high halves), rotating by round
  int32_t *dummy;
  unsigned int *overflow;
  uint32_t cap;             // queue entries usable
  int hops, batch, backoff;
};

__global__ void mshop_fill_kernel(int32_t *D, int64_t total, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    D[i] = v;
}

__global__ void mshop_source_kernel(MsHopArgs a) {
  const int j = threadIdx.x;
  if (j < a.k) {
    const uint32_t s = (uint32_t)a.src[j];
    a.D[(int64_t)j * a.g.n + s] = 0;
    a.q0[j] = QEnt{s, a.g.rowptr[s], a.g.rowptr[s + 1], (uint32_t)j << kHopSrcShift};
  }
  if (j == 0) {
    a.cnt[0] = (unsigned long long)a.k << 32;
    a.cnt[1] = 0;
    a.cnt[2] = 0;
  }
}

__global__ void mshop_finish_kernel(int32_t *D, int64_t total) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
    if (D[i] == kUnvisited) D[i] = -1;
}

// *arrive must be 0 at launch; *err receives 1 if a barrier wait timed out.
__global__ __launch_bounds__(kTBlock) void msbfs_hops_kernel(MsHopArgs a, unsigned int *arrive, unsigned int *err) {
  const unsigned int nb = gridDim.x;
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  HopCtx h;
  h.g = a.g;
  h.lv = a.D;
  h.col = a.g.n;
  h.lvl_mask = (1u << kHopSrcShift) - 1u;
  h.dummy = a.dummy + gtid;
  h.overflow = a.overflow;
  h.bound = (int32_t)(1u << kHopSrcShift) - 2;
  h.cap = a.cap;
  h.hops = a.hops;
  h.batch = a.batch;
  h.wrank = (int64_t)(threadIdx.x / kWave) * gridDim.x + blockIdx.x;
  h.nwaves = (int64_t)gridDim.x * (blockDim.x / kWave);
  __shared__ QEnt locq[2][kTBlock / kWave][kWave];
  __shared__ unsigned int sh_nq;
  unsigned int bar = 0;
  for (int64_t R = 0;; R++) {
    if (threadIdx.x == 0) sh_nq = (unsigned int)(ld_dev(&a.cnt[R % 3]) >> 32);
    __syncthreads();
    const unsigned int nq = sh_nq;
    __syncthreads();
    if (nq == 0) break;
    if (gtid == 0) st_dev(&a.cnt[(R + 2) % 3], 0ull);
    hop_round(h, (R & 1) ? a.q1 : a.q0, nq < a.cap ? nq : a.cap, (R & 1) ? a.q0 : a.q1, &a.cnt[(R + 1) % 3],
              &locq[0][0][0]);
    if (grid_barrier(arrive, nb * ++bar, err, a.backoff)) return;
  }
}

// ------------------------------------------------------------------ host side

// Grid of the multi-source sweep: two 256-thread blocks per CU measured
// best on RGG n=16.7M with 62 sources (455 ms vs 480 / 466 / 517 ms at one,
// four and eight blocks per CU).
static int ms_grid(const phde_ctx *c) {
  static const int per_cu = [] {
    const char *e = getenv("PHDE_MS_BLOCKS_PER_CU");
    return e ? std::max(1, std::min(32, atoi(e))) : 2;
  }();
  return std::max(1, c->num_cus) * per_cu;
}

// Grid of the delta-stepping rounds (blocks per CU: PHDE_SSSP_BLOCKS_PER_CU).
static int sssp_grid(const phde_ctx *c) {
  static const int per_cu = [] {
    const char *e = getenv("PHDE_SSSP_BLOCKS_PER_CU");
    return e ? std::max(1, std::min(32, atoi(e))) : 2;
  }();
  return std::max(1, c->num_cus) * per_cu;
}

// Per-thread scratch words of the masked-off atomics, sized for the larger
// of the two grids above.
static int ensure_dummy(phde_ctx *c) {
  if (!c->ms_dummy) {
    const size_t words = (size_t)std::max(ms_grid(c), sssp_grid(c)) * kTBlock;
    PHDE_HIP(hipMalloc(&c->ms_dummy, sizeof(unsigned long long) * words));
  }
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
  PHDE_TRY_RC(ensure_dummy(c));
  if (!c->ts_far[0]) {
    for (int i = 0; i < 2; i++) PHDE_HIP(hipMalloc(&c->ts_far[i], sizeof(uint32_t) * (n + 1)));
    PHDE_HIP(hipMalloc(&c->ts_nstamp, sizeof(int32_t) * n));
    PHDE_HIP(hipMalloc(&c->ts_fstamp, sizeof(int32_t) * n));
  }
  if (!c->ts_cnt) PHDE_HIP(hipMalloc(&c->ts_cnt, 256));
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
  a.ncnt = c->ts_cnt;
  a.fcnt = c->ts_cnt + 4;
  a.fmin = (int *)(c->ts_cnt + 8);
  a.st = (SsspState *)(c->ts_cnt + 16);
  a.src = src;
  a.delta = delta;
  a.min_dist = min_dist;
  a.kc = c->d_kc;
  a.done_dev = c->d_done;
  a.host_flags = c->d_flags;
  a.dummy = (int32_t *)c->ms_dummy;
  // several hops per round on graphs with a deep BFS
  const char *eh = getenv("PHDE_SSSP_HOPS"), *eb = getenv("PHDE_SSSP_HOP_BATCH");
  // measured on the weighted 160^3 grid: 9.6 ms per SSSP in one-hop rounds,
  // 9.1 at 4 hops from 64-entry batches (9.3 at 3, 9.8 at 8; 10.4 from
  // batches of 32; its wide wavefront keeps every wave busy)
  const int sssp_hops = eh ? std::max(0, std::min(64, atoi(eh))) : 4;
  const int sssp_batch = eb ? std::max(1, std::min(kWave, atoi(eb))) : 64;
  a.hops = c->bfs_hops > 1 ? sssp_hops : 0;
  a.hop_batch = sssp_batch;
  c->h_flags[kFlagDone] = -1;
  c->h_flags[kFlagProgress] = -1;
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_SSSP, &ev);
  const int grid = sssp_grid(c);
  sssp_init_kernel<<<grid, kTBlock, 0, c->stream>>>(a);
  PHDE_HIP(hipGetLastError());
  rc = run_levels(c, [&](int R) -> int {
    sssp_round_kernel<<<grid, kTBlock, 0, c->stream>>>(a, R);
    PHDE_HIP(hipGetLastError());
    return PHDE_OK;
  });
  if (rc) return rc;
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
    for (int i = 0; i < 3; i++) PHDE_HIP(hipMalloc(&c->ms_fr[i], sizeof(unsigned long long) * n));
    PHDE_HIP(hipMalloc(&c->ms_q3, sizeof(uint32_t) * (n + 1)));
    PHDE_HIP(hipMalloc(&c->ms_src, sizeof(int64_t) * 64));
    PHDE_TRY_RC(ensure_dummy(c));
    if (!c->ts_cnt) PHDE_HIP(hipMalloc(&c->ts_cnt, 256));
  }
  PHDE_HIP(hipMemcpyAsync(c->ms_src, sources, sizeof(int64_t) * k, hipMemcpyHostToDevice, c->stream));
  if (c->pg) {
    // a deep graph cut into parts with per-part distance tables: the k
    // searches run together by the boundary-distance method (pbfs.hip)
    bool ran = false;
    hipEvent_t pev;
    c->timer.begin(c->stream, PHDE_KERNEL_MSBFS, &pev);
    rc = pbfs_multi_run(c, sources, k, D, &ran);
    if (!rc && ran) {
      mshop_finish_kernel<<<std::max(1, c->num_cus) * 4, kTBlock, 0, c->stream>>>(D, n * k);
      PHDE_HIP(hipGetLastError());
    }
    c->timer.end(c->stream, PHDE_KERNEL_MSBFS, pev);
    if (rc || ran) return rc;
  }
  // hops per round and first batch of the concurrent searches (PHDE_MS_HOPS=0:
  // the bit-parallel sweep for every graph)
  static const int ms_hops = [] {
    const char *e = getenv("PHDE_MS_HOPS");
    return e ? std::max(0, std::min(64, atoi(e))) : 4;
  }();
  static const int ms_batch = [] {
    // 52 searches, RGG 16.8M: 331 / 289 / 285 / 323 ms from batches of
    // 24 / 32 / 40 / 48; RGG 4.2M: 77.9 ms from 32, 78.4 from 40
    const char *e = getenv("PHDE_MS_HOP_BATCH");
    return e ? std::max(1, std::min(kWave, atoi(e))) : 40;
  }();
  if (ms_hops > 1 && c->bfs_hops > 1 && c->bfs_persistent && n < ((int64_t)1 << kHopSrcShift) && k <= 64) {
    // a graph with a deep BFS: concurrent multi-hop searches, one column each
    MsHopArgs h;
    h.g = c->g;
    h.k = k;
    h.src = c->ms_src;
    h.D = D;
    h.q0 = reinterpret_cast<QEnt *>(c->d_q[0]);
    h.q1 = reinterpret_cast<QEnt *>(c->d_q[1]);
    h.cnt = reinterpret_cast<unsigned long long *>(c->ts_cnt + 48);
    h.dummy = reinterpret_cast<int32_t *>(c->ms_dummy);
    h.overflow = c->d_bar + 2;
    h.cap = (c->bfs_qcap > 0 && c->bfs_qcap < (uint64_t)(n + 1)) ? c->bfs_qcap : (uint32_t)(n + 1);
    h.hops = ms_hops;
    h.batch = ms_batch;
    h.backoff = c->bar_backoff;
    hipEvent_t ev;
    c->timer.begin(c->stream, PHDE_KERNEL_MSBFS, &ev);
    const int fgrid = std::max(1, c->num_cus) * 4;
    mshop_fill_kernel<<<fgrid, kTBlock, 0, c->stream>>>(D, n * k, kUnvisited);
    mshop_source_kernel<<<1, 64, 0, c->stream>>>(h);
    PHDE_HIP(hipGetLastError());
    PHDE_TRY_RC(fill_u32(c, c->d_bar, 3, 0u));
    unsigned int *arrive = c->d_bar, *err = c->d_bar + 1;
    void *args[] = {&h, &arrive, &err};
    // blocks per CU: 52 searches on RGG 4.2M 78.1 ms at one, 79.0 at two;
    // on RGG 16.8M 302.5 ms at one, 291.9 at two
    const char *epc = getenv("PHDE_MS_HOP_BLOCKS_PER_CU");
    const int per_cu = epc ? std::max(1, std::min(2, atoi(epc))) : (n >= ((int64_t)1 << 23) ? 2 : 1);
    int resident = 0;
    PHDE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, msbfs_hops_kernel, kTBlock, 0));
    const int hgrid = std::max(1, c->num_cus) * std::max(1, std::min(per_cu, resident));
    const hipError_t e = launch_resident((const void *)msbfs_hops_kernel, hgrid, kTBlock, args, 0, c->stream);
    if (e == hipSuccess) {
      mshop_finish_kernel<<<fgrid, kTBlock, 0, c->stream>>>(D, n * k);
      PHDE_HIP(hipGetLastError());
      c->timer.end(c->stream, PHDE_KERNEL_MSBFS, ev);
      PHDE_HIP(hipMemcpyAsync(c->h_bar, c->d_bar + 1, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                              c->stream));
      PHDE_HIP(hipStreamSynchronize(c->stream));
      if (!c->h_bar[0] && !c->h_bar[1]) return PHDE_OK;
      // the grid barrier timed out (blocks kept from becoming resident) or
      // the queue overflowed: the bit-parallel sweep below recomputes D
    } else {
      (void)hipGetLastError();
      c->timer.end(c->stream, PHDE_KERNEL_MSBFS, ev);
    }
  }
  MsArgs a;
  a.g = c->g;
  a.k = k;
  a.visit = c->ms_visit;
  for (int i = 0; i < 3; i++) a.fr[i] = c->ms_fr[i];
  a.q[0] = c->d_q[0];
  a.q[1] = c->d_q[1];
  a.q[2] = c->ms_q3;
  a.cnt = c->ts_cnt + 32;
  a.src = c->ms_src;
  a.D = D;
  a.done_dev = c->d_done;
  a.host_flags = c->d_flags;
  a.dummy = c->ms_dummy;
  c->h_flags[kFlagDone] = -1;
  c->h_flags[kFlagProgress] = -1;
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_MSBFS, &ev);
  const int grid = ms_grid(c);
  msbfs_fill_kernel<<<grid, kTBlock, 0, c->stream>>>(a);
  msbfs_source_kernel<<<1, 64, 0, c->stream>>>(a);
  PHDE_HIP(hipGetLastError());
  rc = run_levels(c, [&](int L) -> int {
    msbfs_level_kernel<<<grid, kTBlock, 0, c->stream>>>(a, L);
    PHDE_HIP(hipGetLastError());
    return PHDE_OK;
  });
  if (rc) return rc;
  c->timer.end(c->stream, PHDE_KERNEL_MSBFS, ev);
  return PHDE_OK;
}

void traverse_release(phde_ctx *c) {
  void *p[] = {c->ts_far[0], c->ts_far[1], c->ts_nstamp, c->ts_fstamp, c->ts_cnt,
               c->ms_visit, c->ms_fr[0], c->ms_fr[1], c->ms_fr[2], c->ms_q3, c->ms_src,
               c->d_done, c->ms_dummy};
  if (c->h_flags) (void)hipHostFree(c->h_flags);
  for (void *x : p)
    if (x) (void)hipFree(x);
}

}  // namespace phde
