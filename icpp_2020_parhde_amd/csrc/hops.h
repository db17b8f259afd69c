// hops.h — frontier queue entries and the multi-hop expansion shared by the
// BFS (bfs.hip) and the greedy pivot selection's bounded BFS (learned.hip).
//
// Multi-hop expansion: between two grid barriers each wave expands its batch
// of the frontier for several hops, keeping the vertices it discovers in a
// wave-local LDS queue (64 entries; the rest, and the last hop's
// discoveries, go to the global queue for the next round).  A vertex u
// reached from v (level l) is claimed with atomicMin(lv[u], l + 1) and
// expanded again whenever that lowers its level, so a vertex first reached
// by a longer path (another wave was slower to reach it) is corrected when
// the shorter path arrives: the levels at the end are the exact BFS levels
// (label-correcting search; levels only decrease and every value written is
// the length of a real path).  Only the amount of work depends on timing.
// Claims stop at level `bound`.  Entries carry their level.  An entry that
// would not fit in the global queue sets *overflow; the caller then repeats
// the search level by level.  Several searches can share the queues: an
// entry's level word then holds the search's index above bit kHopSrcShift,
// and search j's levels are the column lv + j * col.
#pragma once

#include "device_util.h"
#include "phde_common.h"

namespace phde {

// Frontier queue entry: the vertex and its adjacency range, so the next
// level reads one 16-byte entry instead of entry + row pointers; lvl is the
// vertex's level when it was queued (multi-hop expansion).
struct QEnt {
  uint32_t v, b, e, lvl;
};

__device__ __forceinline__ QEnt ld_qent(const QEnt *q) {
  const unsigned long long *p = reinterpret_cast<const unsigned long long *>(q);
  const unsigned long long x = ld_dev(p), y = ld_dev(p + 1);
  return QEnt{(uint32_t)x, (uint32_t)(x >> 32), (uint32_t)y, (uint32_t)(y >> 32)};
}
__device__ __forceinline__ void st_qent(QEnt *q, QEnt v) {
  unsigned long long *p = reinterpret_cast<unsigned long long *>(q);
  st_dev(p, (unsigned long long)v.v | ((unsigned long long)v.b << 32));
  st_dev(p + 1, (unsigned long long)v.e | ((unsigned long long)v.lvl << 32));
}

constexpr int kHopBigDegree = 32;  // rows at least this long are expanded by the whole wave
constexpr int kHopSrcShift = 26;   // several searches: index above this bit of the level word

struct HopCtx {
  DevGraph g;
  int32_t *lv;              // levels, lowered by atomicMin (search j: lv + j * col)
  int64_t col;              // column stride of several searches (0: one search)
  uint32_t lvl_mask;        // the level bits of an entry's level word
  int32_t *dummy;           // this thread's scratch word (masked-off claims)
  unsigned int *overflow;   // set when an entry does not fit in the global queue
  int32_t bound;            // largest level claimed
  uint32_t cap;             // global queue entries
  int hops;                 // hops per round
  int batch;                // queue entries a wave starts a round with
  int64_t wrank, nwaves;    // this wave's rank in the round-robin deal, and the wave count
};

// Wave-collective: place each lane's winners (c[j]: vertex vv[j], rows
// [vb[j], ve[j]), level nl) in the wave's local queue loc while it has room
// (only when `local`), the rest in the global queue (count in the high half
// of *qs_next).  loc_n is the local count (wave-uniform), updated.
template <int G>
__device__ __forceinline__ void hop_emit(const HopCtx &h, const bool *c, const uint32_t *vv, const uint32_t *vb,
                                         const uint32_t *ve, const uint32_t *nl, bool local, QEnt *loc,
                                         int &loc_n, QEnt *q_next, unsigned long long *qs_next) {
  int cnt = 0;
#pragma unroll
  for (int j = 0; j < G; j++) cnt += c[j] ? 1 : 0;
  int total = 0;
  const int off = wave_excl_scan(cnt, &total);
  if (total == 0) return;
  const int avail = local ? kWave - loc_n : 0;
  const int nglob = total > avail ? total - avail : 0;
  unsigned int base = 0;
  if (__lane_id() == 0 && nglob) base = (unsigned int)(atomicAdd(qs_next, (unsigned long long)nglob << 32) >> 32);
  base = __shfl(base, 0);
  int r = off;
#pragma unroll
  for (int j = 0; j < G; j++) {
    if (c[j]) {
      const QEnt ent{vv[j], vb[j], ve[j], nl[j]};
      if (r < avail) {
        loc[loc_n + r] = ent;
      } else {
        const unsigned int gi = base + (unsigned int)(r - avail);
        if (gi < h.cap)
          st_qent(&q_next[gi], ent);
        else
          atomicOr(h.overflow, 1u);
      }
      r++;
    }
  }
  loc_n = local ? (loc_n + total < kWave ? loc_n + total : kWave) : 0;
}

// One hop from the lanes' short rows (ds neighbours from e.b; maxd their
// wave maximum), G neighbours per lane per round trip: all loads, level
// checks and claims of a group in flight together.
template <int G>
__device__ __forceinline__ void hop_small(const HopCtx &h, const QEnt &e, int ds, int maxd, bool local, QEnt *loc,
                                          int &loc_n, QEnt *q_next, unsigned long long *qs_next) {
  const uint32_t *rowptr = h.g.rowptr;
  const uint32_t *adj = h.g.adj;
  const uint32_t tag = e.lvl & ~h.lvl_mask;  // the search's index bits
  const int32_t nl = (int32_t)(e.lvl & h.lvl_mask) + 1;
  int32_t *lvc = h.lv + (int64_t)(tag >> kHopSrcShift) * h.col;
  for (int g0 = 0; g0 < maxd; g0 += G) {
    uint32_t vv[G], vb[G], ve[G], lv[G];
    bool c[G];
#pragma unroll
    for (int j = 0; j < G; j++) {
      c[j] = g0 + j < ds;
      vv[j] = adj[c[j] ? e.b + g0 + j : 0u];  // unconditional: no branch per neighbour
    }
#pragma unroll
    for (int j = 0; j < G; j++) vv[j] = c[j] ? vv[j] : 0u;
    // plain loads: levels only decrease, so a stale line can only hold a
    // larger one, and the claim below decides
    int32_t dd[G];
#pragma unroll
    for (int j = 0; j < G; j++) dd[j] = lvc[c[j] ? vv[j] : 0u];
#pragma unroll
    for (int j = 0; j < G; j++) c[j] = c[j] && dd[j] > nl;
    // claims, all issued before any result is used (masked-off lanes aim
    // at the scratch word instead of branching)
    int32_t won[G];
#pragma unroll
    for (int j = 0; j < G; j++) won[j] = atomicMin(c[j] ? &lvc[vv[j]] : h.dummy, nl);
#pragma unroll
    for (int j = 0; j < G; j++) {
      const uint32_t r = c[j] ? vv[j] : 0u;
      vb[j] = rowptr[r];
      ve[j] = rowptr[r + 1];
    }
    // keep the row-pointer loads in flight with the claims
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < G; j++) {
      c[j] = c[j] && won[j] > nl;
      lv[j] = (uint32_t)nl | tag;
    }
    hop_emit<G>(h, c, vv, vb, ve, lv, local, loc, loc_n, q_next, qs_next);
  }
}

// One hop from the lanes' entries e (valid lanes, level below the bound):
// rows of kHopBigDegree or more neighbours by the whole wave, the others
// one lane each.
__device__ __forceinline__ void hop_expand(const HopCtx &h, const QEnt &e, bool valid, bool local, QEnt *loc,
                                           int &loc_n, QEnt *q_next, unsigned long long *qs_next) {
  const uint32_t *rowptr = h.g.rowptr;
  const uint32_t *adj = h.g.adj;
  const int lane = __lane_id();
  const int d = valid ? (int)(e.e - e.b) : 0;
  const bool big = d >= kHopBigDegree;
  unsigned long long bmask = __ballot(big);
  while (bmask) {
    const int l = __ffsll((long long)bmask) - 1;
    bmask &= bmask - 1;
    const uint32_t bb = __shfl(e.b, l), be = __shfl(e.e, l);
    const uint32_t word = __shfl(e.lvl, l);
    const uint32_t tag = word & ~h.lvl_mask;
    const int32_t nl = (int32_t)(word & h.lvl_mask) + 1;
    int32_t *lvc = h.lv + (int64_t)(tag >> kHopSrcShift) * h.col;
    for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
      const uint32_t k = k0 + lane;
      bool f = false;
      uint32_t v = 0, vb = 0, ve = 0, lv = (uint32_t)nl | tag;
      if (k < be) {
        v = adj[k];
        if (lvc[v] > nl && atomicMin(&lvc[v], nl) > nl) {
          f = true;
          vb = rowptr[v];
          ve = rowptr[v + 1];
        }
      }
      hop_emit<1>(h, &f, &v, &vb, &ve, &lv, local, loc, loc_n, q_next, qs_next);
    }
  }
  const int ds = big ? 0 : d;
  const int maxd = wave_max_i32(ds);
  if (maxd > 0) {
    if (maxd <= 8)
      hop_small<8>(h, e, ds, maxd, local, loc, loc_n, q_next, qs_next);
    else if (maxd <= 16 || maxd > 24)
      hop_small<16>(h, e, ds, maxd, local, loc, loc_n, q_next, qs_next);
    else
      hop_small<24>(h, e, ds, maxd, local, loc, loc_n, q_next, qs_next);
  }
}

// One round: the nq entries of q_cur, dealt to the waves in batches of
// h.batch (round-robin over the blocks), each expanded for up to h.hops
// hops.  locq: the block's LDS, 2 x (blockDim / 64) x 64 entries.
__device__ __forceinline__ void hop_round(const HopCtx &h, const QEnt *q_cur, uint32_t nq, QEnt *q_next,
                                          unsigned long long *qs_next, QEnt *locq) {
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  const int wpb = blockDim.x / kWave;
  const int hb = h.batch;
  const int64_t nbatch = ((int64_t)nq + hb - 1) / hb;
  for (int64_t k = h.wrank; k < nbatch; k += h.nwaves) {
    const int64_t idx = k * hb + lane;
    bool valid = lane < hb && idx < (int64_t)nq;
    QEnt e{0u, 0u, 0u, 0u};
    if (valid) e = ld_qent(&q_cur[idx]);
    int sel = 0;
    for (int hop = 0; hop < h.hops; hop++) {
      valid = valid && (int32_t)(e.lvl & h.lvl_mask) < h.bound;
      const bool local = hop + 1 < h.hops;
      int loc_n = 0;
      QEnt *loc = locq + ((sel ^ 1) * wpb + w) * kWave;
      hop_expand(h, e, valid, local, loc, loc_n, q_next, qs_next);
      if (!local || loc_n == 0) break;
      // the wave's own LDS writes, read back by other lanes of the wave
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      valid = lane < loc_n;
      e = valid ? loc[lane] : QEnt{0u, 0u, 0u, 0u};
      __builtin_amdgcn_wave_barrier();
      sel ^= 1;
    }
  }
}

}  // namespace phde
