// learned.hip — learned pivot proposal.  No reference counterpart: the
// reference picks pivots by k-centers or at random (ParHDE.cpp:118-172,
// DrawingBase.cpp:119-194).  The k-centers path stays the judge and the
// fallback (api.hip).
//
// Features   x_v = (log(1 + deg_v) / 4, d1_v / ecc1, d2_v / ecc2, 1), where
//            d1 are the BFS levels from the first pivot and d2 those from
//            the vertex farthest from it: the first two k-centers
//            traversals, whose columns the layout reuses.
// Network    three message-passing layers of width 64 and a linear read-out,
//              h1 = relu([x, A x] W1 + b1)                      (FP32)
//              h_l = relu([h_{l-1}, A h_{l-1}] W_l + b_l), l=2,3 (BF16 MFMA,
//                                                               FP32 accumulate)
//              y  = h3 . wo + bo
//            A = D^-1 Adj is the neighbour mean.  Activations are stored as
//            BF16 (one 128-byte row per vertex).  Training minimises the
//            mean squared error against the k-centers coverage field with
//            hand-written backward kernels and Adam.
// Selection  greedy graph non-maximum suppression in one persistent kernel:
//            repeatedly take the best-scoring vertex farther than r hops from
//            every pivot so far and lower the distance-to-pivots field within
//            r hops of it (a BFS pruned where the field is already smaller).
//
// MFMA: __builtin_amdgcn_mfma_f32_32x32x16_bf16.  Lane l holds A[l%32][8*(l/32)
// + 0..7] and B[8*(l/32) + 0..7][l%32]; result element i of lane l is
// C[8*(i/4) + 4*(l/32) + i%4][l%32].
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "device_util.h"
#include "hops.h"
#include "learned.h"
#include "pbfs.h"
#include "traverse.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

struct phde_mpnn {
  int device = 0;
  hipStream_t stream = nullptr;  // parameter uploads outside a context
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // params, Adam moments, gradient
  uint16_t *wb = nullptr;      // BF16 copies of W2 and W3 ([2][128][64])
  long long step = 0;
  // per-launch partial gradients, one kParams row per workgroup
  float *part = nullptr;
  double *loss_part = nullptr, *loss = nullptr;
  int nblk = 0;
  // activation workspace (grown to the largest graph seen)
  int64_t cap = 0;
  float4 *xa = nullptr;
  uint16_t *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *za2 = nullptr, *za3 = nullptr, *dlt = nullptr;
  float *gs = nullptr, *ga = nullptr, *d32 = nullptr, *y = nullptr;
};

namespace phde {

constexpr int kNB = 256;  // threads per block, forward / read-out / reductions
constexpr int kBB = 128;  // threads per block, MFMA backward
constexpr int kAP = 136;  // padded row (bf16) of a 128-wide LDS tile
constexpr int kWP = 72;   // padded row of a 64-wide LDS tile
constexpr int kTP = 40;   // padded row of a 32-wide (transposed) LDS tile

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf16_lo(uint32_t x) { return __uint_as_float(x << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t x) { return __uint_as_float(x & 0xffff0000u); }
__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  return (uint32_t)bf16_bits(a) | ((uint32_t)bf16_bits(b) << 16);
}
__device__ __forceinline__ float inv_degree(const DevGraph &g, int64_t v) {
  const int d = g.deg32[v];
  return d > 0 ? 1.0f / (float)d : 0.0f;
}

// ------------------------------------------------------------------ features

__global__ void feature_kernel(DevGraph g, const int32_t *d1, const int32_t *d2, float inv1,
                               float inv2, float4 *x) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n;
       v += (int64_t)gridDim.x * blockDim.x)
    x[v] = make_float4(log1pf((float)g.deg32[v]) * 0.25f, (float)d1[v] * inv1, (float)d2[v] * inv2,
                       1.0f);
}

// ------------------------------------------------------------------ forward

// Neighbour gathers of the network layers, kGat neighbours of a row per
// round trip: ids [k0, k0 + kGat) of a row starting at b with rem = d - k0
// of them left.  The loads are unconditional (straight-line code keeps
// them all in flight); past a row's end the id loads read adj[0] and the
// row gathers read vertex 0, so every access stays in bounds.
constexpr int kGat = 16;

__device__ __forceinline__ void gather_ids(const uint32_t *adj, uint32_t b, int rem, uint32_t (&ids)[kGat]) {
#pragma unroll
  for (int j = 0; j < kGat; j++) ids[j] = adj[j < rem ? b + (uint32_t)j : 0u];
}

__device__ __forceinline__ void gather_rows(const uint32_t *hin2, int c2, const uint32_t (&ids)[kGat], int rem,
                                            uint32_t (&x)[kGat]) {
#pragma unroll
  for (int j = 0; j < kGat; j++) x[j] = hin2[(int64_t)(j < rem ? ids[j] : 0u) * 32 + c2];
}

// Adds the chunk's rows to (s0, s1) in the row's order: whole groups of 4
// as ((x0 + x1) + x2) + x3, the last < 4 one at a time.
__device__ __forceinline__ void sum_rows(const uint32_t (&x)[kGat], int rem, float &s0, float &s1) {
#pragma unroll
  for (int q = 0; q < kGat; q += 4) {
    if (q + 4 <= rem) {
      s0 += bf16_lo(x[q]) + bf16_lo(x[q + 1]) + bf16_lo(x[q + 2]) + bf16_lo(x[q + 3]);
      s1 += bf16_hi(x[q]) + bf16_hi(x[q + 1]) + bf16_hi(x[q + 2]) + bf16_hi(x[q + 3]);
    } else {
#pragma unroll
      for (int j = q; j < q + 4; j++) {
        if (j < rem) {
          s0 += bf16_lo(x[j]);
          s1 += bf16_hi(x[j]);
        }
      }
    }
  }
}

// Layer 1 (4 -> 64, FP32): one wave per 64 vertices; lane i first gathers
// vertex i's neighbour mean, then lane j computes output column j of every
// vertex of the group (coalesced 128-byte rows).
__global__ __launch_bounds__(kNB) void net_l1_fwd_kernel(DevGraph g, const float4 *x, const float *P,
                                                        float4 *xa_out, uint16_t *h1) {
  __shared__ float zs[kNB / kWave][64][9];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  float W[8];
#pragma unroll
  for (int k = 0; k < 8; k++) W[k] = P[kOffW1 + k * 64 + lane];
  const float bias = P[kOffB1 + lane];
  const int64_t n = g.n, ngroups = (n + 63) / 64;
  const int wpb = kNB / kWave;
  for (int64_t g0 = (int64_t)blockIdx.x * wpb; g0 < ngroups; g0 += (int64_t)gridDim.x * wpb) {
    const int64_t grp = g0 + w;
    const int64_t v = grp * 64 + lane;
    float4 xv = make_float4(0, 0, 0, 0), ag = make_float4(0, 0, 0, 0);
    const bool own = grp < ngroups && v < n;  // grp is wave-uniform
    uint32_t b = 0;
    int d = 0;
    if (own) {
      b = g.rowptr[v];
      d = (int)(g.rowptr[v + 1] - b);
    }
    // neighbours added one at a time in row order, loaded kGat per round
    // trip (unconditional loads, clamped past the row's end)
    const int dmax = wave_max_i32(d);
    for (int k0 = 0; k0 < dmax; k0 += kGat) {
      const int rem = d - k0;
      uint32_t ids[kGat];
      gather_ids(g.adj, b + (uint32_t)k0, rem, ids);
      float4 xu[kGat];
#pragma unroll
      for (int j = 0; j < kGat; j++) xu[j] = x[j < rem ? ids[j] : 0u];
#pragma unroll
      for (int j = 0; j < kGat; j++) {
        if (j < rem) {
          ag.x += xu[j].x;
          ag.y += xu[j].y;
          ag.z += xu[j].z;
          ag.w += xu[j].w;
        }
      }
    }
    if (own) {
      xv = x[v];
      const float inv = inv_degree(g, v);
      ag.x *= inv;
      ag.y *= inv;
      ag.z *= inv;
      ag.w *= inv;
      if (xa_out) xa_out[v] = ag;
    }
    zs[w][lane][0] = xv.x;
    zs[w][lane][1] = xv.y;
    zs[w][lane][2] = xv.z;
    zs[w][lane][3] = xv.w;
    zs[w][lane][4] = ag.x;
    zs[w][lane][5] = ag.y;
    zs[w][lane][6] = ag.z;
    zs[w][lane][7] = ag.w;
    __syncthreads();
    if (grp < ngroups) {
      const int64_t left = n - grp * 64;
      const int cnt = left < 64 ? (int)left : 64;
      for (int j = 0; j < cnt; j++) {
        float a = bias;
#pragma unroll
        for (int k = 0; k < 8; k++) a += zs[w][j][k] * W[k];
        h1[(grp * 64 + j) * 64 + lane] = bf16_bits(fmaxf(a, 0.0f));
      }
    }
    __syncthreads();
  }
}

// Hidden layer (128 -> 64 on MFMA).  Each wave owns a 32-vertex tile: the
// two half-waves gather [h_v, mean_u h_u] rows (32 lanes x 2 columns) into
// LDS as BF16, then 16 MFMAs multiply the tile by W (held transposed in LDS).
// The tile's 33 row pointers are loaded once (lane r: rowptr[v0 + r]); a
// row's first kGat neighbour ids are loaded while the previous row's
// gathers are in flight.
__global__ __launch_bounds__(kNB) void net_hidden_fwd_kernel(DevGraph g, const uint16_t *hin,
                                                            const uint16_t *Wb, const float *bias,
                                                            uint16_t *hout, uint16_t *za_out) {
  __shared__ __attribute__((aligned(16))) uint16_t Wt[64 * kAP];
  __shared__ __attribute__((aligned(16))) uint16_t As[kNB / kWave][32 * kAP];
  for (int e = threadIdx.x; e < 128 * 64; e += kNB) Wt[(e % 64) * kAP + e / 64] = Wb[e];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int half = lane / 32, c2 = lane % 32;
  const float bj0 = bias[c2], bj1 = bias[32 + c2];
  const uint32_t *hin2 = reinterpret_cast<const uint32_t *>(hin);
  uint16_t *A = As[w];
  __syncthreads();
  const int64_t n = g.n, ntiles = (n + 31) / 32;
  const int wpb = kNB / kWave;
  for (int64_t t0 = (int64_t)blockIdx.x * wpb; t0 < ntiles; t0 += (int64_t)gridDim.x * wpb) {
    const int64_t tile = t0 + w;
    if (tile < ntiles) {  // wave-uniform
      const int64_t v0 = tile * 32;
      const uint32_t rp = (lane <= 32 && v0 + lane <= n) ? g.rowptr[v0 + lane] : 0u;
      // start and degree of the tile's row r (degree 0 past the graph's end)
      auto row = [&](int r, uint32_t &b) {
        b = __shfl(rp, r);
        const uint32_t e = __shfl(rp, r + 1);
        return v0 + r < n ? (int)(e - b) : 0;
      };
      uint32_t b, ids[kGat];
      int d = row(half, b);
      gather_ids(g.adj, b, d, ids);
      for (int it = 0; it < 16; it++) {
        const int r = 2 * it + half;
        const int64_t v = v0 + r;
        uint32_t x[kGat];
        gather_rows(hin2, c2, ids, d, x);
        const uint32_t selfp = v < n ? hin2[v * 32 + c2] : 0u;
        // the next row's ids, loaded while this row's gathers are in flight
        uint32_t bn = 0, nids[kGat];
        int dn = 0;
        if (it + 1 < 16) {
          dn = row(r + 2, bn);
          gather_ids(g.adj, bn, dn, nids);
        }
        float s0 = 0, s1 = 0;
        sum_rows(x, d, s0, s1);
        const int dmax = wave_max_i32(d);
        for (int k0 = kGat; k0 < dmax; k0 += kGat) {  // rows longer than kGat
          const int rem = d - k0;
          uint32_t cids[kGat], cx[kGat];
          gather_ids(g.adj, b + (uint32_t)k0, rem, cids);
          gather_rows(hin2, c2, cids, rem, cx);
          sum_rows(cx, rem, s0, s1);
        }
        uint32_t aggp = 0;
        if (v < n) {
          const float inv = inv_degree(g, v);
          aggp = pack_bf16(s0 * inv, s1 * inv);
          if (za_out) reinterpret_cast<uint32_t *>(za_out)[v * 32 + c2] = aggp;
        }
        reinterpret_cast<uint32_t *>(A + r * kAP)[c2] = selfp;
        reinterpret_cast<uint32_t *>(A + r * kAP + 64)[c2] = aggp;
        b = bn;
        d = dn;
#pragma unroll
        for (int j = 0; j < kGat; j++) ids[j] = nids[j];
      }
    }
    __syncthreads();
    if (tile < ntiles) {
#pragma unroll
      for (int nt = 0; nt < 2; nt++) {
        floatx16 acc = {};
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
          const bf16x8 a = *reinterpret_cast<const bf16x8 *>(A + c2 * kAP + ks * 16 + half * 8);
          const bf16x8 b =
              *reinterpret_cast<const bf16x8 *>(Wt + (nt * 32 + c2) * kAP + ks * 16 + half * 8);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = nt * 32 + c2;
        const float bj = nt ? bj1 : bj0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const int64_t v = tile * 32 + 8 * (i / 4) + 4 * half + (i % 4);
          if (v < n) hout[v * 64 + col] = bf16_bits(fmaxf(acc[i] + bj, 0.0f));
        }
      }
    }
    __syncthreads();
  }
}

// Scores without a target: one lane per vertex, its 64 hidden values read
// as 8 x 16 bytes.  The 64 terms h_k * wo_k are summed in the pairing of
// net_readout_kernel's butterfly as lane 0 sees it: first
// fma(h_k, wo_k, h_{k+32} * wo_{k+32}) for k < 32 (that kernel's compiled
// first step fuses the lane's own product into the add), then k with
// k + 16, ..., k + 1 — so the scores are bit-identical to it.
__global__ __launch_bounds__(kNB) void net_score_kernel(const uint16_t *h3, const float *P, int64_t n,
                                                       float *y) {
#pragma clang fp contract(off)
  const float bo = P[kOffBo];
  float wo[64];
#pragma unroll
  for (int k = 0; k < 64; k++) wo[k] = P[kOffWo + k];
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const uint4 *row = reinterpret_cast<const uint4 *>(h3 + v * 64);
    float h[64], p[32];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const uint4 u = row[q];
      const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
      for (int i = 0; i < 4; i++) {
        h[8 * q + 2 * i] = bf16_lo(wd[i]);
        h[8 * q + 2 * i + 1] = bf16_hi(wd[i]);
      }
    }
#pragma unroll
    for (int k = 0; k < 32; k++) p[k] = __builtin_fmaf(h[k], wo[k], h[k + 32] * wo[k + 32]);
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) {
#pragma unroll
      for (int k = 0; k < o; k++) p[k] = p[k] + p[k + o];
    }
    y[v] = p[0] + bo;
  }
}

// Read-out y = h3 . wo + bo, one wave per vertex (lane = column).  With a
// target it also starts the backward pass: the loss, dL/dy, the read-out
// gradients, delta3 = dL/dy * wo * [h3 > 0] (stored BF16 for the MFMA
// backward) and b3's gradient.  Per-block partials go to part[block].
__global__ __launch_bounds__(kNB) void net_readout_kernel(const uint16_t *h3, const float *P,
                                                         const float *target, float inv_count,
                                                         int64_t n, float *y, uint16_t *dlt,
                                                         float *part, double *loss_part) {
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const float wo = P[kOffWo + lane], bo = P[kOffBo];
  float gwo = 0, gb3 = 0, gbo = 0;
  double loss = 0;
  const int64_t gw = (int64_t)blockIdx.x * (kNB / kWave) + w, nw = (int64_t)gridDim.x * (kNB / kWave);
  for (int64_t v = gw; v < n; v += nw) {
    const float h = __uint_as_float((uint32_t)h3[v * 64 + lane] << 16);
    float s = h * wo;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    const float yv = s + bo;
    if (lane == 0) y[v] = yv;
    if (target) {
      const float diff = yv - target[v];
      const float gy = 2.0f * diff * inv_count;
      const float d = h > 0.0f ? gy * wo : 0.0f;
      dlt[v * 64 + lane] = bf16_bits(d);
      gwo += gy * h;
      gb3 += d;
      if (lane == 0) {
        gbo += gy;
        loss += (double)diff * (double)diff;
      }
    }
  }
  if (!target) return;
  __shared__ float sh[kNB / kWave][130];
  __shared__ double shl[kNB / kWave];
  sh[w][lane] = gwo;
  sh[w][64 + lane] = gb3;
  if (lane == 0) {
    sh[w][128] = gbo;
    shl[w] = loss;
  }
  __syncthreads();
  if (w == 0) {
    float a = 0, b = 0;
    for (int q = 0; q < kNB / kWave; q++) {
      a += sh[q][lane];
      b += sh[q][64 + lane];
    }
    float *pb = part + (int64_t)blockIdx.x * kNetParams;
    pb[kOffWo + lane] = a;
    pb[kOffB3 + lane] = b;
    if (lane == 0) {
      float c = 0;
      double l = 0;
      for (int q = 0; q < kNB / kWave; q++) {
        c += sh[q][128];
        l += shl[q];
      }
      pb[kOffBo] = c;
      loss_part[blockIdx.x] = l;
    }
  }
}

// ------------------------------------------------------------------ backward

// Hidden layer backward on MFMA, one wave per 32-vertex tile:
//   gz = delta W^T           (32 x 128): gs = gz[:, :64], ga = gz[:, 64:] / deg
//   gW += z^T delta          (128 x 64), z = [h_{l-1}, mean h_{l-1}]
// The weight gradient accumulates in registers over all tiles of the wave and
// is reduced over the block into part[block][offW ...].
__global__ __launch_bounds__(kBB) void net_hidden_bwd_kernel(DevGraph g, const uint16_t *dlt,
                                                            const uint16_t *hin, const uint16_t *za,
                                                            const uint16_t *Wb, float *gs, float *ga,
                                                            float *part, int offW) {
  constexpr int kWaves = kBB / kWave;
  constexpr int kTileBytes = (32 * kWP + 64 * kTP + 128 * kTP) * 2;
  __shared__ __attribute__((aligned(16))) uint16_t Wl[128 * kWP];
  __shared__ __attribute__((aligned(16))) unsigned char tiles[kWaves * kTileBytes];
  for (int e = threadIdx.x; e < 128 * 64; e += kBB) Wl[(e / 64) * kWP + e % 64] = Wb[e];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  const int half = lane / 32, c2 = lane % 32;
  uint16_t *Ds = reinterpret_cast<uint16_t *>(tiles + w * kTileBytes);  // [32][kWP]  delta
  uint16_t *Dt = Ds + 32 * kWP;                                          // [64][kTP]  delta^T
  uint16_t *Zt = Dt + 64 * kTP;                                          // [128][kTP] z^T
  const uint32_t *d2 = reinterpret_cast<const uint32_t *>(dlt);
  const uint32_t *h2 = reinterpret_cast<const uint32_t *>(hin);
  const uint32_t *z2 = reinterpret_cast<const uint32_t *>(za);
  floatx16 accW[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) accW[a][b] = floatx16{};
  __syncthreads();
  const int64_t n = g.n, ntiles = (n + 31) / 32;
  for (int64_t t0 = (int64_t)blockIdx.x * kWaves; t0 < ntiles; t0 += (int64_t)gridDim.x * kWaves) {
    const int64_t tile = t0 + w;
    if (tile < ntiles) {
      for (int it = 0; it < 16; it++) {
        const int r = 2 * it + half;
        const int64_t v = tile * 32 + r;
        uint32_t d = 0, hs = 0, ag = 0;
        if (v < n) {
          d = d2[v * 32 + c2];
          hs = h2[v * 32 + c2];
          ag = z2[v * 32 + c2];
        }
        reinterpret_cast<uint32_t *>(Ds + r * kWP)[c2] = d;
        Dt[(2 * c2) * kTP + r] = (uint16_t)d;
        Dt[(2 * c2 + 1) * kTP + r] = (uint16_t)(d >> 16);
        Zt[(2 * c2) * kTP + r] = (uint16_t)hs;
        Zt[(2 * c2 + 1) * kTP + r] = (uint16_t)(hs >> 16);
        Zt[(64 + 2 * c2) * kTP + r] = (uint16_t)ag;
        Zt[(64 + 2 * c2 + 1) * kTP + r] = (uint16_t)(ag >> 16);
      }
    }
    __syncthreads();
    if (tile < ntiles) {
#pragma unroll
      for (int nt = 0; nt < 4; nt++) {
        floatx16 acc = {};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
          const bf16x8 a = *reinterpret_cast<const bf16x8 *>(Ds + c2 * kWP + ks * 16 + half * 8);
          const bf16x8 b =
              *reinterpret_cast<const bf16x8 *>(Wl + (nt * 32 + c2) * kWP + ks * 16 + half * 8);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const int col = nt * 32 + c2;
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const int64_t v = tile * 32 + 8 * (i / 4) + 4 * half + (i % 4);
          if (v < n) {
            if (nt < 2)
              gs[v * 64 + col] = acc[i];
            else
              ga[v * 64 + col - 64] = acc[i] * inv_degree(g, v);
          }
        }
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
          for (int ks = 0; ks < 2; ks++) {
            const bf16x8 a =
                *reinterpret_cast<const bf16x8 *>(Zt + (mt * 32 + c2) * kTP + ks * 16 + half * 8);
            const bf16x8 b =
                *reinterpret_cast<const bf16x8 *>(Dt + (nt * 32 + c2) * kTP + ks * 16 + half * 8);
            accW[mt][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, accW[mt][nt], 0, 0, 0);
          }
    }
    __syncthreads();
  }
  // block reduction of the waves' gW tiles (fixed order) through LDS
  float *red = reinterpret_cast<float *>(tiles);  // 128 x 64 floats fit in the tile area
  static_assert(kWaves * kTileBytes >= 128 * 64 * 4, "reduction buffer");
  for (int q = kWaves - 1; q >= 0; q--) {
    if (w == q) {
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
          for (int i = 0; i < 16; i++) {
            const int row = mt * 32 + 8 * (i / 4) + 4 * half + (i % 4);
            const int col = nt * 32 + c2;
            float *dst = &red[row * 64 + col];
            *dst = (q == kWaves - 1) ? accW[mt][nt][i] : *dst + accW[mt][nt][i];
          }
    }
    __syncthreads();
  }
  float *pb = part + (int64_t)blockIdx.x * kNetParams + offW;
  for (int e = threadIdx.x; e < 128 * 64; e += kBB) pb[e] = red[e];
}

// gh_u = gs_u + sum_{v in N(u)} ga_v  (ga already divided by deg v), then
// delta_u = gh_u * [h_prev_u > 0]; one half-wave per vertex, two columns per
// lane.  Writes delta as BF16 (next MFMA layer) and/or FP32 (layer 1) and
// the bias gradient partials part[block][offB ...].
__global__ __launch_bounds__(kNB) void net_agg_bwd_kernel(DevGraph g, const float *gs, const float *ga,
                                                         const uint16_t *hprev, uint16_t *dlt_out,
                                                         float *d32_out, float *part, int offB) {
  const int half_id = threadIdx.x / 32, c2 = threadIdx.x % 32;
  const float2 *gs2 = reinterpret_cast<const float2 *>(gs);
  const float2 *ga2 = reinterpret_cast<const float2 *>(ga);
  const uint32_t *hp2 = reinterpret_cast<const uint32_t *>(hprev);
  float gb0 = 0, gb1 = 0;
  const int64_t hw = (int64_t)blockIdx.x * (kNB / 32) + half_id, nhw = (int64_t)gridDim.x * (kNB / 32);
  for (int64_t u = hw; u < g.n; u += nhw) {
    float2 acc = gs2[u * 32 + c2];
    const uint32_t b = g.rowptr[u];
    const int d = (int)(g.rowptr[u + 1] - b);
    // kGat neighbours per round trip; added in row order as pairs
    // (t0 + t1), an odd last one alone
    for (int k0 = 0; k0 < d; k0 += kGat) {
      const int rem = d - k0;
      uint32_t ids[kGat];
      gather_ids(g.adj, b + (uint32_t)k0, rem, ids);
      float2 t[kGat];
#pragma unroll
      for (int j = 0; j < kGat; j++) t[j] = ga2[(int64_t)(j < rem ? ids[j] : 0u) * 32 + c2];
#pragma unroll
      for (int q = 0; q < kGat; q += 2) {
        if (q + 2 <= rem) {
          acc.x += t[q].x + t[q + 1].x;
          acc.y += t[q].y + t[q + 1].y;
        } else if (q < rem) {
          acc.x += t[q].x;
          acc.y += t[q].y;
        }
      }
    }
    const uint32_t hp = hp2[u * 32 + c2];
    const float d0 = bf16_lo(hp) > 0.0f ? acc.x : 0.0f;
    const float d1 = bf16_hi(hp) > 0.0f ? acc.y : 0.0f;
    if (dlt_out) reinterpret_cast<uint32_t *>(dlt_out)[u * 32 + c2] = pack_bf16(d0, d1);
    if (d32_out) reinterpret_cast<float2 *>(d32_out)[u * 32 + c2] = make_float2(d0, d1);
    gb0 += d0;
    gb1 += d1;
  }
  __shared__ float sh[kNB / 32][64];
  sh[half_id][2 * c2] = gb0;
  sh[half_id][2 * c2 + 1] = gb1;
  __syncthreads();
  if (threadIdx.x < 64) {
    float s = 0;
    for (int q = 0; q < kNB / 32; q++) s += sh[q][threadIdx.x];
    part[(int64_t)blockIdx.x * kNetParams + offB + threadIdx.x] = s;
  }
}

// gW1 = [x, A x]^T delta1 (8 x 64, FP32); lane = output column.
__global__ __launch_bounds__(kNB) void net_l1_bwd_kernel(const float4 *x, const float4 *xa,
                                                        const float *d32, int64_t n, float *part) {
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t gw = (int64_t)blockIdx.x * (kNB / kWave) + w, nw = (int64_t)gridDim.x * (kNB / kWave);
  for (int64_t v = gw; v < n; v += nw) {
    const float4 a = x[v], b = xa[v];
    const float d = d32[v * 64 + lane];
    acc[0] += a.x * d;
    acc[1] += a.y * d;
    acc[2] += a.z * d;
    acc[3] += a.w * d;
    acc[4] += b.x * d;
    acc[5] += b.y * d;
    acc[6] += b.z * d;
    acc[7] += b.w * d;
  }
  __shared__ float sh[kNB / kWave][8][64];
#pragma unroll
  for (int k = 0; k < 8; k++) sh[w][k][lane] = acc[k];
  __syncthreads();
  for (int e = threadIdx.x; e < 8 * 64; e += kNB) {
    float s = 0;
    for (int q = 0; q < kNB / kWave; q++) s += sh[q][e / 64][e % 64];
    part[(int64_t)blockIdx.x * kNetParams + kOffW1 + e] = s;
  }
}

// grad[p] = sum over blocks of part[b][p] (fixed order); loss = sum of squares.
__global__ void net_reduce_kernel(const float *part, int nblk, float *grad, const double *loss_part,
                                  double *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNetParams) {
    float s = 0;
    for (int b = 0; b < nblk; b++) s += part[(int64_t)b * kNetParams + p];
    grad[p] = s;
  }
  if (p == 0) {
    double l = 0;
    for (int b = 0; b < nblk; b++) l += loss_part[b];
    *loss = l;
  }
}

// Adam (Kingma & Ba 2015) on the FP32 parameters; refreshes the BF16
// copies of W2 and W3.
__global__ void net_adam_kernel(float *p, float *m, float *v, const float *g, float lr, float b1,
                                float b2, float eps, float c1, float c2, uint16_t *wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kNetParams) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  if (i >= kOffW2 && i < kOffW2 + 128 * 64) wb[i - kOffW2] = bf16_bits(pi);
  if (i >= kOffW3 && i < kOffW3 + 128 * 64) wb[128 * 64 + i - kOffW3] = bf16_bits(pi);
}

__global__ void net_bf16_copy_kernel(const float *p, uint16_t *wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 128 * 64) {
    wb[i] = bf16_bits(p[kOffW2 + i]);
    wb[128 * 64 + i] = bf16_bits(p[kOffW3 + i]);
  }
}

// ------------------------------------------------------------------ selection

constexpr int kSelUnroll = 4;  // vertices per thread and round trip in the argmax pass
constexpr int kSelGroup = 8;  // neighbours a lane claims per round trip
constexpr int kSelRound = 8;  // picks per round of the selection

struct SelArgs {
  DevGraph g;
  const float *score;
  int32_t *sd;                // distance to the nearest pick (exact up to r)
  uint32_t *q0, *q1;          // frontier queues
  unsigned int *cnt;          // three rotating frontier counters
  unsigned long long *best;   // one argmax slot per pick (zero at launch)
  int64_t *picks;
  int32_t *dummy;  // one scratch word per thread: target of masked-off claims
  const int32_t *d12;  // the two feature traversals' levels (n each)
  int first, K, r, backoff;
  // multi-hop bounded BFS (hops > 1): 16-byte entries in q0 / q1, counts
  // in the high halves of cnt64[3]; *overflow set when the queue would
  // overflow (the caller repeats the selection level by level)
  int hops, hop_batch;
  uint32_t qcap;  // queue entries usable (0: n + 1)
  unsigned long long *cnt64;
  unsigned int *overflow;
  // one_round: the launch makes the picks of one round, writes their
  // number to *nround and returns (the caller lowers sd from them)
  int one_round;
  int *nround;
};

__device__ __forceinline__ uint32_t orderable(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(v, o);
    v = t > v ? t : v;
  }
  return v;
}

// Lower sd[w] to `nd` when that is an improvement; true for the one claim
// that made it (the vertex joins the next frontier).
__device__ __forceinline__ bool sel_claim(int32_t *sd, uint32_t w, int nd) {
  if (ld_dev(&sd[w]) <= nd) return false;
  return atomicMin(&sd[w], nd) > nd;
}

__device__ __forceinline__ void sel_push(bool push, uint32_t v, uint32_t *qn, unsigned int *cn) {
  const unsigned long long mask = __ballot(push);
  if (!mask) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long)mask) - 1;
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(cn, (unsigned int)__popcll(mask));
  base = __shfl(base, leader);
  if (push) st_dev(&qn[base + __popcll(mask & lanemask_lt())], v);
}

// All picks in one persistent launch (every block resident), in rounds of up to kSelRound picks.
// Within a round each pick maximises (min(sd, r+1), score, -index) over the
// vertices that are provably more than r hops from the round's earlier
// picks — max(|d1_v - d1_c|, |d2_v - d2_c|) > r, a lower bound of the hop
// distance by the triangle inequality — where sd is the distance to the
// picks of earlier rounds, exact within r hops.  A round ends early when no
// such vertex is more than r hops from every earlier pick (a round's first
// pick is always taken: then it is the vertex farthest from them).  After
// each round one BFS from all its picks, bounded to r levels, lowers sd; it
// only enters vertices whose sd it lowers (a path through a vertex already
// at least as close to another pick cannot lower anything behind it).
__global__ __launch_bounds__(256) void select_kernel(SelArgs a, unsigned int *arrive, unsigned int *err) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gsize = (int64_t)gridDim.x * blockDim.x;
  const int lane = __lane_id();
  const int64_t gwave = gtid / kWave, nwaves = gsize / kWave;
  const unsigned int nb = gridDim.x;
  unsigned int bar = 0;
  const int64_t n = a.g.n;
  const uint32_t *rowptr = a.g.rowptr, *adj = a.g.adj;
  const uint32_t rcap = (uint32_t)a.r + 1u;
  const int32_t *d1 = a.d12, *d2 = a.d12 + n;
  unsigned int G = 0;  // running level counter: queue / counter rotation
  int i = a.first;
  while (i < a.K) {
    // ---- picks of one round
    int rp[kSelRound], r1[kSelRound], r2[kSelRound];
    int nr = 0;
    while (nr < kSelRound && i < a.K) {
      // kSelUnroll vertices per thread and round trip: all their loads
      // issued before any is used
      unsigned long long best = 0;
      for (int64_t v0 = gtid; v0 < n; v0 += gsize * kSelUnroll) {
        int x1[kSelUnroll], x2[kSelUnroll];
        uint32_t sdv[kSelUnroll], oid[kSelUnroll];
        float scv[kSelUnroll];
#pragma unroll
        for (int u = 0; u < kSelUnroll; u++) {
          const int64_t v = v0 + (int64_t)u * gsize;
          const uint32_t w = v < n ? (uint32_t)v : 0u;
          x1[u] = d1[w];
          x2[u] = d2[w];
          sdv[u] = (uint32_t)ld_dev(&a.sd[w]);
          scv[u] = a.score[w];
          oid[u] = original_id(a.g, w);
        }
#pragma unroll
        for (int u = 0; u < kSelUnroll; u++) {
          bool far = v0 + (int64_t)u * gsize < n;
          for (int j = 0; j < nr; j++) far = far && max(abs(x1[u] - r1[j]), abs(x2[u] - r2[j])) > a.r;
          if (!far) continue;
          const uint32_t hi = (min(sdv[u], rcap) << 21) | (orderable(scv[u]) >> 11);
          const unsigned long long key = ((unsigned long long)hi << 32) | (0xffffffffull - (uint64_t)oid[u]);
          best = key > best ? key : best;
        }
      }
      best = wave_max_u64(best);
      if (lane == 0 && best) atomicMax(&a.best[i], best);
      grid_barrier(arrive, nb * ++bar, err, a.backoff);
      if (ld_dev(err)) return;
      const unsigned long long bk = ld_dev(&a.best[i]);
      const bool avail = (uint32_t)(bk >> 53) == rcap;  // more than r hops from every earlier pick
      if (nr > 0 && !avail) {                           // (bk == 0: nothing far enough either)
        if (gtid == 0) st_dev(&a.best[i], 0ull);        // slot i is argmaxed again next round
        break;
      }
      const uint32_t po = (uint32_t)(0xffffffffull - (bk & 0xffffffffull));  // caller's id
      const uint32_t p = internal_id(a.g, po);
      rp[nr] = (int)p;
      r1[nr] = d1[p];
      r2[nr] = d2[p];
      nr++;
      if (gtid == 0) a.picks[i] = po;
      i++;
    }
    if (a.one_round) {
      if (gtid == 0) st_dev(a.nround, nr);
      return;
    }
    // ---- bounded BFS from the round's picks
    if (a.hops > 1) {
      // several hops per grid barrier (hops.h): the same sd at the end
      QEnt *hq[2] = {reinterpret_cast<QEnt *>(a.q0), reinterpret_cast<QEnt *>(a.q1)};
      if (gtid == 0) {
        for (int j = 0; j < nr; j++) {
          const uint32_t p = (uint32_t)rp[j];
          st_dev(&a.sd[p], 0);
          st_qent(&hq[G & 1][j], QEnt{p, rowptr[p], rowptr[p + 1], 0u});
        }
        st_dev(&a.cnt64[G % 3], (unsigned long long)nr << 32);
        st_dev(&a.cnt64[(G + 1) % 3], 0ull);
      }
      grid_barrier(arrive, nb * ++bar, err, a.backoff);
      if (ld_dev(err)) return;
      HopCtx h;
      h.g = a.g;
      h.lv = a.sd;
      h.col = 0;
      h.lvl_mask = 0xffffffffu;
      h.dummy = a.dummy + gtid;
      h.overflow = a.overflow;
      h.bound = a.r;
      h.cap = (a.qcap > 0 && a.qcap < (uint64_t)(n + 1)) ? a.qcap : (uint32_t)(n + 1);
      h.hops = a.hops;
      h.batch = a.hop_batch;
      h.wrank = (int64_t)(threadIdx.x / kWave) * gridDim.x + blockIdx.x;
      h.nwaves = nwaves;
      __shared__ QEnt locq[2][256 / kWave][kWave];
      __shared__ unsigned int sh_nq;
      for (;;) {
        if (threadIdx.x == 0) sh_nq = (unsigned int)(ld_dev(&a.cnt64[G % 3]) >> 32);
        __syncthreads();
        const unsigned int nq = sh_nq;
        __syncthreads();
        if (nq == 0) break;
        if (gtid == 0) st_dev(&a.cnt64[(G + 2) % 3], 0ull);
        hop_round(h, hq[G & 1], nq < h.cap ? nq : h.cap, hq[(G + 1) & 1], &a.cnt64[(G + 1) % 3], &locq[0][0][0]);
        grid_barrier(arrive, nb * ++bar, err, a.backoff);
        G++;
        if (ld_dev(err)) return;
      }
      G++;  // the next round's picks go to a fresh queue
      continue;
    }
    uint32_t *q0 = (G & 1) ? a.q1 : a.q0;
    if (gtid == 0) {
      for (int j = 0; j < nr; j++) {
        st_dev(&a.sd[rp[j]], 0);
        st_dev(&q0[j], (uint32_t)rp[j]);
      }
      st_dev(&a.cnt[G % 3], (unsigned int)nr);
      st_dev(&a.cnt[(G + 1) % 3], 0u);
    }
    grid_barrier(arrive, nb * ++bar, err, a.backoff);
    if (ld_dev(err)) return;
    for (int l = 0; l < a.r; l++) {
      if (gtid == 0) st_dev(&a.cnt[(G + 2) % 3], 0u);
      const int nd = l + 1;
      uint32_t *qn = (G & 1) ? a.q0 : a.q1;
      unsigned int *cn = &a.cnt[(G + 1) % 3];
      const uint32_t *qc = (G & 1) ? a.q1 : a.q0;
      const unsigned int nq = ld_dev(&a.cnt[G % 3]);
      for (int64_t base = gwave * kWave; base < (int64_t)nq; base += nwaves * kWave) {
        const int64_t idx = base + lane;
        const bool act = idx < (int64_t)nq;
        const uint32_t u = act ? ld_dev(&qc[idx]) : 0u;
        const uint32_t b = act ? rowptr[u] : 0u, e = act ? rowptr[u + 1] : 0u;
        const int d = (int)(e - b);
        const bool big = d >= 32;
        unsigned long long bmask = __ballot(big);
        while (bmask) {
          const int l2 = __ffsll((long long)bmask) - 1;
          bmask &= bmask - 1;
          const uint32_t bb = __shfl(b, l2), be = __shfl(e, l2);
          for (uint32_t k0 = bb; k0 < be; k0 += kWave) {
            const uint32_t k = k0 + lane;
            bool push = false;
            uint32_t w = 0;
            if (k < be) {
              w = adj[k];
              push = sel_claim(a.sd, w, nd);
            }
            sel_push(push, w, qn, cn);
          }
        }
        // small rows: kSelGroup neighbours per lane per round trip, loads
        // and claims issued before use, masked-off claims to scratch
        const int ds = big ? 0 : d;
        const int maxd = wave_max_i32(ds);
        int32_t *dum = a.dummy + gtid;
        for (int t0 = 0; t0 < maxd; t0 += kSelGroup) {
          uint32_t vv[kSelGroup];
          int32_t cur[kSelGroup];
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) vv[j] = adj[(t0 + j < ds) ? b + t0 + j : 0u];
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) vv[j] = (t0 + j < ds) ? vv[j] : 0u;
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) cur[j] = ld_dev(&a.sd[vv[j]]);
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) cur[j] = (t0 + j < ds && cur[j] > nd) ? 1 : 0;
          // every claim issued before any result is used (a claim under a
          // lane condition would wait for its result in that branch)
          int32_t old[kSelGroup];
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) old[j] = atomicMin(cur[j] ? &a.sd[vv[j]] : dum, nd);
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) cur[j] = (cur[j] && old[j] > nd) ? 1 : 0;
          int npush = 0;
#pragma unroll
          for (int j = 0; j < kSelGroup; j++) npush += cur[j];
          int total = 0;
          const int off = wave_excl_scan(npush, &total);
          if (total) {
            unsigned int base2 = 0;
            if (lane == 0) base2 = atomicAdd(cn, (unsigned int)total);
            base2 = __shfl(base2, 0) + (unsigned int)off;
#pragma unroll
            for (int j = 0; j < kSelGroup; j++)
              if (cur[j]) st_dev(&qn[base2++], vv[j]);
          }
        }
      }
      grid_barrier(arrive, nb * ++bar, err, a.backoff);
      G++;
      if (ld_dev(err)) return;
      if (ld_dev(&a.cnt[G % 3]) == 0) break;
    }
    G++;  // the next round's picks go to a fresh queue
  }
}

// Per column: exact 128-bit sum of squares into out[2 col], out[2 col + 1];
// unreached entries counted into out[2 k].
__global__ void sumsq_cols_kernel(const int32_t *D, int64_t n, unsigned long long *out) {
  const int col = blockIdx.y;
  const int32_t *c = D + (int64_t)col * n;
  U128 s{0, 0};
  unsigned long long neg = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t d = c[i];
    if (d < 0 || d >= 0x3fffffff)  // -1 (BFS) or 2^30-1 (SSSP): unreached
      neg++;
    else
      u128_add(s, (unsigned long long)d * (unsigned long long)d);
  }
  s = wave_sum_u128(s);
  neg = wave_sum_u64(neg);
  if (__lane_id() == 0) {
    atomic_add_u128(out + 2 * col, s);
    if (neg) atomicAdd(&out[2 * gridDim.y], neg);
  }
}

// ------------------------------------------------------------------ host side

static int grow(void **p, size_t bytes) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  PHDE_HIP(hipMalloc(p, std::max<size_t>(bytes, 16)));
  return PHDE_OK;
}

int pivot_features(phde_ctx *c, int64_t first, int64_t *second, int *ecc1, int *ecc2) {
  const int64_t n = c->g.n;
  if (!c->lp_feat) {
    PHDE_HIP(hipMalloc(&c->lp_feat, sizeof(float4) * n));
    PHDE_HIP(hipMalloc(&c->lp_d12, sizeof(int32_t) * 2 * n));
    PHDE_HIP(hipMalloc(&c->lp_sd, sizeof(int32_t) * n));
    PHDE_HIP(hipMalloc(&c->lp_score, sizeof(float) * n));
  }
  if (first < 0 || first >= n) {
    set_error("first pivot out of range");
    return PHDE_ERR_ARG;
  }
  int32_t *d1 = c->lp_d12, *d2 = c->lp_d12 + n;
  PHDE_TRY_RC(fill_i32(c, c->lp_sd, n, INT32_MAX));
  PHDE_TRY_RC(bfs_run(c, to_internal(c, first), 15, 18, d1, c->lp_sd, true));
  KcResult r1 = *c->h_kc;
  if (r1.unreached) {
    set_error("graph is not connected: " + std::to_string(r1.unreached) + " vertices unreached from " +
              std::to_string(first));
    return PHDE_ERR_UNREACHED;
  }
  const int64_t p2 = (int64_t)(0xffffffffull - (r1.key & 0xffffffffull));  // caller's id
  PHDE_TRY_RC(bfs_run(c, to_internal(c, p2), 15, 18, d2, c->lp_sd, true));
  const KcResult r2 = *c->h_kc;
  *second = p2;
  *ecc1 = (int)r1.depth;
  *ecc2 = (int)r2.depth;
  const float inv1 = r1.depth > 0 ? 1.0f / (float)r1.depth : 0.0f;
  const float inv2 = r2.depth > 0 ? 1.0f / (float)r2.depth : 0.0f;
  const int nb = std::max(1, std::min<int>((int)((n + 255) / 256), c->num_cus * 8));
  feature_kernel<<<nb, 256, 0, c->stream>>>(c->g, d1, d2, inv1, inv2, (float4 *)c->lp_feat);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

static int net_workspace(phde_mpnn *net, int64_t n) {
  if (n <= net->cap) return PHDE_OK;
  PHDE_TRY_RC(grow((void **)&net->xa, sizeof(float4) * n));
  uint16_t **hb[] = {&net->h1, &net->h2, &net->h3, &net->za2, &net->za3, &net->dlt};
  for (auto p : hb) PHDE_TRY_RC(grow((void **)p, sizeof(uint16_t) * 64 * n));
  float **fb[] = {&net->gs, &net->ga, &net->d32};
  for (auto p : fb) PHDE_TRY_RC(grow((void **)p, sizeof(float) * 64 * n));
  PHDE_TRY_RC(grow((void **)&net->y, sizeof(float) * n));
  net->cap = n;
  return PHDE_OK;
}

static int net_forward(phde_mpnn *net, phde_ctx *c, bool train) {
  const int64_t n = c->g.n;
  PHDE_TRY_RC(net_workspace(net, n));
  const int nb = net->nblk;
  const float4 *x = (const float4 *)c->lp_feat;
  net_l1_fwd_kernel<<<nb, kNB, 0, c->stream>>>(c->g, x, net->p, train ? net->xa : nullptr, net->h1);
  // the hidden layers' LDS (51 KB per block) leaves room for 3 blocks per
  // CU, and their gathers are latency-bound: fill all three
  const int nbh = nb / 2 * 3;
  net_hidden_fwd_kernel<<<nbh, kNB, 0, c->stream>>>(c->g, net->h1, net->wb, net->p + kOffB2, net->h2,
                                                   train ? net->za2 : nullptr);
  net_hidden_fwd_kernel<<<nbh, kNB, 0, c->stream>>>(c->g, net->h2, net->wb + 128 * 64, net->p + kOffB3,
                                                   net->h3, train ? net->za3 : nullptr);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int net_score(phde_ctx *c, phde_mpnn *net) {
  if (!c->lp_feat) {
    set_error("pivot features not computed");
    return PHDE_ERR_ARG;
  }
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_OTHER, &ev);
  PHDE_TRY_RC(net_forward(net, c, false));
  const int64_t n = c->g.n;
  const int nbs = (int)std::max<int64_t>(1, std::min<int64_t>((n + kNB - 1) / kNB, (int64_t)net->nblk * 8));
  net_score_kernel<<<nbs, kNB, 0, c->stream>>>(net->h3, net->p, n, c->lp_score);
  PHDE_HIP(hipGetLastError());
  c->timer.end(c->stream, PHDE_KERNEL_OTHER, ev);
  return PHDE_OK;
}

// Forward + backward on the context's graph and training target; leaves the
// gradient in net->g and the summed squared error in net->loss (device).
static int net_backward(phde_mpnn *net, phde_ctx *c) {
  if (!c->lp_feat || !c->lp_target) {
    set_error("training data not set on this context");
    return PHDE_ERR_ARG;
  }
  const int64_t n = c->g.n;
  const int nb = net->nblk;
  PHDE_TRY_RC(net_forward(net, c, true));
  net_readout_kernel<<<nb, kNB, 0, c->stream>>>(net->h3, net->p, c->lp_target,
                                               1.0f / (float)std::max<int64_t>(1, n), n, net->y, net->dlt,
                                               net->part, net->loss_part);
  // layer 3
  net_hidden_bwd_kernel<<<nb, kBB, 0, c->stream>>>(c->g, net->dlt, net->h2, net->za3,
                                                  net->wb + 128 * 64, net->gs, net->ga, net->part, kOffW3);
  net_agg_bwd_kernel<<<nb, kNB, 0, c->stream>>>(c->g, net->gs, net->ga, net->h2, net->dlt, nullptr,
                                               net->part, kOffB2);
  // layer 2
  net_hidden_bwd_kernel<<<nb, kBB, 0, c->stream>>>(c->g, net->dlt, net->h1, net->za2, net->wb, net->gs,
                                                  net->ga, net->part, kOffW2);
  net_agg_bwd_kernel<<<nb, kNB, 0, c->stream>>>(c->g, net->gs, net->ga, net->h1, nullptr, net->d32,
                                               net->part, kOffB1);
  // layer 1
  net_l1_bwd_kernel<<<nb, kNB, 0, c->stream>>>((const float4 *)c->lp_feat, net->xa, net->d32, n,
                                              net->part);
  net_reduce_kernel<<<(kNetParams + 255) / 256, 256, 0, c->stream>>>(net->part, nb, net->g,
                                                                    net->loss_part, net->loss);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int greedy_select(phde_ctx *c, const float *score_dev, int first, int K, int r, int64_t *picks) {
  const int64_t n = c->g.n;
  if (K <= first) return PHDE_OK;
  PHDE_TRY_RC(bfs_setup(c));
  DBuf<unsigned long long> best;
  DBuf<int64_t> pk;
  PHDE_TRY_RC(best.alloc(K));
  PHDE_TRY_RC(pk.alloc(K));
  PHDE_HIP(hipMemsetAsync(best.p, 0, sizeof(unsigned long long) * K, c->stream));
  SelArgs a;
  a.g = c->g;
  a.score = score_dev;
  a.sd = c->lp_sd;
  // the BFS queues of the context hold n 16-byte entries each: reuse them
  a.q0 = c->d_q[0];
  a.q1 = c->d_q[1];
  a.cnt = reinterpret_cast<unsigned int *>(c->d_slots);
  a.best = best.p;
  a.picks = pk.p;
  a.dummy = c->d_dummy;
  a.d12 = c->lp_d12;
  a.first = first;
  a.K = K;
  a.r = std::max(1, std::min(r, 2046));
  a.backoff = c->bar_backoff;
  a.hops = c->bfs_hops;
  a.hop_batch = c->bfs_hop_batch;
  a.qcap = c->bfs_qcap;
  a.cnt64 = reinterpret_cast<unsigned long long *>(c->d_slots) + 4;
  a.overflow = c->d_bar + 2;
  a.one_round = 0;
  a.nround = nullptr;
  unsigned int *arrive = c->d_bar, *err = c->d_bar + 1;
  void *args[] = {&a, &arrive, &err};
  if (pbfs_seeds_ok(c)) {
    // round by round: the kernel makes a round's picks, then one search
    // from all of them together (the boundary-distance method, bounded at
    // r) lowers sd to the exact distance to the nearest pick within r
    // hops, and elsewhere leaves it above r or as it was -- as the
    // kernel's own bounded search does up to values above r, which a key
    // does not tell apart (it takes min(sd, r + 1)).
    DBuf<int32_t> col;
    DBuf<int> nr_dev;
    PHDE_TRY_RC(col.alloc(n));
    PHDE_TRY_RC(nr_dev.alloc(1));
    a.one_round = 1;
    a.nround = nr_dev.p;
    std::vector<int64_t> h(K);
    for (int i = first; i < K;) {
      a.first = i;
      PHDE_TRY_RC(fill_u32(c, c->d_bar, 3, 0u));
      hipEvent_t ev;
      c->timer.begin(c->stream, PHDE_KERNEL_OTHER, &ev);
      const hipError_t e = launch_resident((const void *)select_kernel, c->bfs_grid, 256, args, 0, c->stream);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string("pivot selection needs a persistent kernel: ") + hipGetErrorString(e));
        return kSelectUnavailable;
      }
      c->timer.end(c->stream, PHDE_KERNEL_OTHER, ev);
      int nr = 0;
      unsigned int herr = 0;
      // the count, the flag and every pick slot from here on, in one wait
      PHDE_HIP(hipMemcpyAsync(&nr, nr_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      PHDE_HIP(hipMemcpyAsync(&herr, err, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
      PHDE_HIP(hipMemcpyAsync(h.data() + i, pk.p + i, sizeof(int64_t) * (K - i), hipMemcpyDeviceToHost, c->stream));
      PHDE_HIP(hipStreamSynchronize(c->stream));
      if (herr) {
        set_error("pivot selection grid barrier timed out");
        return kSelectUnavailable;
      }
      if (nr < 1 || nr > K - i) {
        set_error("pivot selection made no pick");
        return PHDE_ERR_DEVICE;
      }
      std::vector<int64_t> seeds(nr);
      for (int j = 0; j < nr; j++) seeds[j] = to_internal(c, h[i + j]);
      bool ran = false;
      PHDE_TRY_RC(pbfs_seeds_run(c, seeds.data(), nr, col.p, c->lp_sd, a.r, &ran));
      if (!ran) {
        // the method stopped: a search from each pick (with the fused
        // min into sd) gives the same sd
        for (int j = 0; j < nr; j++) PHDE_TRY_RC(bfs_run(c, seeds[j], 15, 18, col.p, c->lp_sd, true));
      }
      i += nr;
    }
    for (int i = first; i < K; i++) picks[i] = h[i];
    return PHDE_OK;
  }
  // the multi-hop search can (rarely) outgrow its queue: keep the state it
  // starts from, to repeat the selection level by level then
  DBuf<int32_t> sd0;
  if (a.hops > 1) {
    PHDE_TRY_RC(sd0.alloc(n));
    PHDE_HIP(hipMemcpyAsync(sd0.p, c->lp_sd, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  PHDE_TRY_RC(fill_u32(c, c->d_bar, 3, 0u));
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_OTHER, &ev);
  const hipError_t e = launch_resident((const void *)select_kernel, c->bfs_grid, 256, args, 0, c->stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error(std::string("pivot selection needs a persistent kernel: ") + hipGetErrorString(e));
    return kSelectUnavailable;
  }
  c->timer.end(c->stream, PHDE_KERNEL_OTHER, ev);
  std::vector<int64_t> h(K);
  unsigned int herr[2] = {0u, 0u};
  PHDE_HIP(hipMemcpyAsync(h.data(), pk.p, sizeof(int64_t) * K, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(herr, err, 2 * sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (herr[0]) {
    // blocks kept from becoming resident: the caller selects without the
    // persistent kernel
    set_error("pivot selection grid barrier timed out");
    return kSelectUnavailable;
  }
  if (a.hops > 1 && herr[1]) {
    PHDE_HIP(hipMemcpyAsync(c->lp_sd, sd0.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, c->stream));
    const int saved = c->bfs_hops;
    c->bfs_hops = 0;
    const int rc = greedy_select(c, score_dev, first, K, r, picks);
    c->bfs_hops = saved;
    return rc;
  }
  for (int i = first; i < K; i++) picks[i] = h[i];
  return PHDE_OK;
}

int column_sumsq(phde_ctx *c, const int32_t *D, int k, double *sumsq, unsigned long long *unreached) {
  DBuf<unsigned long long> acc;
  PHDE_TRY_RC(acc.alloc(2 * (size_t)k + 1));
  PHDE_HIP(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long) * (2 * k + 1), c->stream));
  const int bx = std::max(1, std::min<int>((int)((c->g.n + 255) / 256), std::max(1, c->num_cus * 4 / k)));
  sumsq_cols_kernel<<<dim3(bx, k), 256, 0, c->stream>>>(D, c->g.n, acc.p);
  PHDE_HIP(hipGetLastError());
  std::vector<unsigned long long> h(2 * (size_t)k + 1);
  PHDE_HIP(hipMemcpyAsync(h.data(), acc.p, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < k; i++) sumsq[i] = u128_to_double(&h[2 * (size_t)i]);
  if (unreached) *unreached = h[2 * (size_t)k];
  return PHDE_OK;
}

void learned_release(phde_ctx *c) {
  void *p[] = {c->lp_feat, c->lp_d12, c->lp_sd, c->lp_score, c->lp_target};
  for (void *x : p)
    if (x) (void)hipFree(x);
}

}  // namespace phde

// ------------------------------------------------------------------ C ABI

using namespace phde;

static void net_free(phde_mpnn *net) {
  void *p[] = {net->p, net->m, net->v, net->g, net->wb, net->part, net->loss_part, net->loss,
               net->xa, net->h1, net->h2, net->h3, net->za2, net->za3, net->dlt, net->gs,
               net->ga, net->d32, net->y};
  for (void *x : p)
    if (x) (void)hipFree(x);
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
}

// Glorot-uniform initial weights from a splitmix64 stream; zero biases.
static void net_init_host(uint64_t seed, std::vector<float> &p) {
  p.assign(kNetParams, 0.0f);
  uint64_t s = seed ^ 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int fan_in, int fan_out, int count) {
    const double lim = std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < count; i++) p[off + i] = (float)((2.0 * next() - 1.0) * lim);
  };
  fill(kOffW1, 8, 64, 8 * 64);
  fill(kOffW2, 128, 64, 128 * 64);
  fill(kOffW3, 128, 64, 128 * 64);
  fill(kOffWo, 64, 1, 64);
}

static int net_upload(phde_mpnn *net, const float *host) {
  PHDE_HIP(hipMemcpyAsync(net->p, host, sizeof(float) * kNetParams, hipMemcpyHostToDevice, net->stream));
  PHDE_HIP(hipMemsetAsync(net->m, 0, sizeof(float) * kNetParams, net->stream));
  PHDE_HIP(hipMemsetAsync(net->v, 0, sizeof(float) * kNetParams, net->stream));
  net_bf16_copy_kernel<<<(128 * 64 + 255) / 256, 256, 0, net->stream>>>(net->p, net->wb);
  PHDE_HIP(hipGetLastError());
  PHDE_HIP(hipStreamSynchronize(net->stream));
  net->step = 0;
  return PHDE_OK;
}

extern "C" {

int phde_net_num_params(void) { return kNetParams; }

int phde_net_create(int device, uint64_t seed, phde_mpnn **out) {
  if (!out) return PHDE_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no GPU available");
    return PHDE_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return PHDE_ERR_ARG;
  PHDE_HIP(hipSetDevice(device));
  phde_mpnn *net = new phde_mpnn();
  net->device = device;
  int cus = 0;
  auto fail = [&](int rc) {
    net_free(net);
    return rc;
  };
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return fail(PHDE_ERR_DEVICE);
  net->nblk = std::max(1, cus) * 2;
  if (hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&net->p, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&net->m, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&net->v, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&net->g, sizeof(float) * kNetParams) != hipSuccess ||
      hipMalloc(&net->wb, sizeof(uint16_t) * 2 * 128 * 64) != hipSuccess ||
      hipMalloc(&net->part, sizeof(float) * kNetParams * (size_t)net->nblk) != hipSuccess ||
      hipMalloc(&net->loss_part, sizeof(double) * net->nblk) != hipSuccess ||
      hipMalloc(&net->loss, sizeof(double)) != hipSuccess) {
    set_error("pivot network: device allocation failed");
    return fail(PHDE_ERR_NOMEM);
  }
  std::vector<float> p;
  net_init_host(seed, p);
  const int rc = net_upload(net, p.data());
  if (rc) return fail(rc);
  *out = net;
  return PHDE_OK;
}

int phde_net_destroy(phde_mpnn *net) {
  if (!net) return PHDE_ERR_ARG;
  (void)hipSetDevice(net->device);
  net_free(net);
  return PHDE_OK;
}

int phde_net_get_params(phde_mpnn *net, float *out) {
  if (!net || !out) return PHDE_ERR_ARG;
  PHDE_HIP(hipMemcpy(out, net->p, sizeof(float) * kNetParams, hipMemcpyDeviceToHost));
  return PHDE_OK;
}

int phde_net_set_params(phde_mpnn *net, const float *in) {
  if (!net || !in) return PHDE_ERR_ARG;
  return net_upload(net, in);
}

// File format: 8-byte magic "PHDENET1", int32 parameter count, then the
// FP32 parameters (little endian) in the order of learned.h.
int phde_net_save(phde_mpnn *net, const char *path) {
  if (!net || !path) return PHDE_ERR_ARG;
  std::vector<float> p(kNetParams);
  PHDE_TRY_RC(phde_net_get_params(net, p.data()));
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cannot write ") + path);
    return PHDE_ERR_ARG;
  }
  const int32_t cnt = kNetParams;
  const bool ok = std::fwrite("PHDENET1", 1, 8, f) == 8 && std::fwrite(&cnt, 4, 1, f) == 1 &&
                  std::fwrite(p.data(), 4, kNetParams, f) == (size_t)kNetParams;
  std::fclose(f);
  if (!ok) {
    set_error(std::string("short write to ") + path);
    return PHDE_ERR_ARG;
  }
  return PHDE_OK;
}

int phde_net_load(phde_mpnn *net, const char *path) {
  if (!net || !path) return PHDE_ERR_ARG;
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    set_error(std::string("cannot read ") + path);
    return PHDE_ERR_ARG;
  }
  char magic[8];
  int32_t cnt = 0;
  std::vector<float> p(kNetParams);
  const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "PHDENET1", 8) == 0 &&
                  std::fread(&cnt, 4, 1, f) == 1 && cnt == kNetParams &&
                  std::fread(p.data(), 4, kNetParams, f) == (size_t)kNetParams;
  std::fclose(f);
  if (!ok) {
    set_error(std::string("not a pivot network file: ") + path);
    return PHDE_ERR_ARG;
  }
  return net_upload(net, p.data());
}

int phde_set_pivot_net(phde_ctx *c, phde_mpnn *net) {
  if (!c) return PHDE_ERR_ARG;
  if (net && net->device != c->device) {
    set_error("pivot network lives on another device");
    return PHDE_ERR_ARG;
  }
  c->net = net;
  return PHDE_OK;
}

int phde_pivot_features(phde_ctx *c, int64_t first, float *feat, int64_t *sources) {
  if (!c) return PHDE_ERR_ARG;
  int64_t p2 = 0;
  int e1 = 0, e2 = 0;
  PHDE_TRY_RC(pivot_features(c, first, &p2, &e1, &e2));
  DBuf<float4> fo;  // rows in the caller's numbering
  if (feat && renumbered(c)) {
    PHDE_TRY_RC(fo.alloc(c->g.n));
    PHDE_TRY_RC(rows_to_original(c, c->lp_feat, fo.p, 1, sizeof(float4)));
  }
  if (feat)
    PHDE_HIP(hipMemcpyAsync(feat, fo.p ? (const void *)fo.p : (const void *)c->lp_feat,
                            sizeof(float) * 4 * c->g.n, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (sources) {
    sources[0] = first;
    sources[1] = p2;
  }
  return PHDE_OK;
}

int phde_set_training_data(phde_ctx *c, const float *feat, const float *target) {
  if (!c || !feat || !target) return PHDE_ERR_ARG;
  const int64_t n = c->g.n;
  if (!c->lp_feat) {
    PHDE_HIP(hipMalloc(&c->lp_feat, sizeof(float4) * n));
    PHDE_HIP(hipMalloc(&c->lp_d12, sizeof(int32_t) * 2 * n));
    PHDE_HIP(hipMalloc(&c->lp_sd, sizeof(int32_t) * n));
    PHDE_HIP(hipMalloc(&c->lp_score, sizeof(float) * n));
  }
  if (!c->lp_target) PHDE_HIP(hipMalloc(&c->lp_target, sizeof(float) * n));
  if (renumbered(c)) {
    DBuf<float4> fi;
    DBuf<float> ti;
    PHDE_TRY_RC(fi.alloc(n));
    PHDE_TRY_RC(ti.alloc(n));
    PHDE_HIP(hipMemcpyAsync(fi.p, feat, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c->stream));
    PHDE_HIP(hipMemcpyAsync(ti.p, target, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    PHDE_TRY_RC(rows_to_internal(c, fi.p, c->lp_feat, 1, sizeof(float4)));
    PHDE_TRY_RC(rows_to_internal(c, ti.p, c->lp_target, 1, sizeof(float)));
    PHDE_HIP(hipStreamSynchronize(c->stream));
    return PHDE_OK;
  }
  PHDE_HIP(hipMemcpyAsync(c->lp_feat, feat, sizeof(float) * 4 * n, hipMemcpyHostToDevice, c->stream));
  PHDE_HIP(hipMemcpyAsync(c->lp_target, target, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

int phde_net_scores(phde_mpnn *net, phde_ctx *c, float *score) {
  if (!net || !c || !score) return PHDE_ERR_ARG;
  PHDE_TRY_RC(net_score(c, net));
  DBuf<float> so;  // rows in the caller's numbering
  if (renumbered(c)) {
    PHDE_TRY_RC(so.alloc(c->g.n));
    PHDE_TRY_RC(rows_to_original(c, c->lp_score, so.p, 1, sizeof(float)));
  }
  PHDE_HIP(hipMemcpyAsync(score, so.p ? so.p : c->lp_score, sizeof(float) * c->g.n, hipMemcpyDeviceToHost,
                          c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

int phde_net_gradient(phde_mpnn *net, phde_ctx *c, double *loss, float *grad) {
  if (!net || !c) return PHDE_ERR_ARG;
  PHDE_TRY_RC(net_backward(net, c));
  double l = 0;
  if (grad)
    PHDE_HIP(hipMemcpyAsync(grad, net->g, sizeof(float) * kNetParams, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(&l, net->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (loss) *loss = l / (double)std::max<int64_t>(1, c->g.n);
  return PHDE_OK;
}

int phde_net_adam(phde_mpnn *net, phde_ctx *c, float lr) {
  if (!net || !c || !(lr > 0)) return PHDE_ERR_ARG;
  const float b1 = 0.9f, b2 = 0.999f;
  net->step++;
  const float c1 = 1.0f - std::pow(b1, (float)net->step), c2 = 1.0f - std::pow(b2, (float)net->step);
  net_adam_kernel<<<(kNetParams + 255) / 256, 256, 0, c->stream>>>(net->p, net->m, net->v, net->g, lr, b1,
                                                                  b2, 1e-8f, c1, c2, net->wb);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int phde_net_train_step(phde_mpnn *net, phde_ctx *c, float lr, double *loss) {
  if (!net || !c) return PHDE_ERR_ARG;
  PHDE_TRY_RC(net_backward(net, c));
  PHDE_TRY_RC(phde_net_adam(net, c, lr));
  if (loss) {
    double l = 0;
    PHDE_HIP(hipMemcpyAsync(&l, net->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
    *loss = l / (double)std::max<int64_t>(1, c->g.n);
  }
  return PHDE_OK;
}

int phde_select_pivots(phde_ctx *c, const float *score, int64_t first, int k, int radius, int64_t *pivots) {
  if (!c || !score || !pivots || k < 2 || k > 64 || radius < 1) return PHDE_ERR_ARG;
  if (k > c->g.n) {
    set_error("more pivots than vertices");
    return PHDE_ERR_ARG;
  }
  int64_t p2 = 0;
  int e1 = 0, e2 = 0;
  PHDE_TRY_RC(pivot_features(c, first, &p2, &e1, &e2));
  if (renumbered(c)) {
    DBuf<float> si;
    PHDE_TRY_RC(si.alloc(c->g.n));
    PHDE_HIP(hipMemcpyAsync(si.p, score, sizeof(float) * c->g.n, hipMemcpyHostToDevice, c->stream));
    PHDE_TRY_RC(rows_to_internal(c, si.p, c->lp_score, 1, sizeof(float)));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  } else {
    PHDE_HIP(hipMemcpyAsync(c->lp_score, score, sizeof(float) * c->g.n, hipMemcpyHostToDevice, c->stream));
  }
  pivots[0] = first;
  pivots[1] = p2;
  const int rc = greedy_select(c, c->lp_score, 2, k, radius, pivots);
  return rc == kSelectUnavailable ? PHDE_ERR_UNSUPPORTED : rc;
}

}  // extern "C"
