// orth.hip — degree-weighted Gram-Schmidt (D-orthogonalisation) of pivot
// distance columns, as tall-skinny n x s reductions on gfx950.
//
// Reference behaviour followed:
//  - ParHDE CGS (ParHDE.cpp:190-285): every candidate is projected against
//    the ones vector and all accepted columns using coefficients computed
//    from the original candidate, a -= sum_k (num_k / den_k) S_k with
//    num_k = sum_i a_i S_ik deg_i and den_k = sum_i S_ik^2 deg_i; it is
//    discarded when ||a||_2 < 0.001, else S_j = a / ||a||_2.
//  - WParHDE MGS (WParHDE.cpp:167-214): the same projection done one column
//    at a time on the updated candidate.
//
// Layout: S is column-major (column k at S + k*n) in FP64 or FP32 storage;
// every reduction accumulates in FP64 and is combined in a fixed order
// (per-block partials, then one block), so results are deterministic.
// The candidate a is an FP64 vector of n entries.
#include <algorithm>
#include <vector>

#include "phde_common.h"
#include "orth.h"

namespace phde {

constexpr int kOrthBlock = 256;
constexpr int kColChunk = 16;

__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x / kWave); i++) t += sh[i];
  return t;  // valid in thread 0
}

__global__ void fill_i32_kernel(int32_t *p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u32_kernel(unsigned int *p, int64_t n, unsigned int v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_f64_kernel(double *p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// a_i = dist_i * scale (the column normalisation of ParHDE.cpp:176-185 /
// WParHDE.cpp:154-162 applied to an exact integer distance column).
__global__ void cand_from_dist_kernel(const int32_t *dist, double scale, double *a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = (double)dist[i] * scale;
}

// Partial sums over a row range of num_k = sum_i (a_i deg_i) S_ik for the
// columns of one chunk (blockIdx.y), plus the ones-vector term (u0 constant)
// for chunk 0.  part layout: [col][rowblock], col ncols = ones vector.
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_dots_kernel(const double *a, const double *deg,
                                                             const T *S, int64_t n, int ncols,
                                                             double u0, double *part) {
  __shared__ double sh[kOrthBlock / kWave];
  const int k0 = blockIdx.y * kColChunk;
  const int kc = min(kColChunk, ncols - k0);
  double acc[kColChunk];
#pragma unroll
  for (int c = 0; c < kColChunk; c++) acc[c] = 0;
  double acc0 = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ad = a[i] * deg[i];
    acc0 = fma(ad, u0, acc0);
#pragma unroll
    for (int c = 0; c < kColChunk; c++)
      if (c < kc) acc[c] = fma(ad, (double)S[(int64_t)(k0 + c) * n + i], acc[c]);
  }
  for (int c = 0; c < kColChunk; c++) {
    if (c >= kc) break;
    const double t = block_sum(acc[c], sh);
    if (threadIdx.x == 0) part[(int64_t)(k0 + c) * gridDim.x + blockIdx.x] = t;
  }
  if (blockIdx.y == 0) {
    const double t = block_sum(acc0, sh);
    if (threadIdx.x == 0) part[(int64_t)ncols * gridDim.x + blockIdx.x] = t;
  }
}

// coef_k = num_k * (1 / den_k) for k < ncols, and the ones vector in slot
// kOnesSlot (ParHDE.cpp:239-244).  One block per column; the partials are
// summed in a fixed order (strided per lane, then a wave tree).
__global__ void cgs_coef_kernel(const double *part, int nblk, int ncols, const double *den,
                                double *coef) {
  __shared__ double sh[kOrthBlock / kWave];
  const int k = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)k * nblk + b];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) {
    const int slot = k == ncols ? kOnesSlot : k;
    coef[slot] = t * (1 / den[slot]);
  }
}

// As cgs_coef_kernel, for partials laid out by slot (ones at kOnesSlot).
__global__ void cgs_coef_kernel_slots(const double *part, int nblk, int ncols, const double *den, double *coef) {
  __shared__ double sh[kOrthBlock / kWave];
  const int slot = blockIdx.x == ncols ? kOnesSlot : blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)slot * nblk + b];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) coef[slot] = t * (1 / den[slot]);
}

// a_i -= sum_k coef_k S_ik + coef_ones * u0, and per-block sums of a_i^2
// (ParHDE.cpp:245-265).
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_update_kernel(double *a, const T *S, int64_t n,
                                                               int ncols, double u0,
                                                               const double *coef,
                                                               double *part) {
  __shared__ double sc[kOnesSlot + 1];
  __shared__ double sh[kOrthBlock / kWave];
  for (int k = threadIdx.x; k <= kOnesSlot; k += blockDim.x) sc[k] = coef[k];
  __syncthreads();
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc = 0;
    for (int k = 0; k < ncols; k++) acc = fma(sc[k], (double)S[(int64_t)k * n + i], acc);
    acc = fma(sc[kOnesSlot], u0, acc);
    const double v = a[i] - acc;
    a[i] = v;
    ss = fma(v, v, ss);
  }
  const double t = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out[0] = sqrt(sum of partials) in block order; out[1] = the sum itself.
__global__ void norm_finalize_kernel(const double *part, int nblk, double *out) {
  __shared__ double sh[kOrthBlock / kWave];
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[b];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) {
    out[0] = sqrt(t);
    out[1] = t;
  }
}

// S_j = a * (1 / ||a||) (ParHDE.cpp:271-274) and partial sums of
// S_j^2 deg for the cached denominator of column j.
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_accept_kernel(const double *a,
                                                               const double *norm, const double *deg,
                                                               T *Sj, int64_t n, double *part) {
  __shared__ double sh[kOrthBlock / kWave];
  const double inv = 1 / norm[0];
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const T v = (T)(a[i] * inv);
    Sj[i] = v;
    const double dv = (double)v;
    ss += dv * dv * deg[i];
  }
  const double t = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void den_finalize_kernel(const double *part, int nblk, double *den_slot) {
  __shared__ double sh[kOrthBlock / kWave];
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[b];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) *den_slot = t;
}

// Weighted dot sum_i x_i y_i deg_i (x, y may be FP64 or T storage); used by
// MGS for num and den of one column.
template <typename TX, typename TY>
__global__ __launch_bounds__(kOrthBlock) void wdot_kernel(const TX *x, const TY *y,
                                                         const double *deg, int64_t n,
                                                         double *part) {
  __shared__ double sh[kOrthBlock / kWave];
  double s = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double yv = (double)y[i];
    s += (double)x[i] * (yv * deg[i]);
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// MGS coefficient: out = num * (1 / den) with num from partials and den
// the cached denominator of the column (WParHDE.cpp:175-183).
__global__ void mgs_coef_kernel(const double *part_num, int nblk, const double *den_slot,
                                double *out) {
  __shared__ double sh[kOrthBlock / kWave];
  double sn = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) sn += part_num[b];
  const double tn = block_sum(sn, sh);
  if (threadIdx.x == 0) out[0] = tn * (1 / den_slot[0]);
}

// Per-block sums of a_i^2.
__global__ __launch_bounds__(kOrthBlock) void sumsq_kernel(const double *a, int64_t n,
                                                          double *part) {
  __shared__ double sh[kOrthBlock / kWave];
  double s = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s += a[i] * a[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// a_i -= c * y_i (WParHDE.cpp:184-188: (num * U_k) * (1/den)).
template <typename TY>
__global__ void mgs_axpy_kernel(double *a, const TY *y, const double *c, int64_t n) {
  const double cc = c[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] -= cc * (double)y[i];
}


// ---------------------------------------------------------------- blocked CGS
// A block of up to kCgsBlock candidates shares the passes over the columns
// accepted before the block: one pass for their dot products, one for the
// columns' part of their updates.  Every per-element sum runs in the same
// order, over the same rows per thread and through the same reduction tree
// as in cgs_dots_kernel / cgs_update_kernel, so each candidate ends exactly
// as if it had been taken alone (the remaining work per candidate covers
// the columns accepted inside the block).

// Dot products of the candidates A_b (b < nb) with columns [0, m) and the
// ones vector.  part layout: [b][slot][rowblock] with slot = column or
// kOnesSlot.  Columns are taken Cols per blockIdx.y; every (candidate,
// column) sum runs over the same rows in the same order whatever Cols is,
// so the choice changes time, not bits.
template <typename T, int Cols>
__global__ __launch_bounds__(kOrthBlock) void cgs_dots_multi_kernel(const double *A, const double *deg,
                                                                   const T *S, int64_t n, int m, int nb,
                                                                   double u0, double *part) {
  __shared__ double sh[kCgsBlock * (Cols + 1)][kOrthBlock / kWave];
  const int k0 = blockIdx.y * Cols;
  const int kc = max(0, min(Cols, m - k0));
  const bool ones = blockIdx.y == 0;
  double acc[kCgsBlock][Cols + 1];
#pragma unroll
  for (int b = 0; b < kCgsBlock; b++)
#pragma unroll
    for (int c = 0; c <= Cols; c++) acc[b][c] = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double dg = deg[i];
    double sv[Cols];
#pragma unroll
    for (int c = 0; c < Cols; c++) sv[c] = c < kc ? (double)S[(int64_t)(k0 + c) * n + i] : 0.0;
#pragma unroll
    for (int b = 0; b < kCgsBlock; b++) {
      if (b < nb) {
        const double ad = A[(int64_t)b * n + i] * dg;
        acc[b][Cols] = fma(ad, u0, acc[b][Cols]);
#pragma unroll
        for (int c = 0; c < Cols; c++) acc[b][c] = fma(ad, sv[c], acc[b][c]);
      }
    }
  }
  // every value's wave sum as in block_sum, then the waves' sums in order
  const int w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
#pragma unroll
  for (int b = 0; b < kCgsBlock; b++)
#pragma unroll
    for (int c = 0; c <= Cols; c++) {
      double v = acc[b][c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0) sh[b * (Cols + 1) + c][w] = v;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < kCgsBlock * (Cols + 1); t += blockDim.x) {
    const int b = t / (Cols + 1), c = t % (Cols + 1);
    if (b >= nb) continue;
    if (c < Cols && c >= kc) continue;
    if (c == Cols && !ones) continue;
    double tot = 0;
    for (int q = 0; q < (int)(blockDim.x / kWave); q++) tot += sh[t][q];
    const int slot = c == Cols ? kOnesSlot : k0 + c;
    part[((int64_t)b * (kOnesSlot + 1) + slot) * gridDim.x + blockIdx.x] = tot;
  }
}

// Dot products of one candidate a with columns [k_lo, k_hi) (the block's
// own accepted columns); part: [slot][rowblock].
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_dots_range_kernel(const double *a, const double *deg,
                                                                   const T *S, int64_t n, int k_lo, int k_hi,
                                                                   double *part) {
  __shared__ double sh[kOrthBlock / kWave];
  double acc[kCgsBlock];
#pragma unroll
  for (int c = 0; c < kCgsBlock; c++) acc[c] = 0;
  const int kc = k_hi - k_lo;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ad = a[i] * deg[i];
#pragma unroll
    for (int c = 0; c < kCgsBlock; c++)
      if (c < kc) acc[c] = fma(ad, (double)S[(int64_t)(k_lo + c) * n + i], acc[c]);
  }
  for (int c = 0; c < kc; c++) {
    const double t = block_sum(acc[c], sh);
    if (threadIdx.x == 0) part[(int64_t)(k_lo + c) * gridDim.x + blockIdx.x] = t;
  }
}

// coef[b][k] = num / den for k < m and the ones slot, every candidate.
__global__ void cgs_coef_multi_kernel(const double *part, int nblk, int m, const double *den, double *coef) {
  __shared__ double sh[kOrthBlock / kWave];
  const int b = blockIdx.y;
  const int k = blockIdx.x == m ? kOnesSlot : blockIdx.x;
  const double *p = part + ((int64_t)b * (kOnesSlot + 1) + k) * nblk;
  double s = 0;
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) s += p[q];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) coef[(int64_t)b * (kOnesSlot + 1) + k] = t * (1 / den[k]);
}

// acc[b][i] = sum_{k < m} coef[b][k] S_ik, in column order (the first part
// of cgs_update_kernel's sum).
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_update_old_kernel(const T *S, int64_t n, int m, int nb,
                                                                   const double *coef, double *acc_out) {
  __shared__ double sc[kCgsBlock][kOnesSlot + 1];
  for (int t = threadIdx.x; t < kCgsBlock * (kOnesSlot + 1); t += blockDim.x)
    sc[t / (kOnesSlot + 1)][t % (kOnesSlot + 1)] = t / (kOnesSlot + 1) < nb ? coef[t] : 0.0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc[kCgsBlock];
#pragma unroll
    for (int b = 0; b < kCgsBlock; b++) acc[b] = 0;
    for (int k = 0; k < m; k++) {
      const double sv = (double)S[(int64_t)k * n + i];
#pragma unroll
      for (int b = 0; b < kCgsBlock; b++) acc[b] = fma(sc[b][k], sv, acc[b]);
    }
#pragma unroll
    for (int b = 0; b < kCgsBlock; b++)
      if (b < nb) acc_out[(int64_t)b * n + i] = acc[b];
  }
}

// The rest of one candidate's update: acc continues over columns [m, ncols)
// and the ones term, a = A_b - acc, per-block sums of a^2.
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_update_tail_kernel(double *a, const double *Ab, const double *acc_old,
                                                                    const T *S, int64_t n, int m, int ncols,
                                                                    double u0, const double *coef, double *part) {
  __shared__ double sc[kOnesSlot + 1];
  __shared__ double sh[kOrthBlock / kWave];
  for (int k = threadIdx.x; k <= kOnesSlot; k += blockDim.x) sc[k] = coef[k];
  __syncthreads();
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc = acc_old[i];
    for (int k = m; k < ncols; k++) acc = fma(sc[k], (double)S[(int64_t)k * n + i], acc);
    acc = fma(sc[kOnesSlot], u0, acc);
    const double v = Ab[i] - acc;
    a[i] = v;
    ss = fma(v, v, ss);
  }
  const double t = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---------------------------------------------------------------- blocked CGS2
// FP32 storage path.  Each candidate is projected twice (CGS2) against the
// ones vector and the accepted columns; for a block of candidates the two
// passes over the m columns accepted before the block (and the ones
// vector) are shared: A_b -= sum_k coef_bk S_k + coef_b,ones u0, twice,
// with coefficients from the block's dot products of that pass.  Each
// candidate is then projected twice against the columns accepted inside
// the block before it, then normed and judged.  In exact arithmetic this is
// the same projection as candidate-by-candidate CGS2 (the accepted columns
// and the ones vector are D-orthogonal); the order of the FP64 sums differs.

// A_b -= sum_{k < m} coef[b][k] S_k + coef[b][ones] u0, every candidate b.
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_block_sub_kernel(double *A, const T *S, int64_t n, int m, int nb,
                                                                  double u0, const double *coef) {
  __shared__ double sc[kCgsBlock][kOnesSlot + 1];
  for (int t = threadIdx.x; t < kCgsBlock * (kOnesSlot + 1); t += blockDim.x)
    sc[t / (kOnesSlot + 1)][t % (kOnesSlot + 1)] = t / (kOnesSlot + 1) < nb ? coef[t] : 0.0;
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc[kCgsBlock];
#pragma unroll
    for (int b = 0; b < kCgsBlock; b++) acc[b] = sc[b][kOnesSlot] * u0;
    for (int k = 0; k < m; k++) {
      const double sv = (double)S[(int64_t)k * n + i];
#pragma unroll
      for (int b = 0; b < kCgsBlock; b++) acc[b] = fma(sc[b][k], sv, acc[b]);
    }
#pragma unroll
    for (int b = 0; b < kCgsBlock; b++)
      if (b < nb) A[(int64_t)b * n + i] -= acc[b];
  }
}

// coef[k] = num_k / den_k for the columns k in [k_lo, k_hi) (partials laid
// out by column as cgs_dots_range_kernel writes them).
__global__ void cgs_coef_range_kernel(const double *part, int nblk, int k_lo, const double *den, double *coef) {
  __shared__ double sh[kOrthBlock / kWave];
  const int k = k_lo + blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(int64_t)k * nblk + b];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) coef[k] = t * (1 / den[k]);
}

// a -= sum_{k in [k_lo, k_hi)} coef_k S_k, per-block sums of a^2.
template <typename T>
__global__ __launch_bounds__(kOrthBlock) void cgs_sub_range_kernel(double *a, const T *S, int64_t n, int k_lo,
                                                                  int k_hi, const double *coef, double *part) {
  __shared__ double sc[kCgsBlock];
  __shared__ double sh[kOrthBlock / kWave];
  const int kc = k_hi - k_lo;
  for (int k = threadIdx.x; k < kc; k += blockDim.x) sc[k] = coef[k_lo + k];
  __syncthreads();
  double ss = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double acc = 0;
    for (int k = 0; k < kc; k++) acc = fma(sc[k], (double)S[(int64_t)(k_lo + k) * n + i], acc);
    const double v = a[i] - acc;
    a[i] = v;
    ss = fma(v, v, ss);
  }
  const double t = block_sum(ss, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---------------------------------------------------------------- host side

static int grid_for(int64_t n, int cap) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + kOrthBlock - 1) / kOrthBlock, cap));
}

int fill_i32(phde_ctx *c, int32_t *p, int64_t n, int32_t v) {
  fill_i32_kernel<<<grid_for(n, 4096), kOrthBlock, 0, c->stream>>>(p, n, v);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}
int fill_u32(phde_ctx *c, unsigned int *p, int64_t n, unsigned int v) {
  fill_u32_kernel<<<grid_for(n, 4096), kOrthBlock, 0, c->stream>>>(p, n, v);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}
int fill_f64(phde_ctx *c, double *p, int64_t n, double v) {
  fill_f64_kernel<<<grid_for(n, 4096), kOrthBlock, 0, c->stream>>>(p, n, v);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Orth::setup(phde_ctx *c, int64_t n_, int s_, bool fp32_) {
  ctx = c;
  n = n_;
  s = s_;
  fp32 = fp32_;
  nblk = grid_for(n, 2 * std::max(c->num_cus, 1) * 4);
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  PHDE_HIP(hipMalloc(&S, esz * (size_t)n * s));
  PHDE_HIP(hipMalloc(&a, sizeof(double) * n));
  PHDE_HIP(hipMalloc(&part, sizeof(double) * (size_t)nblk * (kOnesSlot + 2) * 2));
  PHDE_HIP(hipMalloc(&den, sizeof(double) * (kOnesSlot + 1)));
  PHDE_HIP(hipMalloc(&coef, sizeof(double) * (kOnesSlot + 1)));
  PHDE_HIP(hipMalloc(&scal, sizeof(double) * 8));
  PHDE_HIP(hipHostMalloc(&h_scal, sizeof(double) * 8, hipHostMallocDefault));
  return PHDE_OK;
}

void Orth::release() {
  if (S) (void)hipFree(S);
  if (a) (void)hipFree(a);
  if (part) (void)hipFree(part);
  if (den) (void)hipFree(den);
  if (coef) (void)hipFree(coef);
  if (scal) (void)hipFree(scal);
  if (ones) (void)hipFree(ones);
  if (h_scal) (void)hipHostFree(h_scal);
  if (blkA) (void)hipFree(blkA);
  if (blkAcc) (void)hipFree(blkAcc);
  if (blkPart) (void)hipFree(blkPart);
  if (blkCoef) (void)hipFree(blkCoef);
  S = nullptr;
  a = part = den = coef = scal = h_scal = ones = nullptr;
  blkA = blkAcc = blkPart = blkCoef = nullptr;
}

// den of the ones vector: sum_i u0*u0*deg_i (ParHDE.cpp:226).
int Orth::init_ones(const double *deg, double u0_) {
  deg_ = deg;
  u0 = u0_;
  // a = u0 everywhere, then reuse the weighted dot.
  int rc = fill_f64(ctx, a, n, u0);
  if (rc) return rc;
  wdot_kernel<double, double><<<nblk, kOrthBlock, 0, ctx->stream>>>(a, a, deg, n, part);
  den_finalize_kernel<<<1, kOrthBlock, 0, ctx->stream>>>(part, nblk, den + kOnesSlot);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Orth::load_candidate(const int32_t *dist, double scale) {
  cand_from_dist_kernel<<<grid_for(n, 4096), kOrthBlock, 0, ctx->stream>>>(dist, scale, a, n);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

template <typename T>
static int cgs_pass_t(Orth &o, int ncols) {
  hipStream_t st = o.ctx->stream;
  const T *S = (const T *)o.S;
  const int nchunk = std::max(1, (ncols + kColChunk - 1) / kColChunk);
  cgs_dots_kernel<T><<<dim3(o.nblk, nchunk), kOrthBlock, 0, st>>>(o.a, o.deg_, S, o.n, ncols,
                                                                  o.u0, o.part);
  cgs_coef_kernel<<<ncols + 1, kOrthBlock, 0, st>>>(o.part, o.nblk, ncols, o.den, o.coef);
  cgs_update_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(o.a, S, o.n, ncols, o.u0, o.coef,
                                                      o.part);
  norm_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.scal);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Orth::cgs_pass(int ncols) { return fp32 ? cgs_pass_t<float>(*this, ncols) : cgs_pass_t<double>(*this, ncols); }

template <typename T>
static int mgs_pass_t(Orth &o, int ncols) {
  hipStream_t st = o.ctx->stream;
  const T *S = (const T *)o.S;
  // ones vector first (U[:,0] of WParHDE.cpp:170), then accepted columns
  // in order, each against the updated candidate.
  wdot_kernel<double, double><<<o.nblk, kOrthBlock, 0, st>>>(o.a, o.ones, o.deg_, o.n, o.part);
  mgs_coef_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.den + kOnesSlot, o.coef);
  mgs_axpy_kernel<double><<<grid_for(o.n, 4096), kOrthBlock, 0, st>>>(o.a, o.ones, o.coef, o.n);
  for (int k = 0; k < ncols; k++) {
    const T *Sk = S + (int64_t)k * o.n;
    wdot_kernel<double, T><<<o.nblk, kOrthBlock, 0, st>>>(o.a, Sk, o.deg_, o.n, o.part);
    mgs_coef_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.den + k, o.coef);
    mgs_axpy_kernel<T><<<grid_for(o.n, 4096), kOrthBlock, 0, st>>>(o.a, Sk, o.coef, o.n);
  }
  sumsq_kernel<<<o.nblk, kOrthBlock, 0, st>>>(o.a, o.n, o.part);
  norm_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.scal);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Orth::mgs_pass(int ncols) {
  if (!ones) {
    PHDE_HIP(hipMalloc(&ones, sizeof(double) * n));
    int rc = fill_f64(ctx, ones, n, u0);
    if (rc) return rc;
  }
  return fp32 ? mgs_pass_t<float>(*this, ncols) : mgs_pass_t<double>(*this, ncols);
}

int Orth::read_norm(double *norm) {
  PHDE_HIP(hipMemcpyAsync(h_scal, scal, sizeof(double) * 2, hipMemcpyDeviceToHost, ctx->stream));
  PHDE_HIP(hipStreamSynchronize(ctx->stream));
  *norm = h_scal[0];
  return PHDE_OK;
}

template <typename T>
static int accept_t(Orth &o, int j) {
  hipStream_t st = o.ctx->stream;
  T *Sj = (T *)o.S + (int64_t)j * o.n;
  cgs_accept_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(o.a, o.scal, o.deg_, Sj, o.n, o.part);
  den_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.den + j);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Orth::accept(int j) { return fp32 ? accept_t<float>(*this, j) : accept_t<double>(*this, j); }

// Columns per block of the blocked dot products.  Each block re-reads the
// candidates, but wider blocks hold more accumulators and run fewer waves:
// measured on RGG n=4.2M, s=50 (bit-identical results), orth 12.4 / 12.8 /
// 15.0 ms (FP64) and 20.9 / 23.9 / 32.1 ms (FP32 storage, CGS2) at 4 / 8 /
// 16 columns.
constexpr int kCgsCols = 4;

template <typename T>
static void dots_multi(Orth &o, const T *S, int m, int nc) {
  const int nchunk = std::max(1, (m + kCgsCols - 1) / kCgsCols);
  cgs_dots_multi_kernel<T, kCgsCols><<<dim3(o.nblk, nchunk), kOrthBlock, 0, o.ctx->stream>>>(
      o.blkA, o.deg_, S, o.n, m, nc, o.u0, o.blkPart);
}

template <typename T>
static int cgs_block_t(Orth &o, const int32_t *const *cols, const double *scales, int nc, int s_target,
                       int &accepted, std::vector<bool> &taken, std::vector<double> &norms) {
  hipStream_t st = o.ctx->stream;
  const T *S = (const T *)o.S;
  const int64_t n = o.n;
  const int m = accepted;
  for (int b = 0; b < nc; b++)
    cand_from_dist_kernel<<<grid_for(n, 4096), kOrthBlock, 0, st>>>(cols[b], scales[b], o.blkA + (int64_t)b * n, n);
  dots_multi<T>(o, S, m, nc);
  cgs_coef_multi_kernel<<<dim3(m + 1, nc), kOrthBlock, 0, st>>>(o.blkPart, o.nblk, m, o.den, o.blkCoef);
  cgs_update_old_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(S, n, m, nc, o.blkCoef, o.blkAcc);
  PHDE_HIP(hipGetLastError());
  taken.assign(nc, false);
  norms.assign(nc, 0.0);
  for (int b = 0; b < nc && accepted < s_target; b++) {
    const int ncols = accepted;
    double *pb = o.blkPart + (int64_t)b * (kOnesSlot + 1) * o.nblk;
    const double *Ab = o.blkA + (int64_t)b * n;
    if (ncols > m)
      cgs_dots_range_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(Ab, o.deg_, S, n, m, ncols, pb);
    // coefficients of every column so far (the ones in slot kOnesSlot)
    cgs_coef_kernel_slots<<<ncols + 1, kOrthBlock, 0, st>>>(pb, o.nblk, ncols, o.den, o.coef);
    cgs_update_tail_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(o.a, Ab, o.blkAcc + (int64_t)b * n, S, n, m, ncols,
                                                             o.u0, o.coef, o.part);
    norm_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.scal);
    PHDE_HIP(hipGetLastError());
    double nd = 0;
    PHDE_TRY_RC(o.read_norm(&nd));
    norms[b] = nd;
    if (nd >= 0.001) {
      PHDE_TRY_RC(o.accept(accepted));
      accepted++;
      taken[b] = true;
    }
  }
  return PHDE_OK;
}

template <typename T>
static int cgs2_block_t(Orth &o, const int32_t *const *cols, const double *scales, int nc, int s_target,
                        int &accepted, std::vector<bool> &taken, std::vector<double> &norms) {
  hipStream_t st = o.ctx->stream;
  T *S = (T *)o.S;
  const int64_t n = o.n;
  const int m = accepted;
  for (int b = 0; b < nc; b++)
    cand_from_dist_kernel<<<grid_for(n, 4096), kOrthBlock, 0, st>>>(cols[b], scales[b], o.blkA + (int64_t)b * n, n);
  for (int pass = 0; pass < 2; pass++) {
    dots_multi<T>(o, S, m, nc);
    cgs_coef_multi_kernel<<<dim3(m + 1, nc), kOrthBlock, 0, st>>>(o.blkPart, o.nblk, m, o.den, o.blkCoef);
    cgs_block_sub_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(o.blkA, S, n, m, nc, o.u0, o.blkCoef);
  }
  PHDE_HIP(hipGetLastError());
  taken.assign(nc, false);
  norms.assign(nc, 0.0);
  for (int b = 0; b < nc && accepted < s_target; b++) {
    double *ab = o.blkA + (int64_t)b * n;
    if (accepted > m) {
      for (int pass = 0; pass < 2; pass++) {
        cgs_dots_range_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(ab, o.deg_, S, n, m, accepted, o.part);
        cgs_coef_range_kernel<<<accepted - m, kOrthBlock, 0, st>>>(o.part, o.nblk, m, o.den, o.coef);
        cgs_sub_range_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(ab, S, n, m, accepted, o.coef, o.part);
      }
    } else {
      sumsq_kernel<<<o.nblk, kOrthBlock, 0, st>>>(ab, n, o.part);
    }
    norm_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.scal);
    PHDE_HIP(hipGetLastError());
    double nd = 0;
    PHDE_TRY_RC(o.read_norm(&nd));
    norms[b] = nd;
    if (nd >= 0.001) {
      T *Sj = S + (int64_t)accepted * n;
      cgs_accept_kernel<T><<<o.nblk, kOrthBlock, 0, st>>>(ab, o.scal, o.deg_, Sj, n, o.part);
      den_finalize_kernel<<<1, kOrthBlock, 0, st>>>(o.part, o.nblk, o.den + accepted);
      PHDE_HIP(hipGetLastError());
      accepted++;
      taken[b] = true;
    }
  }
  return PHDE_OK;
}

int Orth::cgs_block(const int32_t *const *cols, const double *scales, int nc, int s_target, int &accepted,
                    std::vector<bool> &taken, std::vector<double> &norms) {
  if (nc < 1 || nc > kCgsBlock) {
    set_error("internal: CGS block size");
    return PHDE_ERR_ARG;
  }
  if (!blkA) {
    PHDE_HIP(hipMalloc(&blkA, sizeof(double) * (size_t)n * kCgsBlock));
    PHDE_HIP(hipMalloc(&blkAcc, sizeof(double) * (size_t)n * kCgsBlock));
    PHDE_HIP(hipMalloc(&blkPart, sizeof(double) * (size_t)nblk * (kOnesSlot + 1) * kCgsBlock));
    PHDE_HIP(hipMalloc(&blkCoef, sizeof(double) * (size_t)(kOnesSlot + 1) * kCgsBlock));
  }
  // FP32 storage: CGS2 (two projections per candidate); FP64: one, bit for
  // bit as the candidate-by-candidate kernels
  return fp32 ? cgs2_block_t<float>(*this, cols, scales, nc, s_target, accepted, taken, norms)
              : cgs_block_t<double>(*this, cols, scales, nc, s_target, accepted, taken, norms);
}

template <typename T>
static int ones_dots_t(Orth &o, double *out) {
  hipStream_t st = o.ctx->stream;
  PHDE_TRY_RC(fill_f64(o.ctx, o.a, o.n, o.u0));
  const int nchunk = std::max(1, (o.s + kColChunk - 1) / kColChunk);
  cgs_dots_kernel<T><<<dim3(o.nblk, nchunk), kOrthBlock, 0, st>>>(o.a, o.deg_, (const T *)o.S, o.n, o.s, o.u0,
                                                                  o.part);
  PHDE_HIP(hipGetLastError());
  std::vector<double> h((size_t)o.nblk * (o.s + 1));
  PHDE_HIP(hipMemcpyAsync(h.data(), o.part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
  PHDE_HIP(hipStreamSynchronize(st));
  for (int k = 0; k <= o.s; k++) {
    double v = 0;
    for (int b = 0; b < o.nblk; b++) v += h[(size_t)k * o.nblk + b];
    out[k] = v;
  }
  return PHDE_OK;
}

int Orth::ones_dots(double *out) { return fp32 ? ones_dots_t<float>(*this, out) : ones_dots_t<double>(*this, out); }


}  // namespace phde
