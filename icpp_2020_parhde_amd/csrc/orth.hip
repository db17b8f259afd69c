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
#include "spmm.h"

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

// ---------------------------------------------------------------- Gram CGS
// FP64 storage, unweighted graphs.  The candidates are distance columns, so
// C = [1, d_1 .. d_N] holds small non-negative integers and the Gram
// matrices G_I = C^T C and G_D = C^T D C (D: the integer vertex degrees)
// are sums of integer products.  While max(sum deg, n) * max(d)^2 < 2^53
// every partial sum is an integer below 2^53, which FP64 holds exactly:
// one pass over C gives both matrices exactly, whatever the order of the
// sums.  The reference's classical GS (ParHDE.cpp:190-285) then runs on
// coefficient vectors in one wave, lane e holding coefficient e: candidate
// j is a = C (scale_j e_j); the ones vector is C (u0 e_0); an accepted
// column is S_k = C m_k, with P_k = G_D m_k and den_k = m_k^T P_k kept; the
// projection coefficient of S_k is num_k / den_k with num_k = scale_j
// P_k[j]; the residual r has ||C r||_2^2 = r^T G_I r.  The products G r
// and G_D m sum large terms of both signs down to small results (the
// residual of a nearly dependent candidate), so they run in double-double
// arithmetic (error-free products and sums, one rounding at the end); the
// outer dot products then lose no more than the vector CGS does.  S = C M
// is formed in one pass at the end.

constexpr int kGcThreads = 256;  // four waves per block of the Gram pass
constexpr int kGcTiles = 10;     // 16 x 16 tiles of the upper triangle of a 64 x 64 matrix
constexpr int kGcPart = 2 * kGcTiles * 256;  // partial sums per wave: two matrices' tiles, 4 per lane


// Upper-triangle tile number t < 10 -> (p, q), p <= q < 4.
__device__ __forceinline__ void gram_tile(int t, int &p, int &q) {
  p = 0;
  while (t >= 4 - p) {
    t -= 4 - p;
    p++;
  }
  q = p + t;
}

// Partial sums of G_I and G_D on the FP64 matrix cores.  C (n x 64, column
// 0 the ones, columns 1..N the candidates, zero past N) is read straight
// from the distance columns: a wave takes 4 rows per step, lane l supplying
// C[row + l/16][16 p + l%16] for p < 4, and issues one
// v_mfma_f64_16x16x4f64 per tile (p <= q) of each matrix (G_D's left
// operand scaled by the row's degree).  Operands and products are integers
// and so is every partial sum below 2^53, which FP64 holds exactly; the
// result is the same in any order.  Each wave walks a contiguous range of
// steps (its column segments stay in cache across steps), kGcSteps at a
// time, the next group's loads issued before the current group's products;
// it writes its sums to part[wave][tile][lane][4].  part_max[block] = the
// largest distance the block read.  With pack: the row's 2-bit residues
// d mod 4 of the 64 columns, word p holding columns 16 p .. 16 p + 15 (two
// bits each, gathered across the 16 lanes of the row).
constexpr int kGcSteps = 4;
typedef double gc_v4d __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kGcThreads) void gram_tiles_kernel(GramCols cols, int N, const double *deg, int64_t n,
                                                               double *part, int *part_max, uint4 *pack) {
  __shared__ int smax[kGcThreads / kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int kr = lane / 16, cl = lane % 16;
  const int32_t *cp[4];
  int keep[4];  // lane's column p: 0 zero, 1 ones, 2 a candidate
#pragma unroll
  for (int p = 0; p < 4; p++) {
    const int c = 16 * p + cl;
    keep[p] = c == 0 ? 1 : c <= N ? 2 : 0;
    cp[p] = cols.col[min(max(c - 1, 0), N - 1)];
  }
  gc_v4d acc[2 * kGcTiles];
#pragma unroll
  for (int t = 0; t < 2 * kGcTiles; t++) acc[t] = gc_v4d{0.0, 0.0, 0.0, 0.0};
  int dmax = 0;
  const int64_t nsteps = (n + 3) / 4;
  const int64_t gw = (int64_t)blockIdx.x * (kGcThreads / kWave) + wave;
  const int64_t nw = (int64_t)gridDim.x * (kGcThreads / kWave);
  const int64_t s0 = nsteps * gw / nw, s1 = nsteps * (gw + 1) / nw;
  int vn[kGcSteps][4];
  double dn[kGcSteps];
  // unconditional loads (rows clamped in range), masked when used
  auto load = [&](int64_t g0) {
#pragma unroll
    for (int u = 0; u < kGcSteps; u++) {
      const int64_t rr = min((g0 + u) * 4 + kr, n - 1);
      dn[u] = deg[rr];
#pragma unroll
      for (int p = 0; p < 4; p++) vn[u][p] = cp[p][rr];
    }
  };
  if (s0 < s1) load(s0);
  for (int64_t g0 = s0; g0 < s1; g0 += kGcSteps) {
    int vc[kGcSteps][4];
    double dc[kGcSteps];
#pragma unroll
    for (int u = 0; u < kGcSteps; u++) {
      dc[u] = dn[u];
#pragma unroll
      for (int p = 0; p < 4; p++) vc[u][p] = vn[u][p];
    }
    if (g0 + kGcSteps < s1) load(g0 + kGcSteps);
#pragma unroll
    for (int u = 0; u < kGcSteps; u++) {
      const bool rok = g0 + u < s1 && (g0 + u) * 4 + kr < n;
      double a[4], ad[4];
      unsigned res[4];
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const int v = !rok ? 0 : keep[p] == 1 ? 1 : keep[p] == 2 ? vc[u][p] : 0;
        dmax = max(dmax, v);
        a[p] = (double)v;
        ad[p] = a[p] * dc[u];  // integer product below 2^53: exact
        res[p] = (unsigned)(v & 3) << (2 * cl);
      }
      if (pack) {
#pragma unroll
        for (int off = 1; off < 16; off <<= 1)
#pragma unroll
          for (int p = 0; p < 4; p++) res[p] |= __shfl_xor(res[p], off);
        if (cl == 0 && rok) pack[(g0 + u) * 4 + kr] = make_uint4(res[0], res[1], res[2], res[3]);
      }
#pragma unroll
      for (int t = 0; t < kGcTiles; t++) {
        const int p = t < 4 ? 0 : t < 7 ? 1 : t < 9 ? 2 : 3;
        const int q = t < 4 ? t : t < 7 ? t - 3 : t < 9 ? t - 5 : 3;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[p], a[q], acc[t], 0, 0, 0);
        acc[kGcTiles + t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad[p], a[q], acc[kGcTiles + t], 0, 0, 0);
      }
    }
  }
  double *o = part + gw * kGcPart;
#pragma unroll
  for (int t = 0; t < 2 * kGcTiles; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) o[(t * kWave + lane) * 4 + v] = acc[t][v];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dmax = max(dmax, __shfl_xor(dmax, off));
  if (lane == 0) smax[wave] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int w = 0; w < kGcThreads / kWave; w++) m = max(m, smax[w]);
    part_max[blockIdx.x] = m;
  }
}

// G[mat][i][j] (W x W, both triangles; mat 0 = G_I, 1 = G_D) = the sum of
// the waves' partial tiles, one wave per entry; gmax = the largest
// distance.  The partials are integers and so is every partial sum: any
// order gives the same result.  Lane l's output v of a 16 x 16 tile is its
// row l/16 + 4 v, column l%16 (the lane layout of v_mfma_f64_16x16x4f64).
__global__ void gram_sum_kernel(const double *part, const int *part_max, int nblk, int W, double *G, int *gmax) {
  const int nparts = nblk * (kGcThreads / kWave);
  const int lane = threadIdx.x % kWave;
  const int e = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (e == 0) {
    int m = 0;
    for (int b = lane; b < nblk; b += kWave) m = max(m, part_max[b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if (lane == 0) *gmax = m;
  }
  if (e >= kGcPart) return;
  const int t = e / 256, l = (e / 4) % kWave, v = e % 4;
  const int mat = t / kGcTiles;
  int p, q;
  gram_tile(t % kGcTiles, p, q);
  const int i = 16 * p + l / 16 + 4 * v, j = 16 * q + l % 16;
  if (i >= W || j >= W) return;
  double s = 0;
  for (int b = lane; b < nparts; b += kWave) s += part[(int64_t)b * kGcPart + e];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane != 0) return;
  G[(int64_t)mat * W * W + i * W + j] = s;
  G[(int64_t)mat * W * W + j * W + i] = s;
}

// hi + lo += a * b without rounding error in the product or the leading sum
// (Dekker / Knuth two-product and two-sum on FMA hardware; no contraction,
// which would fuse a * b into the sums and break the error-free steps).
__device__ __forceinline__ void dd_fma(double &hi, double &lo, double a, double b) {
#pragma clang fp contract(off)
  const double p = a * b;
  const double pe = fma(a, b, -p);
  const double s = hi + p;
  const double bb = s - hi;
  const double se = (hi - (s - bb)) + (p - bb);
  hi = s;
  lo += se + pe;
}

// sum_f G[f][lane] x[f] (G rows of stride 64, x in LDS) in double-double:
// four interleaved chains, combined error-free, rounded once.
__device__ __forceinline__ double dd_dot(const double *G, const double *x, int W, int lane) {
#pragma clang fp contract(off)
  double h[4] = {0, 0, 0, 0}, l[4] = {0, 0, 0, 0};
  int f = 0;
  for (; f + 4 <= W; f += 4)
#pragma unroll
    for (int u = 0; u < 4; u++) dd_fma(h[u], l[u], G[(f + u) * 64 + lane], x[f + u]);
  for (; f < W; f++) dd_fma(h[0], l[0], G[f * 64 + lane], x[f]);
  double hi = h[0], lo = l[0];
#pragma unroll
  for (int u = 1; u < 4; u++) {
    const double s = hi + h[u];
    const double bb = s - hi;
    lo += ((hi - (s - bb)) + (h[u] - bb)) + l[u];
    hi = s;
  }
  return hi + lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;  // the same value in every lane (each step adds a commuted pair)
}

// The coefficient-space CGS, one wave.  LDS: GI, GD with row stride 64
// (row f holds G[f][*]), M [k][64], PT [e][64] (P_k[e] at e*64 + k), and a
// broadcast row.  Outputs: M_out [k][64], den[k] (the accepted columns'
// D-norms squared), out[0..N) the candidates' norms, out[N..2N) 1 when
// taken, out[2N] candidates looked at, out[2N+1] columns accepted.
__global__ __launch_bounds__(kWave) void gram_cgs_kernel(const double *G, const int *gmax, int N, int W,
                                                         GramCols cols, double u0, int s_target,
                                                         double *M_out, double *den, double *out) {
  extern __shared__ double csm[];
  double *GI = csm, *GD = GI + W * 64, *Ml = GD + W * 64, *PT = Ml + kGramMax * 64, *row = PT + 64 * 64;
  const int lane = threadIdx.x;
  for (int e = lane; e < W * 64; e += kWave) {
    const int f = e / 64, c = e % 64;
    GI[e] = c < W ? G[f * W + c] : 0.0;
    GD[e] = c < W ? G[W * W + f * W + c] : 0.0;
  }
  for (int e = lane; e < 64 * 64; e += kWave) PT[e] = 0.0;
  __syncthreads();
  // exact sums: every partial sum an integer below 2^53, and the diagonal
  // of G_I equal to the exact sums of squares the traversals counted
  const double dm = (double)*gmax;
  bool exact = max(GD[0], GI[0]) * dm * dm < 9007199254740992.0;
  for (int j = 1; j <= N; j++) exact = exact && GI[j * 64 + j] == cols.sumsq[j - 1];
  if (!exact) {
    if (lane == 0) {
      out[2 * N] = 0;
      out[2 * N + 1] = 0;
    }
    return;
  }
  const double den1 = u0 * u0 * GD[0];  // sum_i u0^2 deg_i
  double my_den = 1.0;                  // lane k: den_k
  int acc = 0, looked = 0;
  for (int j = 1; j <= N && acc < s_target; j++) {
    const double sc = cols.scale[j - 1];
    // projection coefficients: lane k < acc of column k; the ones vector's
    const double ck = lane < acc ? (PT[j * 64 + lane] * sc) * (1 / my_den) : 0.0;
    const double c1 = (GD[j] * u0 * sc) * (1 / den1);
    row[lane] = ck;
    __syncthreads();
    // sum_k ck M_k, four interleaved partial sums
    double t4[4] = {0, 0, 0, 0};
    int k = 0;
    for (; k + 4 <= acc; k += 4)
#pragma unroll
      for (int u = 0; u < 4; u++) t4[u] = fma(row[k + u], Ml[(k + u) * 64 + lane], t4[u]);
    for (; k < acc; k++) t4[0] = fma(row[k], Ml[k * 64 + lane], t4[0]);
    double tmp = (t4[0] + t4[1]) + (t4[2] + t4[3]);
    if (lane == 0) tmp = fma(c1, u0, tmp);
    const double r = (lane == j ? sc : 0.0) - tmp;
    __syncthreads();
    row[lane] = r;
    __syncthreads();
    const double q = wave_sum(r * dd_dot(GI, row, W, lane));
    const double nd = sqrt(max(q, 0.0));
    looked = j;
    if (lane == 0) {
      out[j - 1] = nd;
      out[N + j - 1] = nd >= 0.001 ? 1.0 : 0.0;
    }
    __syncthreads();
    if (nd >= 0.001) {
      const double m = r * (1 / nd);
      Ml[acc * 64 + lane] = m;
      M_out[acc * 64 + lane] = m;
      row[lane] = m;
      __syncthreads();
      const double p = dd_dot(GD, row, W, lane);
      PT[lane * 64 + acc] = p;
      const double d = wave_sum(m * p);
      if (lane == acc) my_den = d;
      if (lane == 0) den[acc] = d;
      acc++;
      __syncthreads();
    }
  }
  if (lane == 0) {
    out[2 * N] = looked;
    out[2 * N + 1] = acc;
  }
}

// ------------------------------------------------------------ Gram tail
// S^T L S without S.  L = D - A and the graph is symmetric, so
// C^T L C = 1/2 sum over directed edges (u, v) of delta delta^T with
// delta = C_u - C_v, that is the sum over the edges with u < v.  BFS levels differ by at most one across an edge, so
// every entry of delta is -1, 0 or 1 and follows from the levels mod 4:
// delta_c = ((d_c(u) - d_c(v)) mod 4) read as 0, 1 or -1 (3); a residue
// difference of 2 cannot come from BFS levels and flags the columns as
// unusable here.  The sum runs on the int8 matrix cores with exact int32
// accumulation; K = M^T (C^T L C) M and the other s x s products of the
// coefficient vectors are then formed in double-double arithmetic (the
// coefficient vectors of nearly dependent columns are large and of mixed
// sign, the products small).

// Per 2-bit field: (x - y) mod 4 (fields do not borrow from each other).
__device__ __forceinline__ unsigned field_sub(unsigned x, unsigned y) {
  constexpr unsigned H = 0xAAAAAAAAu;
  return ((x | H) - (y & ~H)) ^ ((x ^ ~y) & H);
}
// Fields equal to 2 (high bit set, low bit clear).
__device__ __forceinline__ unsigned field_two(unsigned d) { return d & ~(d << 1) & 0xAAAAAAAAu; }
// A field 0, 1, 3 as the int8 0, 1, -1 (0x00, 0x01, 0xff).
__device__ __forceinline__ unsigned field_i8(unsigned f) {
  const unsigned lo = f & 1u, hi = (f >> 1) & 1u;
  return lo | ((lo & hi) * 0xfeu);
}

constexpr int kLgThreads = 256;
constexpr int kLgPart = 3 * 1024;  // per wave: three 32 x 32 int32 tiles
typedef int lg_v16i __attribute__((ext_vector_type(16)));

// Per-wave partial sums of delta delta^T over a contiguous range of
// directed edges, 64 at a time: lane l forms the residue differences of
// edge l (gathering the two endpoints' packed residues; zero unless its
// source is below its target) into LDS, then for
// each 16-edge group lane l (column c = l % 32, edges 8 (l / 32) .. + 7)
// packs delta_c and delta_{c+32} of its 8 edges as int8 and the wave
// issues v_mfma_i32_32x32x16_i8 for the tiles (0,0), (0,1), (1,1) of the
// 64 x 64 result (lane l supplies A[l%32][8(l/32) + b] and B[8(l/32) + b]
// [l%32]; its output v is row 8(v/4) + 4(l/32) + v%4, column l%32).
// part[wave][tile][lane][16]; *bad |= 1 when a residue difference is 2.
__global__ __launch_bounds__(kLgThreads) void lap_gram_kernel(const uint32_t *esrc, const uint32_t *adj, int64_t m,
                                                             const uint4 *pack, int *part, int *bad) {
  __shared__ uint4 sd[kLgThreads / kWave][kWave];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int h = lane / 32, c = lane % 32;
  const int shift = 2 * (c % 16), wsel = c / 16;
  lg_v16i a00 = {}, a01 = {}, a11 = {};
  unsigned badl = 0;
  const int64_t gw = (int64_t)blockIdx.x * (kLgThreads / kWave) + wave;
  const int64_t nw = (int64_t)gridDim.x * (kLgThreads / kWave);
  const int64_t e0 = m * gw / nw, e1 = m * (gw + 1) / nw;
  const unsigned *sw = (const unsigned *)&sd[wave][0];
  for (int64_t eb = e0; eb < e1; eb += kWave) {
    const int64_t e = eb + lane;
    uint4 d = make_uint4(0u, 0u, 0u, 0u);
    // each undirected edge once (u < v): the graph is symmetric, so the
    // sum over the directed edges is twice this one
    const uint32_t u = e < e1 ? esrc[e] : 0u, v = e < e1 ? adj[e] : 0u;
    if (u < v) {
      const uint4 pu = pack[u], pv = pack[v];
      d = make_uint4(field_sub(pu.x, pv.x), field_sub(pu.y, pv.y), field_sub(pu.z, pv.z), field_sub(pu.w, pv.w));
      badl |= field_two(d.x) | field_two(d.y) | field_two(d.z) | field_two(d.w);
    }
    sd[wave][lane] = d;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int g = 0; g < 4; g++) {
      long x0 = 0, x1 = 0;
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const int k = 16 * g + 8 * h + b;
        const unsigned w0 = sw[4 * k + wsel], w1 = sw[4 * k + 2 + wsel];
        x0 |= (long)field_i8((w0 >> shift) & 3u) << (8 * b);
        x1 |= (long)field_i8((w1 >> shift) & 3u) << (8 * b);
      }
      a00 = __builtin_amdgcn_mfma_i32_32x32x16_i8(x0, x0, a00, 0, 0, 0);
      a01 = __builtin_amdgcn_mfma_i32_32x32x16_i8(x0, x1, a01, 0, 0, 0);
      a11 = __builtin_amdgcn_mfma_i32_32x32x16_i8(x1, x1, a11, 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  int *o = part + gw * kLgPart + lane * 16;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    o[v] = a00[v];
    o[1024 + v] = a01[v];
    o[2048 + v] = a11[v];
  }
  if (badl) atomicOr(bad, 1);
}

// KC[i][j] (64 x 64, both triangles) = the sum of the waves' tiles (each
// undirected edge taken once), one wave per entry, in 64-bit integers
// (exact).
__global__ void lap_sum_kernel(const int *part, int nparts, double *KC) {
  const int lane = threadIdx.x % kWave;
  const int e = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (e >= kLgPart) return;
  long long s = 0;
  for (int b = lane; b < nparts; b += kWave) s += part[(int64_t)b * kLgPart + e];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane != 0) return;
  const int t = e / 1024, l = (e / 16) % kWave, v = e % 16;
  const int bi = t == 2 ? 1 : 0, bj = t == 0 ? 0 : 1;
  const int i = 32 * bi + 8 * (v / 4) + 4 * (l / 32) + v % 4, j = 32 * bj + l % 32;
  const double val = (double)s;
  KC[i * 64 + j] = val;
  KC[j * 64 + i] = val;
}

// R = M^T X M (s x s, both triangles, R[a][b] at a * s + b) for a symmetric
// X (64 x 64 in LDS order [e][f]) and the coefficient vectors M ([k][64]),
// both steps in double-double: T = X M (rows e, column b, hi and lo kept),
// then R = M^T T.  With ones: also R1[b] = (X M)[0][b] (the D-products of
// the ones column of C with the columns).  One block of kGtThreads.
constexpr int kGtThreads = 256;
__device__ void gram_mxm(const double *X, const double *M, int s, double *Th, double *Tl, double *R, double *R1) {
#pragma clang fp contract(off)
  for (int q = threadIdx.x; q < 64 * s; q += kGtThreads) {
    const int e = q / s, b = q % s;
    double hi = 0, lo = 0;
    for (int f = 0; f < 64; f++) dd_fma(hi, lo, X[e * 64 + f], M[b * 64 + f]);
    const double t = hi + lo;
    Th[e * 64 + b] = t;
    Tl[e * 64 + b] = lo - (t - hi);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < s * s; q += kGtThreads) {
    const int a = q / s, b = q % s;
    if (a > b) continue;
    double hi = 0, lo = 0;
    for (int e = 0; e < 64; e++) {
      const double mv = M[a * 64 + e];
      dd_fma(hi, lo, mv, Th[e * 64 + b]);
      lo += mv * Tl[e * 64 + b];
    }
    const double r = hi + lo;
    R[a * s + b] = r;
    R[b * s + a] = r;
  }
  if (R1)
    for (int b = threadIdx.x; b < s; b += kGtThreads) R1[b] = Th[b];
  __syncthreads();
}

// K = M^T KC M; with B: B = M^T G_D M, and with Bf ((s+1) x (s+1)) also
// the D-Gram of [u0 1, S] (B bordered by the ones column).  G_D is W x W.
__global__ __launch_bounds__(kGtThreads) void gram_tail_kernel(const double *KC, const double *G, int W, const double *Mg,
                                                              int s, double u0, double *K, double *B, double *Bf) {
  extern __shared__ double tsm[];
  double *X = tsm, *M = X + 64 * 64, *Th = M + kGramMax * 64, *Tl = Th + 64 * 64, *R1 = Tl + 64 * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += kGtThreads) X[e] = KC[e];
  for (int e = threadIdx.x; e < s * 64; e += kGtThreads) M[e] = Mg[e];
  __syncthreads();
  gram_mxm(X, M, s, Th, Tl, K, nullptr);
  if (!B) return;
  for (int e = threadIdx.x; e < 64 * 64; e += kGtThreads) {
    const int i = e / 64, j = e % 64;
    X[e] = i < W && j < W ? G[W * W + i * W + j] : 0.0;
  }
  __syncthreads();
  gram_mxm(X, M, s, Th, Tl, B, R1);
  if (!Bf) return;
  for (int q = threadIdx.x; q < s * s; q += kGtThreads) Bf[(q / s + 1) * (s + 1) + q % s + 1] = B[q];
  if (threadIdx.x == 0) Bf[0] = u0 * u0 * X[0];
  for (int b = threadIdx.x; b < s; b += kGtThreads) {
    Bf[b + 1] = u0 * R1[b];
    Bf[(b + 1) * (s + 1)] = u0 * R1[b];
  }
}

// W[c][e] = sum_k M[k][e] V[k + c * s] (the coordinates' coefficients on
// the columns of C), in double-double, rounded once; lane e, one wave.
__global__ void gram_w_kernel(const double *M, const double *V, int s, int nc, double *W) {
#pragma clang fp contract(off)
  const int e = threadIdx.x;
  for (int c = 0; c < nc; c++) {
    double hi = 0, lo = 0;
    for (int k = 0; k < s; k++) dd_fma(hi, lo, M[k * 64 + e], V[k + c * s]);
    W[c * 64 + e] = hi + lo;
  }
}

// X[v][c] = sum_e C[v][e] W[c][e] (column-major n x nc), one row per
// thread, with per-block partial sums of X^2 (part[c * gridDim.x + b]).
// Column pointers past N repeat column N - 1 (W is zero there).
template <int WT>
__global__ __launch_bounds__(kOrthBlock) void gram_coords_kernel(GramCols cols, int64_t n, const double *__restrict__ W,
                                                                int nc, double *X, double *part) {
  __shared__ double sh[kOrthBlock / kWave][8];
  double ss[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    double d[WT];
    d[0] = 1.0;
#pragma unroll
    for (int e = 1; e < WT; e++) d[e] = (double)cols.col[e - 1][v];
    for (int c = 0; c < nc; c++) {
      const double *w = W + c * 64;
      double x = 0;
#pragma unroll
      for (int e = 0; e < WT; e++) x = fma(w[e], d[e], x);
      X[v + (int64_t)c * n] = x;
      ss[c] = fma(x, x, ss[c]);
    }
  }
  const int lane = threadIdx.x % kWave, wv = threadIdx.x / kWave;
  for (int c = 0; c < nc; c++) {
    double v = ss[c];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) sh[wv][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    double v = 0;
    for (int q = 0; q < kOrthBlock / kWave; q++) v += sh[q][threadIdx.x];
    part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = v;
  }
}

// S[:, k] = C M_k for the accepted columns k < ncols, one row per thread,
// the coefficients summed in order.  M_k is exactly zero past coefficient
// last[k] (its own candidate), so the sum stops at the 8-wide group that
// holds it (adding fma(0, d, s) = s changes nothing).  Column pointers past
// N repeat column N - 1, whose coefficients there are zero.

template <int WT>
__global__ __launch_bounds__(kOrthBlock) void gram_form_kernel(GramCols cols, GramForm fm, int64_t n,
                                                              const double *__restrict__ M, int ncols, double *S) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  double d[WT];
  d[0] = 1.0;
#pragma unroll
  for (int e = 1; e < WT; e++) d[e] = (double)cols.col[e - 1][v];
  for (int k = 0; k < ncols; k++) {
    const double *m = M + k * 64;
    const int last = fm.last[k];
    double s = 0;
#pragma unroll
    for (int g = 0; g < WT / 8; g++) {
      if (g * 8 <= last) {
#pragma unroll
        for (int i = 0; i < 8; i++) s = fma(m[g * 8 + i], d[g * 8 + i], s);
      }
    }
    S[(int64_t)k * n + v] = s;
  }
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
  PHDE_TRY_RC(dalloc((void **)&S, esz * (size_t)n * s));
  PHDE_TRY_RC(dalloc((void **)&a, sizeof(double) * n));
  PHDE_TRY_RC(dalloc((void **)&part, sizeof(double) * (size_t)nblk * (kOnesSlot + 2) * 2));
  PHDE_TRY_RC(dalloc((void **)&den, sizeof(double) * (kOnesSlot + 1)));
  PHDE_TRY_RC(dalloc((void **)&coef, sizeof(double) * (kOnesSlot + 1)));
  PHDE_TRY_RC(dalloc((void **)&scal, sizeof(double) * 8));
  PHDE_TRY_RC(halloc((void **)&h_scal, sizeof(double) * 8));
  return PHDE_OK;
}

void Orth::release() {
  for (auto &b : dev_bufs) tmp_free(b.first, b.second);
  for (auto &b : host_bufs) tmp_host_free(b.first, b.second);
  dev_bufs.clear();
  host_bufs.clear();
  S = nullptr;
  a = part = den = coef = scal = h_scal = ones = nullptr;
  blkA = blkAcc = blkPart = blkCoef = nullptr;
  gPart = gG = gM = gOut = h_gOut = gKC = gW = nullptr;
  gMax = gLapPart = gBad = nullptr;
  gPack = nullptr;
  gram_pending = false;
}

int Orth::dalloc(void **p, size_t bytes) {
  PHDE_TRY_RC(tmp_malloc(p, bytes));
  dev_bufs.emplace_back(*p, bytes);
  return PHDE_OK;
}

int Orth::halloc(void **p, size_t bytes) {
  PHDE_TRY_RC(tmp_host_malloc(p, bytes));
  host_bufs.emplace_back(*p, bytes);
  return PHDE_OK;
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
    PHDE_TRY_RC(dalloc((void **)&ones, sizeof(double) * n));
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
    PHDE_TRY_RC(dalloc((void **)&blkA, sizeof(double) * (size_t)n * kCgsBlock));
    PHDE_TRY_RC(dalloc((void **)&blkAcc, sizeof(double) * (size_t)n * kCgsBlock));
    PHDE_TRY_RC(dalloc((void **)&blkPart, sizeof(double) * (size_t)nblk * (kOnesSlot + 1) * kCgsBlock));
    PHDE_TRY_RC(dalloc((void **)&blkCoef, sizeof(double) * (size_t)(kOnesSlot + 1) * kCgsBlock));
  }
  // FP32 storage: CGS2 (two projections per candidate); FP64: one, bit for
  // bit as the candidate-by-candidate kernels
  return fp32 ? cgs2_block_t<float>(*this, cols, scales, nc, s_target, accepted, taken, norms)
              : cgs_block_t<double>(*this, cols, scales, nc, s_target, accepted, taken, norms);
}

template <int WT>
static void gram_form(Orth &o, const GramCols &gc, const GramForm &fm, int ncols) {
  const int64_t blocks = (o.n + kOrthBlock - 1) / kOrthBlock;
  gram_form_kernel<WT><<<(unsigned)blocks, kOrthBlock, 0, o.ctx->stream>>>(gc, fm, o.n, o.gM, ncols, (double *)o.S);
}

template <int WT>
static void gram_coords_launch(Orth &o, const GramCols &gc, int nb, int nc, double *X, double *part) {
  gram_coords_kernel<WT><<<nb, kOrthBlock, 0, o.ctx->stream>>>(gc, o.n, o.gW, nc, X, part);
}

// Width of the unrolled coefficient loops: N + 1 columns rounded up to 8.
static int gram_width(int N) { return (N + 1 + 7) / 8 * 8; }

int Orth::gram_cgs(const int32_t *const *cols, const double *sumsq, int nc, int s_target, int &accepted,
                   std::vector<bool> &taken, std::vector<double> &norms, int &consumed) {
  if (nc < 1 || nc > kGramMax || accepted != 0 || fp32) {
    set_error("internal: Gram CGS arguments");
    return PHDE_ERR_ARG;
  }
  hipStream_t st = ctx->stream;
  const int W = (nc + 1 + 3) / 4 * 4;
  if (!gPart) {
    const int64_t nsteps = (n + 3) / 4;
    gBlocks = (int)std::max<int64_t>(1, std::min<int64_t>((nsteps + 3) / 4, 2 * std::max(ctx->num_cus, 1)));
    PHDE_TRY_RC(dalloc((void **)&gPart, sizeof(double) * (size_t)gBlocks * (kGcThreads / kWave) * kGcPart));
    PHDE_TRY_RC(dalloc((void **)&gMax, sizeof(int) * (gBlocks + 1)));
    PHDE_TRY_RC(dalloc((void **)&gG, sizeof(double) * 2 * 64 * 64));
    PHDE_TRY_RC(dalloc((void **)&gM, sizeof(double) * kGramMax * 64));
    PHDE_TRY_RC(dalloc((void **)&gOut, sizeof(double) * (2 * kGramMax + 2)));
    PHDE_TRY_RC(halloc((void **)&h_gOut, sizeof(double) * (2 * kGramMax + 2)));
    PHDE_TRY_RC(dalloc((void **)&gPack, sizeof(uint4) * (size_t)n));
    PHDE_HIP(hipFuncSetAttribute((const void *)gram_cgs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * (2 * 64 * 64 + kGramMax * 64 + 64 * 64 + 64))));
  }
  GramCols gc{};
  for (int b = 0; b < kGramMax; b++) {
    gc.col[b] = cols[std::min(b, nc - 1)];
    gc.scale[b] = b < nc ? 1 / std::sqrt(sumsq[b]) : 0.0;
    gc.sumsq[b] = b < nc ? sumsq[b] : 0.0;
  }
  gram_tiles_kernel<<<gBlocks, kGcThreads, 0, st>>>(gc, nc, deg_, n, gPart, gMax, (uint4 *)gPack);
  gram_sum_kernel<<<kGcPart / 4, 4 * kWave, 0, st>>>(gPart, gMax, gBlocks, W, gG, gMax + gBlocks);
  const size_t cgs_lds = sizeof(double) * (2 * W * 64 + kGramMax * 64 + 64 * 64 + 64);
  gram_cgs_kernel<<<1, kWave, cgs_lds, st>>>(gG, gMax + gBlocks, nc, W, gc, u0, s_target, gM, den, gOut);
  PHDE_HIP(hipGetLastError());
  PHDE_HIP(hipMemcpyAsync(h_gOut, gOut, sizeof(double) * (2 * nc + 2), hipMemcpyDeviceToHost, st));
  PHDE_HIP(hipStreamSynchronize(st));
  consumed = (int)h_gOut[2 * nc];
  const int acc = (int)h_gOut[2 * nc + 1];
  taken.assign(consumed, false);
  norms.assign(consumed, 0.0);
  for (int b = 0; b < consumed; b++) {
    norms[b] = h_gOut[b];
    taken[b] = h_gOut[nc + b] != 0.0;
  }
  gCols = gc;
  gForm = GramForm{};
  for (int b = 0, k = 0; b < consumed; b++)
    if (taken[b]) gForm.last[k++] = b + 1;
  gN = nc;
  gAcc = acc;
  accepted = acc;
  // every column accepted: S stays in coefficient form for the tail;
  // otherwise the vector CGS continues on S
  gram_pending = acc > 0;
  if (acc < s_target) return form_S();
  return PHDE_OK;
}

int Orth::form_S() {
  if (!gram_pending) return PHDE_OK;
  gram_pending = false;
  switch (gram_width(gN)) {
    case 8: gram_form<8>(*this, gCols, gForm, gAcc); break;
    case 16: gram_form<16>(*this, gCols, gForm, gAcc); break;
    case 24: gram_form<24>(*this, gCols, gForm, gAcc); break;
    case 32: gram_form<32>(*this, gCols, gForm, gAcc); break;
    case 40: gram_form<40>(*this, gCols, gForm, gAcc); break;
    case 48: gram_form<48>(*this, gCols, gForm, gAcc); break;
    case 56: gram_form<56>(*this, gCols, gForm, gAcc); break;
    default: gram_form<64>(*this, gCols, gForm, gAcc); break;
  }
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

// Source vertex of every directed edge (once per context).
__global__ void edge_source_kernel(const uint32_t *rowptr, int64_t n, uint32_t *esrc) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride)
    for (uint32_t e = rowptr[u]; e < rowptr[u + 1]; e++) esrc[e] = (uint32_t)u;
}

int Orth::gram_lap(double *K_dev, double *B_dev, double *Bfull_host, bool &ok) {
  ok = false;
  if (!gram_pending) {
    set_error("internal: Gram tail without Gram columns");
    return PHDE_ERR_ARG;
  }
  hipStream_t st = ctx->stream;
  const int64_t m = ctx->g.m;
  if (!ctx->d_esrc) {
    PHDE_HIP(hipMalloc(&ctx->d_esrc, sizeof(uint32_t) * (size_t)std::max<int64_t>(m, 1)));
    edge_source_kernel<<<grid_for(n, 16 * std::max(ctx->num_cus, 1)), kOrthBlock, 0, st>>>(ctx->g.rowptr, n,
                                                                                           ctx->d_esrc);
  }
  if (!gLapPart) {
    gLapBlocks = (int)std::max<int64_t>(1, std::min<int64_t>((m + 1023) / 1024, 2 * std::max(ctx->num_cus, 1)));
    PHDE_TRY_RC(dalloc((void **)&gLapPart, sizeof(int) * (size_t)gLapBlocks * (kLgThreads / kWave) * kLgPart));
    PHDE_TRY_RC(dalloc((void **)&gBad, sizeof(int) * 4));
    PHDE_TRY_RC(dalloc((void **)&gKC, sizeof(double) * 64 * 64));
    PHDE_TRY_RC(dalloc((void **)&gW, sizeof(double) * 8 * 64));
    PHDE_HIP(hipFuncSetAttribute((const void *)gram_tail_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(sizeof(double) * (3 * 64 * 64 + kGramMax * 64 + 64))));
  }
  PHDE_TRY_RC(fill_i32(ctx, gBad, 1, 0));
  lap_gram_kernel<<<gLapBlocks, kLgThreads, 0, st>>>(ctx->d_esrc, ctx->g.adj, m, (const uint4 *)gPack, gLapPart, gBad);
  lap_sum_kernel<<<kLgPart / 4, 4 * kWave, 0, st>>>(gLapPart, gLapBlocks * (kLgThreads / kWave), gKC);
  PHDE_HIP(hipGetLastError());
  int bad = 0;
  PHDE_HIP(hipMemcpyAsync(&bad, gBad, sizeof(int), hipMemcpyDeviceToHost, st));
  PHDE_HIP(hipStreamSynchronize(st));
  if (bad) return PHDE_OK;
  const int s_ = gAcc;
  const int W = (gN + 1 + 3) / 4 * 4;
  DBuf<double> Bs, Bf;
  double *Bp = B_dev;
  if (Bfull_host) {
    PHDE_TRY_RC(Bf.alloc((size_t)(s_ + 1) * (s_ + 1)));
    if (!Bp) {
      PHDE_TRY_RC(Bs.alloc((size_t)s_ * s_));
      Bp = Bs.p;
    }
  }
  const size_t lds = sizeof(double) * (3 * 64 * 64 + kGramMax * 64 + 64);
  gram_tail_kernel<<<1, kGtThreads, lds, st>>>(gKC, gG, W, gM, s_, u0, K_dev, Bp, Bf.p);
  PHDE_HIP(hipGetLastError());
  if (Bfull_host) {
    PHDE_HIP(hipMemcpyAsync(Bfull_host, Bf.p, sizeof(double) * (s_ + 1) * (s_ + 1), hipMemcpyDeviceToHost, st));
    PHDE_HIP(hipStreamSynchronize(st));
  }
  ok = true;
  return PHDE_OK;
}

int Orth::gram_coords(const double *V_dev, int nc, double *X_dev, double *part, double *inv) {
  if (!gram_pending || nc < 1 || nc > 8) {
    set_error("internal: Gram coordinates arguments");
    return PHDE_ERR_ARG;
  }
  hipStream_t st = ctx->stream;
  hipEvent_t ev;
  ctx->timer.begin(st, PHDE_KERNEL_COORDS, &ev);
  gram_w_kernel<<<1, kWave, 0, st>>>(gM, V_dev, gAcc, nc, gW);
  const int nb = grid_for(n, 4 * std::max(ctx->num_cus, 1));
  switch (gram_width(gN)) {
    case 8: gram_coords_launch<8>(*this, gCols, nb, nc, X_dev, part); break;
    case 16: gram_coords_launch<16>(*this, gCols, nb, nc, X_dev, part); break;
    case 24: gram_coords_launch<24>(*this, gCols, nb, nc, X_dev, part); break;
    case 32: gram_coords_launch<32>(*this, gCols, nb, nc, X_dev, part); break;
    case 40: gram_coords_launch<40>(*this, gCols, nb, nc, X_dev, part); break;
    case 48: gram_coords_launch<48>(*this, gCols, nb, nc, X_dev, part); break;
    case 56: gram_coords_launch<56>(*this, gCols, nb, nc, X_dev, part); break;
    default: gram_coords_launch<64>(*this, gCols, nb, nc, X_dev, part); break;
  }
  PHDE_HIP(hipGetLastError());
  PHDE_TRY_RC(coords_normalise(ctx, X_dev, nc, part, nb, inv));
  ctx->timer.end(st, PHDE_KERNEL_COORDS, ev);
  return PHDE_OK;
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
