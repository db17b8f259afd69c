// mds.hip — the arithmetic of the PivotMDS / PHDE stage (centring and
// C^T C), compiled without FMA contraction (-ffp-contract=off) so that
// every product and sum rounds on its own, in the reference's order.
//
#include <algorithm>

#include "device_util.h"
#include "mds.h"

namespace phde {

// The power method the reference runs on C^T C (Largest) stops at a loose
// |cos| >= 0.999 and C^T C has a near-null direction (the centring makes
// C 1 ~ 0), so the coordinates follow the rounding of C^T C closely.  This
// stage therefore reproduces the reference's arithmetic exactly: the
// distance columns are moved to the caller's numbering once, the centring
// uses individually rounded products and sums (no FMA contraction, in the
// reference's operation order), and every entry of C^T C is one sequential
// j-ascending chain of rounded products and sums (DrawingBase.cpp:221-244).

// Column sums (of d^2 when squares) exactly as 128-bit integers (acc[2 i],
// acc[2 i + 1]; every term is a non-negative integer, so a total below
// 2^53 is the reference's sequential double sum exactly), 16 columns per
// blockIdx.y; with squares, blockIdx.y == 0 also writes the row sums of
// d^2 over all columns in order (PivotMDS.cpp:39-50: row j accumulates
// column by column).
constexpr int kCenterCols = 16;
__global__ void center_sums_kernel(const int32_t *D, int64_t n, int s, bool squares, double *rowterm,
                                   unsigned long long *acc) {
  const int i0 = blockIdx.y * kCenterCols;
  const int ic = min(kCenterCols, s - i0);
  U128 cp[kCenterCols];
#pragma unroll
  for (int q = 0; q < kCenterCols; q++) cp[q] = U128{0, 0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < kCenterCols; q++) {
      if (q < ic) {
        const int64_t d = D[(int64_t)(i0 + q) * n + j];
        u128_add(cp[q], squares ? (unsigned long long)(d * d) : (unsigned long long)d);
      }
    }
    if (squares && blockIdx.y == 0) {
      double rs = 0;
      for (int i = 0; i < s; i++) {
        const double d = (double)D[(int64_t)i * n + j];
        rs = __dadd_rn(rs, __dmul_rn(d, d));
      }
      rowterm[j] = rs;
    }
  }
#pragma unroll
  for (int q = 0; q < kCenterCols; q++) {
    if (q < ic) {
      const U128 v = wave_sum_u128(cp[q]);
      if (__lane_id() == 0) atomic_add_u128(acc + 2 * (i0 + q), v);
    }
  }
}

// The sequential double sum of a column (of d^2 when squares), for columns
// whose total reaches 2^53 (there the order of the roundings matters): one
// thread per column, j ascending.
__global__ void column_sum_seq_kernel(const int32_t *D, int64_t n, int s, bool squares, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s) return;
  const int32_t *col = D + (int64_t)i * n;
  double v = 0;
  for (int64_t j = 0; j < n; j++) {
    const double d = (double)col[j];
    v = __dadd_rn(v, squares ? __dmul_rn(d, d) : d);
  }
  out[i] = v;
}

// C[j + i n] = -0.5 * (d^2 - rN colsum_i - rK rowsum_j + rK rN total)
// (PivotMDS.cpp:51-63) when dbl, else d - colmean_i (PHDE.cpp:30-47); each
// operation rounded on its own, left to right.
__global__ void center_kernel(const int32_t *D, int64_t n, int s, const double *colterm,
                              const double *rowterm, double rN, double rK, double total, bool dbl,
                              double *C) {
  const double rKN_total = __dmul_rn(__dmul_rn(rK, rN), total);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * s;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e % n;
    const int i = (int)(e / n);
    const double d = (double)D[e];
    if (dbl) {
      double v = __dsub_rn(__dmul_rn(d, d), __dmul_rn(rN, colterm[i]));
      v = __dsub_rn(v, __dmul_rn(rK, rowterm[j]));
      v = __dadd_rn(v, rKN_total);
      C[e] = __dmul_rn(-0.5, v);
    } else {
      C[e] = __dsub_rn(d, colterm[i]);
    }
  }
}

// K[k + i s] = K[i + k s] = sum_j C[j + i n] C[j + k n], j ascending, each
// product and partial sum rounded (matrixTransMultiplication,
// DrawingBase.cpp:221-244, with both operands C).  One wave per 8 x 8 tile
// (i in column group p, k in group q, p <= q), lane (i % 8, k % 8); rows
// stream through LDS 128 at a time.
constexpr int kCtcRows = 128;
__global__ __launch_bounds__(64) void ctc_seq_kernel(const double *C, int64_t n, int s, double *K) {
  __shared__ double tile[kCtcRows][17];
  // blockIdx.x -> (p, q), p <= q, in row-major upper-triangle order
  const int G = (s + 7) / 8;
  int p = 0, rem = blockIdx.x;
  while (rem >= G - p) {
    rem -= G - p;
    p++;
  }
  const int q = p + rem;
  const int lane = threadIdx.x;
  const int ii = lane / 8, kk = lane % 8;
  const int i = 8 * p + ii, k = 8 * q + kk;
  const bool on = i < s && k < s && (p < q || ii <= kk);
  double acc = 0;
  for (int64_t j0 = 0; j0 < n; j0 += kCtcRows) {
    const int rows = (int)min((int64_t)kCtcRows, n - j0);
    for (int t = lane; t < 16 * kCtcRows; t += 64) {
      const int cc = t / kCtcRows, r = t % kCtcRows;
      const int col = cc < 8 ? 8 * p + cc : 8 * q + cc - 8;
      tile[r][cc] = (r < rows && col < s) ? C[(int64_t)col * n + j0 + r] : 0.0;
    }
    __syncthreads();
    for (int r = 0; r < rows; r++) acc = __dadd_rn(acc, __dmul_rn(tile[r][ii], tile[r][8 + kk]));
    __syncthreads();
  }
  if (on) {
    K[k + (int64_t)i * s] = acc;
    K[i + (int64_t)k * s] = acc;
  }
}

int mds_center_sums(phde_ctx *c, const int32_t *D, int64_t n, int s, bool squares, double *rowterm,
                    unsigned long long *acc) {
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)c->num_cus * 4));
  center_sums_kernel<<<dim3(nb, (s + kCenterCols - 1) / kCenterCols), 256, 0, c->stream>>>(D, n, s, squares,
                                                                                            rowterm, acc);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int mds_column_sums_seq(phde_ctx *c, const int32_t *D, int64_t n, int s, bool squares, double *out) {
  column_sum_seq_kernel<<<1, 64, 0, c->stream>>>(D, n, s, squares, out);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int mds_center(phde_ctx *c, const int32_t *D, int64_t n, int s, const double *colterm, const double *rowterm,
               double rN, double rK, double total, bool dbl, double *C) {
  center_kernel<<<(int)std::min<int64_t>((n * s + 255) / 256, 16384), 256, 0, c->stream>>>(
      D, n, s, colterm, rowterm, rN, rK, total, dbl, C);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int mds_ctc(phde_ctx *c, const double *C, int64_t n, int s, double *K) {
  const int G = (s + 7) / 8;
  ctc_seq_kernel<<<G * (G + 1) / 2, 64, 0, c->stream>>>(C, n, s, K);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

}  // namespace phde
