// spmm.hip — Laplacian times the pivot matrix, the s x s Gram product,
// coordinates and Hall energy, for gfx950.
//
// Reference behaviour followed:
//  - L*S (ParHDE.cpp:288-300, WParHDE.cpp:229-241): LS_ik = deg_i S_ik minus
//    sum over neighbours j of (w_ij) S_jk, subtracted in adjacency order
//    (rows longer than kHubRow: chunk by chunk, see below).
//  - S^T(LS) (DrawingBase.cpp:221-244 matrixTransMultiplication):
//    C[k + i*s] = sum_j S_ji LS_jk, the transposed storage the eigensolver
//    consumes.
//  - computeCoordinates (DrawingBase.cpp:207-219) and getCoordinates'
//    column normalisation (DrawingBase.cpp:196-205).
//
// Layout: the SpMM gathers whole neighbour rows, so it works on a row-major
// copy of S (n x s, row i contiguous): one wave per vertex, lane k holds
// column k (s <= 64 = wave width), each neighbour costs one coalesced
// s-wide row read.  Rows of more than kHubRow neighbours (the hubs of a
// power-law graph) would leave one wave gathering for the whole kernel:
// they are cut into chunks of kHubRow neighbours, one wave each, whose
// in-order partial sums a second pass subtracts in chunk order.  The Gram product reads row tiles of S and LS through
// LDS; per-block s*s partials are summed in block order (deterministic).
#include <algorithm>
#include <vector>

#include "phde_common.h"
#include "spmm.h"
#include "traverse.h"

namespace phde {

constexpr int kBlk = 256;
constexpr uint32_t kHubRow = 256;  // SpMM: longer rows are split into chunks of this many neighbours

// S column-major [s][n] -> row-major [n][s]; 64 rows per block via LDS.
template <typename T>
__global__ __launch_bounds__(kBlk) void col_to_row_kernel(const T *Sc, T *Sr, int64_t n, int s) {
  __shared__ T tile[64][65];
  const int t = threadIdx.x;
  for (int64_t r0 = (int64_t)blockIdx.x * 64; r0 < n; r0 += (int64_t)gridDim.x * 64) {
    const int rows = (int)std::min<int64_t>(64, n - r0);
    for (int k = t / 64; k < s; k += kBlk / 64) {
      const int r = t % 64;
      if (r < rows) tile[r][k] = Sc[(int64_t)k * n + r0 + r];
    }
    __syncthreads();
    for (int e = t; e < rows * s; e += kBlk) {
      const int r = e / s, k = e % s;
      Sr[(r0 + r) * s + k] = tile[r][k];
    }
    __syncthreads();
  }
}

// v -= (w_ij) S[j][lane] for the neighbours j in adj[b, e), in order; the
// gathers of kSpmmGroup neighbours are in flight together.
constexpr int kSpmmGroup = 8;
template <typename T, bool W>
__device__ __forceinline__ double subtract_neighbours(const DevGraph &g, const T *S, int s, int lane,
                                                      uint32_t b, uint32_t e, double v) {
  const bool on = lane < s;
  for (uint32_t e0 = b; e0 < e; e0 += kWave) {
    const uint32_t cnt = min((uint32_t)kWave, e - e0);
    const uint32_t my = (lane < (int)cnt) ? g.adj[e0 + lane] : 0u;
    int32_t myw = 0;
    if (W) myw = (lane < (int)cnt) ? g.w[e0 + lane] : 0;
    uint32_t t = 0;
    for (; t + kSpmmGroup <= cnt; t += kSpmmGroup) {
      double x[kSpmmGroup];
#pragma unroll
      for (int q = 0; q < kSpmmGroup; q++) {
        const uint32_t j = __shfl(my, t + q);
        x[q] = on ? (double)S[(int64_t)j * s + lane] : 0.0;
      }
      if (W) {
#pragma unroll
        for (int q = 0; q < kSpmmGroup; q++) x[q] *= (double)__shfl(myw, t + q);
      }
#pragma unroll
      for (int q = 0; q < kSpmmGroup; q++) v -= x[q];
    }
    for (; t < cnt; t++) {
      const uint32_t j = __shfl(my, t);
      double x = on ? (double)S[(int64_t)j * s + lane] : 0.0;
      if (W) x *= (double)__shfl(myw, t);
      v -= x;
    }
  }
  return v;
}

// LS = L * S, row-major in and out.  One wave per row, lane = column; rows
// longer than kHubRow are left to the two hub kernels below.
template <typename T, bool W>
__global__ __launch_bounds__(kBlk) void laplacian_spmm_kernel(DevGraph g, const T *S, T *LS, int s) {
  const int lane = threadIdx.x % kWave;
  const int64_t wave = ((int64_t)blockIdx.x * kBlk + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlk / kWave);
  const bool on = lane < s;
  for (int64_t i = wave; i < g.n; i += nwaves) {
    const uint32_t b = g.rowptr[i], e = g.rowptr[i + 1];
    if (e - b > kHubRow) continue;
    double v = on ? g.degw[i] * (double)S[i * s + lane] : 0.0;
    v = subtract_neighbours<T, W>(g, S, s, lane, b, e, v);
    if (on) LS[i * s + lane] = (T)v;
  }
}

// Hub rows, pass 1: one wave per chunk of kHubRow neighbours, its in-order
// partial (0 minus the chunk's terms) to part[chunk][lane].
template <typename T, bool W>
__global__ __launch_bounds__(kBlk) void hub_partial_kernel(DevGraph g, const T *S, int s,
                                                           const uint32_t *chunk_row,
                                                           const uint32_t *chunk_beg, int64_t nchunks,
                                                           double *part) {
  const int lane = threadIdx.x % kWave;
  const int64_t wave = ((int64_t)blockIdx.x * kBlk + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlk / kWave);
  for (int64_t c = wave; c < nchunks; c += nwaves) {
    const uint32_t r = chunk_row[c], b = chunk_beg[c];
    const uint32_t e = min(b + kHubRow, g.rowptr[r + 1]);
    const double p = subtract_neighbours<T, W>(g, S, s, lane, b, e, 0.0);
    if (lane < s) part[c * s + lane] = p;
  }
}

// Hub rows, pass 2: deg_i S_i plus the row's chunk partials in chunk order.
template <typename T>
__global__ __launch_bounds__(kBlk) void hub_finish_kernel(DevGraph g, const T *S, T *LS, int s,
                                                          const uint32_t *hub_row, const uint32_t *hub_c0,
                                                          int64_t nhubs, const double *part) {
  const int lane = threadIdx.x % kWave;
  const int64_t wave = ((int64_t)blockIdx.x * kBlk + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlk / kWave);
  if (lane >= s) return;
  for (int64_t h = wave; h < nhubs; h += nwaves) {
    const int64_t i = hub_row[h];
    double v = g.degw[i] * (double)S[i * s + lane];
    for (uint32_t c = hub_c0[h]; c < hub_c0[h + 1]; c++) v += part[(int64_t)c * s + lane];
    LS[i * s + lane] = (T)v;
  }
}

// Blocks of the matrix-core Gram: two per CU, fewer for a small n (4 rows
// per wave and step).
static int gram_mfma_blocks(const phde_ctx *c, int64_t n) {
  const int64_t need = ((n + 3) / 4 + (kBlk / kWave) - 1) / (kBlk / kWave);
  return (int)std::max<int64_t>(1, std::min<int64_t>(need, 2 * (int64_t)std::max(c->num_cus, 1)));
}

// Per-block partial C[k + i*s] = sum_rows A[r][i] * B[r][k] over row tiles
// staged in LDS (A, B row-major n x s).
constexpr int kGramRows = 16;
constexpr int kGramPer = 16;  // outputs per thread (s*s <= 4096)
template <typename T>
__global__ __launch_bounds__(kBlk) void gram_partial_kernel(const T *A, const T *B, int64_t n, int s,
                                                           double *part, const double *rw = nullptr) {
  __shared__ double ta[kGramRows][64];
  __shared__ double tb[kGramRows][64];
  const int t = threadIdx.x;
  const int ss = s * s;
  double acc[kGramPer];
  int ii[kGramPer], kk[kGramPer];
#pragma unroll
  for (int q = 0; q < kGramPer; q++) {
    acc[q] = 0;
    const int e = t + q * kBlk;
    ii[q] = e < ss ? e / s : 0;
    kk[q] = e < ss ? e % s : 0;
  }
  for (int64_t r0 = (int64_t)blockIdx.x * kGramRows; r0 < n; r0 += (int64_t)gridDim.x * kGramRows) {
    const int rows = (int)std::min<int64_t>(kGramRows, n - r0);
    for (int e = t; e < kGramRows * 64; e += kBlk) {
      const int r = e / 64, k = e % 64;
      const bool ok = r < rows && k < s;
      ta[r][k] = ok ? (double)A[(r0 + r) * s + k] * (rw ? rw[r0 + r] : 1.0) : 0.0;
      tb[r][k] = ok ? (double)B[(r0 + r) * s + k] : 0.0;
    }
    __syncthreads();
    for (int r = 0; r < rows; r++) {
#pragma unroll
      for (int q = 0; q < kGramPer; q++) acc[q] += ta[r][ii[q]] * tb[r][kk[q]];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < kGramPer; q++) {
    const int e = t + q * kBlk;
    if (e < ss) part[(int64_t)blockIdx.x * ss + e] = acc[q];
  }
}

__global__ void gram_reduce_kernel(const double *part, int nblk, int ss, double *C) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ss) return;
  double s = 0;
  for (int b = 0; b < nblk; b++) s += part[(int64_t)b * ss + e];
  // part holds element (k = e % s, i = e / s) of sum_j A[j][i] B[j][k]; the
  // reference stores it at C[k + i*s] which is exactly index e.
  C[e] = s;
}

// The same partial Gram sums on the FP64 matrix cores: a wave takes 4 rows
// at a time and issues one v_mfma_f64_16x16x4f64 per 16 x 16 tile of the
// (s <= 64, padded to 64) x 64 result, the 16 tiles' sums in registers; each
// wave writes its own partial (part[wave][i*s + k]), which gram_reduce_kernel
// sums in a fixed order.  Lane l supplies A[row r0 + l/16][16*mi + l%16]
// and B[row r0 + l/16][16*ni + l%16]; its output v is row l/16 + 4*v,
// column l%16 of each tile (measured lane layout of the instruction).
typedef double gram_v4d __attribute__((ext_vector_type(4)));
template <typename T>
__global__ __launch_bounds__(kBlk) void gram_mfma_kernel(const T *A, const T *B, int64_t n, int s, double *part,
                                                        const double *rw = nullptr) {
  const int lane = threadIdx.x % kWave;
  const int64_t wave = ((int64_t)blockIdx.x * kBlk + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * (kBlk / kWave);
  const int col = lane % 16, kr = lane / 16;
  gram_v4d acc[4][4];
#pragma unroll
  for (int mi = 0; mi < 4; mi++)
#pragma unroll
    for (int ni = 0; ni < 4; ni++) acc[mi][ni] = gram_v4d{0.0, 0.0, 0.0, 0.0};
  const int64_t nsteps = (n + 3) / 4;
  for (int64_t st = wave; st < nsteps; st += nwaves) {
    const int64_t r = st * 4 + kr;
    const bool rok = r < n;
    const double wr = (rw != nullptr && rok) ? rw[r] : 1.0;
    double av[4], bv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int c = q * 16 + col;
      const bool ok = rok && c < s;
      av[q] = ok ? (double)A[r * s + c] * wr : 0.0;
      bv[q] = ok ? (double)B[r * s + c] : 0.0;
    }
#pragma unroll
    for (int mi = 0; mi < 4; mi++)
#pragma unroll
      for (int ni = 0; ni < 4; ni++) acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mi], bv[ni], acc[mi][ni], 0, 0, 0);
  }
  double *out = part + wave * (int64_t)s * s;
#pragma unroll
  for (int mi = 0; mi < 4; mi++)
#pragma unroll
    for (int ni = 0; ni < 4; ni++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = mi * 16 + kr + 4 * v, k = ni * 16 + col;
        if (i < s && k < s) out[i * s + k] = acc[mi][ni][v];
      }
}

// Partial Gram sums of A^T B (row-major n x s; A's rows scaled by rw when
// given) into part, then C = their sum in a fixed order.  The matrix-core
// kernel unless PHDE_GRAM_MFMA=0 or s > 64.
template <typename T>
static int gram_launch(phde_ctx *c, const T *A, const T *B, int64_t n, int s, double *part, int nblk_gram,
                       double *C, const double *rw = nullptr) {
  static const bool mfma = [] {
    const char *e = getenv("PHDE_GRAM_MFMA");
    return !(e && atoi(e) == 0);
  }();
  int nparts = nblk_gram;
  if (mfma && s <= 64) {
    const int blocks = gram_mfma_blocks(c, n);
    gram_mfma_kernel<T><<<blocks, kBlk, 0, c->stream>>>(A, B, n, s, part, rw);
    nparts = blocks * (kBlk / kWave);
  } else {
    gram_partial_kernel<T><<<nblk_gram, kBlk, 0, c->stream>>>(A, B, n, s, part, rw);
  }
  gram_reduce_kernel<<<(s * s + 255) / 256, 256, 0, c->stream>>>(part, nparts, s * s, C);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

// X[i + k*n] = sum_j S[i + j*n] V[j + k*s] (S column-major), and per-block
// sums of X^2 per coordinate.
template <typename T>
__global__ __launch_bounds__(kBlk) void coords_kernel(const T *S, const double *V, int64_t n, int s,
                                                     int c, double *X, double *part) {
  __shared__ double sv[64 * 8];
  __shared__ double sh[kBlk / kWave][8];
  for (int e = threadIdx.x; e < s * c; e += kBlk) sv[e] = V[e];
  __syncthreads();
  double ss[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * kBlk + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlk) {
    for (int k = 0; k < c; k++) {
      double x = 0;
      for (int j = 0; j < s; j++) x += (double)S[(int64_t)j * n + i] * sv[j + k * s];
      X[i + (int64_t)k * n] = x;
      ss[k] += x * x;
    }
  }
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  for (int k = 0; k < c; k++) {
    double v = ss[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) sh[w][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < c) {
    double v = 0;
    for (int q = 0; q < kBlk / kWave; q++) v += sh[q][threadIdx.x];
    part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = v;
  }
}

// inv[k] = 1 / sqrt(sum_b part[k][b]) (getCoordinates, DrawingBase.cpp:200);
// wave k sums column k's partials (lane-strided, then a fixed shuffle tree).
__global__ void coords_norm_kernel(const double *part, int nblk, int c, double *inv) {
  const int k = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (k >= c) return;
  double s = 0;
  for (int b = lane; b < nblk; b += kWave) s += part[(int64_t)k * nblk + b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) inv[k] = 1 / sqrt(s);
}

__global__ void scale_cols_kernel(double *X, int64_t n, int c, const double *inv) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * c;
       e += (int64_t)gridDim.x * blockDim.x)
    X[e] *= inv[e / n];
}

// Hall energy partials: num = sum over directed edges of w (y_i - y_j)^2 / 2,
// den = sum_i degw_i y_i^2, over the c columns of Y (column-major).
__global__ __launch_bounds__(kBlk) void hall_kernel(DevGraph g, const double *Y, int c, double *part) {
  __shared__ double sh[kBlk / kWave][2];
  double num = 0, den = 0;
  // one vertex per lane; rows of 64 or more neighbours (hubs) are summed by
  // the whole wave so that no lane walks a long row alone
  const int lane0 = threadIdx.x % kWave;
  const int64_t stride = (int64_t)gridDim.x * kBlk;
  for (int64_t i0 = (int64_t)blockIdx.x * kBlk + (threadIdx.x - lane0); i0 < g.n; i0 += stride) {
    const int64_t i = i0 + lane0;
    const bool act = i < g.n;
    const uint32_t b = act ? g.rowptr[i] : 0u, e = act ? g.rowptr[i + 1] : 0u;
    const bool big = e - b >= (uint32_t)kWave;
    unsigned long long bmask = __ballot(big);
    while (bmask) {
      const int l = __ffsll((long long)bmask) - 1;
      bmask &= bmask - 1;
      const int64_t v = __shfl(i, l);
      const uint32_t bb = __shfl(b, l), be = __shfl(e, l);
      for (int k = 0; k < c; k++) {
        const double yv = Y[v + (int64_t)k * g.n];
        for (uint32_t x = bb + lane0; x < be; x += kWave) {
          const double d = yv - Y[g.adj[x] + (int64_t)k * g.n];
          num += 0.5 * (g.w ? (double)g.w[x] : 1.0) * d * d;
        }
      }
    }
    if (!act) continue;
    for (int k = 0; k < c; k++) {
      const double yi = Y[i + (int64_t)k * g.n];
      if (!big)
        for (uint32_t x = b; x < e; x++) {
          const double d = yi - Y[g.adj[x] + (int64_t)k * g.n];
          num += 0.5 * (g.w ? (double)g.w[x] : 1.0) * d * d;
        }
      den += g.degw[i] * yi * yi;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    num += __shfl_xor(num, o);
    den += __shfl_xor(den, o);
  }
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  if (lane == 0) {
    sh[w][0] = num;
    sh[w][1] = den;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int q = 0; q < kBlk / kWave; q++) {
      a += sh[q][0];
      b += sh[q][1];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// ---------------------------------------------------------------- host side

static int blocks_for(phde_ctx *c, int64_t work, int per_cu = 8) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(work, (int64_t)c->num_cus * per_cu));
}

// Rows longer than kHubRow and their chunks (built once per context, at
// creation, from a host copy of the row pointers).
int spmm_hubs(phde_ctx *c) {
  if (c->sp_nhubs >= 0) return PHDE_OK;
  const int64_t n = c->g.n;
  std::vector<uint32_t> rp(n + 1);
  PHDE_HIP(hipMemcpyAsync(rp.data(), c->g.rowptr, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  std::vector<uint32_t> hr, c0, cr, cb;
  for (int64_t i = 0; i < n; i++) {
    if (rp[i + 1] - rp[i] <= kHubRow) continue;
    hr.push_back((uint32_t)i);
    c0.push_back((uint32_t)cr.size());
    for (uint32_t b = rp[i]; b < rp[i + 1]; b += kHubRow) {
      cr.push_back((uint32_t)i);
      cb.push_back(b);
    }
  }
  c0.push_back((uint32_t)cr.size());
  if (!hr.empty()) {
    struct {
      uint32_t **dst;
      const std::vector<uint32_t> *src;
    } up[] = {{&c->sp_hub_row, &hr}, {&c->sp_hub_c0, &c0}, {&c->sp_chunk_row, &cr}, {&c->sp_chunk_beg, &cb}};
    for (auto &u : up) {
      PHDE_HIP(hipMalloc(u.dst, sizeof(uint32_t) * u.src->size()));
      PHDE_HIP(hipMemcpy(*u.dst, u.src->data(), sizeof(uint32_t) * u.src->size(), hipMemcpyHostToDevice));
    }
  }
  c->sp_nhubs = (int64_t)hr.size();
  c->sp_nchunks = (int64_t)cr.size();
  return PHDE_OK;
}

template <typename T>
static int laplacian_t(phde_ctx *c, const void *Scol, void *Srow, void *LSrow, int s, double *SLS,
                       double *part, int nblk_gram, double *hub_part) {
  const int64_t n = c->g.n;
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_OTHER, &ev);
  col_to_row_kernel<T><<<blocks_for(c, (n + 63) / 64), kBlk, 0, c->stream>>>((const T *)Scol,
                                                                           (T *)Srow, n, s);
  c->timer.end(c->stream, PHDE_KERNEL_OTHER, ev);
  c->timer.begin(c->stream, PHDE_KERNEL_SPMM, &ev);
  const int sb = blocks_for(c, (n + 3) / 4, 16);
  if (c->g.w)
    laplacian_spmm_kernel<T, true><<<sb, kBlk, 0, c->stream>>>(c->g, (const T *)Srow, (T *)LSrow, s);
  else
    laplacian_spmm_kernel<T, false><<<sb, kBlk, 0, c->stream>>>(c->g, (const T *)Srow, (T *)LSrow, s);
  if (c->sp_nhubs > 0) {
    const int hb = blocks_for(c, (c->sp_nchunks + 3) / 4, 16);
    if (c->g.w)
      hub_partial_kernel<T, true><<<hb, kBlk, 0, c->stream>>>(c->g, (const T *)Srow, s, c->sp_chunk_row,
                                                               c->sp_chunk_beg, c->sp_nchunks, hub_part);
    else
      hub_partial_kernel<T, false><<<hb, kBlk, 0, c->stream>>>(c->g, (const T *)Srow, s, c->sp_chunk_row,
                                                                c->sp_chunk_beg, c->sp_nchunks, hub_part);
    hub_finish_kernel<T><<<blocks_for(c, (c->sp_nhubs + 3) / 4, 16), kBlk, 0, c->stream>>>(
        c->g, (const T *)Srow, (T *)LSrow, s, c->sp_hub_row, c->sp_hub_c0, c->sp_nhubs, hub_part);
  }
  c->timer.end(c->stream, PHDE_KERNEL_SPMM, ev);
  c->timer.begin(c->stream, PHDE_KERNEL_GRAM, &ev);
  PHDE_TRY_RC(gram_launch<T>(c, (const T *)Srow, (const T *)LSrow, n, s, part, nblk_gram, SLS));
  c->timer.end(c->stream, PHDE_KERNEL_GRAM, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int Lap::setup(phde_ctx *c, int64_t n, int s_, bool fp32_) {
  ctx = c;
  s = s_;
  fp32 = fp32_;
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  nblk_gram = blocks_for(c, (n + kGramRows - 1) / kGramRows, 2);
  PHDE_HIP(hipMalloc(&Srow, esz * (size_t)n * s));
  PHDE_HIP(hipMalloc(&LSrow, esz * (size_t)n * s));
  const int nparts = std::max(nblk_gram, gram_mfma_blocks(c, n) * (kBlk / kWave));
  PHDE_HIP(hipMalloc(&part, sizeof(double) * (size_t)std::max(nparts, 1) * s * s));
  PHDE_HIP(hipMalloc(&SLS, sizeof(double) * s * s));
  PHDE_TRY_RC(spmm_hubs(c));
  if (c->sp_nchunks > 0) PHDE_HIP(hipMalloc(&hub_part, sizeof(double) * (size_t)c->sp_nchunks * s));
  return PHDE_OK;
}

void Lap::release() {
  if (Srow) (void)hipFree(Srow);
  if (LSrow) (void)hipFree(LSrow);
  if (part) (void)hipFree(part);
  if (SLS) (void)hipFree(SLS);
  if (hub_part) (void)hipFree(hub_part);
  Srow = LSrow = nullptr;
  part = SLS = hub_part = nullptr;
}

int Lap::run(const void *Scol) {
  return fp32 ? laplacian_t<float>(ctx, Scol, Srow, LSrow, s, SLS, part, nblk_gram, hub_part)
              : laplacian_t<double>(ctx, Scol, Srow, LSrow, s, SLS, part, nblk_gram, hub_part);
}

// C = A^T A for a column-major n x s matrix (PivotMDS / PHDE C^T C,
// PivotMDS.cpp:149-152) using the same row-tiled Gram.
int Lap::gram_self(const void *Acol) {
  const int64_t n = ctx->g.n;
  hipEvent_t ev;
  ctx->timer.begin(ctx->stream, PHDE_KERNEL_GRAM, &ev);
  if (fp32) {
    col_to_row_kernel<float><<<blocks_for(ctx, (n + 63) / 64), kBlk, 0, ctx->stream>>>(
        (const float *)Acol, (float *)Srow, n, s);
    PHDE_TRY_RC(gram_launch<float>(ctx, (const float *)Srow, (const float *)Srow, n, s, part, nblk_gram, SLS));
  } else {
    col_to_row_kernel<double><<<blocks_for(ctx, (n + 63) / 64), kBlk, 0, ctx->stream>>>(
        (const double *)Acol, (double *)Srow, n, s);
    PHDE_TRY_RC(gram_launch<double>(ctx, (const double *)Srow, (const double *)Srow, n, s, part, nblk_gram, SLS));
  }
  ctx->timer.end(ctx->stream, PHDE_KERNEL_GRAM, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int coords_compute(phde_ctx *c, const void *Scol, bool fp32, int s, int nc, const double *V_dev,
                   double *X_dev, double *part, double *inv) {
  const int64_t n = c->g.n;
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_COORDS, &ev);
  const int nb = blocks_for(c, (n + kBlk - 1) / kBlk, 4);
  if (fp32)
    coords_kernel<float><<<nb, kBlk, 0, c->stream>>>((const float *)Scol, V_dev, n, s, nc, X_dev, part);
  else
    coords_kernel<double><<<nb, kBlk, 0, c->stream>>>((const double *)Scol, V_dev, n, s, nc, X_dev,
                                                      part);
  coords_norm_kernel<<<1, 8 * kWave, 0, c->stream>>>(part, nb, nc, inv);
  scale_cols_kernel<<<blocks_for(c, (n * nc + kBlk - 1) / kBlk, 4), kBlk, 0, c->stream>>>(X_dev, n,
                                                                                        nc, inv);
  c->timer.end(c->stream, PHDE_KERNEL_COORDS, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

// X[:, k] *= 1 / ||X[:, k]||_2 from per-block partial sums of squares
// (part[k * nblk + b]), as coords_compute finishes.
int coords_normalise(phde_ctx *c, double *X_dev, int nc, const double *part, int nblk, double *inv) {
  const int64_t n = c->g.n;
  coords_norm_kernel<<<1, 8 * kWave, 0, c->stream>>>(part, nblk, nc, inv);
  scale_cols_kernel<<<blocks_for(c, (n * nc + kBlk - 1) / kBlk, 4), kBlk, 0, c->stream>>>(X_dev, n, nc, inv);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

// S^T D S for the generalised eigenproblem: the row-tiled Gram with the
// rows of the left operand scaled by deg.
int weighted_gram(phde_ctx *c, const void *S_dev, bool fp32, int s, double *out_dev) {
  Lap tmp;
  int rc = tmp.setup(c, c->g.n, s, fp32);
  if (rc) return rc;
  const int64_t n = c->g.n;
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_GRAM, &ev);
  if (fp32) {
    col_to_row_kernel<float><<<blocks_for(c, (n + 63) / 64), kBlk, 0, c->stream>>>(
        (const float *)S_dev, (float *)tmp.Srow, n, s);
    PHDE_TRY_RC(gram_launch<float>(c, (const float *)tmp.Srow, (const float *)tmp.Srow, n, s, tmp.part,
                                   tmp.nblk_gram, out_dev, c->g.degw));
  } else {
    col_to_row_kernel<double><<<blocks_for(c, (n + 63) / 64), kBlk, 0, c->stream>>>(
        (const double *)S_dev, (double *)tmp.Srow, n, s);
    PHDE_TRY_RC(gram_launch<double>(c, (const double *)tmp.Srow, (const double *)tmp.Srow, n, s, tmp.part,
                                    tmp.nblk_gram, out_dev, c->g.degw));
  }
  c->timer.end(c->stream, PHDE_KERNEL_GRAM, ev);
  PHDE_HIP(hipGetLastError());
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

int hall_energy_dev(phde_ctx *c, const double *Y_dev, int nc, double *part_dev, double *h_out2) {
  const int nb = blocks_for(c, (c->g.n + kBlk - 1) / kBlk, 4);
  hall_kernel<<<nb, kBlk, 0, c->stream>>>(c->g, Y_dev, nc, part_dev);
  PHDE_HIP(hipGetLastError());
  std::vector<double> h(2 * (size_t)nb);
  PHDE_HIP(hipMemcpyAsync(h.data(), part_dev, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost,
                          c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  double a = 0, b = 0;
  for (int i = 0; i < nb; i++) {
    a += h[2 * i];
    b += h[2 * i + 1];
  }
  h_out2[0] = a;
  h_out2[1] = b;
  return PHDE_OK;
}

}  // namespace phde
