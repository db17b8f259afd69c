// eig.hip — s x s eigensolvers in one gfx950 workgroup.  Compiled with
// -ffp-contract=off: every product and sum is rounded on its own, as the
// reference's x86-64 build does, so the power method below reproduces the
// reference's FP64 results bit for bit from the same input.
//
// power_eig_kernel restates eigenValuePowerMethod
// (variants/ParallelComponents.cpp:48-158): subspace iteration E = K T,
// classical Gram-Schmidt deflation of each column against the earlier
// (un-normalised) ones, column normalisation, stop when
// min_i |<e_i, t_i>| >= 1 - 1e-3, then eigenvalue reversal and the
// reference's sign rule.  Dot products are summed sequentially in index
// order (the one-thread reduction order).  The column deflations are the
// serial part; they run as a wavefront: wave w owns columns w, w+16, ...,
// and column i waits (LDS flag) only for the columns j < i it subtracts.
//
// sym_eig_kernel is the exact solve: cyclic Jacobi with the round-robin
// parallel ordering on the symmetric part of A, optionally after a
// Cholesky reduction B = L L^T of a generalised problem.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "phde_common.h"
#include "eig.h"

namespace phde {

constexpr int kEigThreads = 1024;
constexpr int kEigWaves = kEigThreads / kWave;

// Power method: every sum of the reference runs in one lane, in index
// order, each product and sum rounded on its own.  Lane i holds column i of
// the iterate in registers, padded with zeros to P rows (P = dim rounded up
// to 8, at least 16): a padded term adds +0.0 to a sum that started at +0.0
// -- such a sum is never -0.0 -- so every sum equals the reference's over
// dim terms, and the loops have no bounds to test.  The product E = K T is
// split over the block's waves by rows and exchanged through LDS ([k][i]:
// row k of every column contiguous); the deflation, normalisation and
// convergence test then run in wave 0 alone, column j of the deflation
// broadcast to its lanes through an LDS row, and the new iterate goes back
// to the other waves through LDS.  Measured at dim 50 (29 iterations):
// 8.8 -> 3.2 ms with the columns in lanes; then 2.68 -> 2.33 ms when the
// deflation lost three redundant copies (each step's broadcast row was read
// by every wave, the CU's LDS bandwidth shared four ways) and the padding
// went from 64 to 56 rows.  What is left is one wave's instruction stream:
// per step, two passes of P multiplies and P adds or subtracts.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kPowerThreads = 256;
constexpr int kPowerWaves = kPowerThreads / kWave;

// LDS of power_eig_kernel<P>: K padded ([k + j*P]), the iterate and the
// previous iterate ([k*P + i]), wave 0's broadcast row, the convergence
// measure.
constexpr size_t power_lds_bytes(int P) { return sizeof(double) * (3 * (size_t)P * P + (size_t)P + 2); }

template <int P>
__global__ __launch_bounds__(kPowerThreads) void power_eig_kernel(double *K, const double *init, double *evals,
                                                                  int dim, int max_iter, int *status) {
  static_assert(P % 8 == 0 && P >= 16 && P <= 64, "rows: a multiple of 8, at most one per lane");
  extern __shared__ __attribute__((aligned(16))) double lds[];
  double *Kp = lds;
  double *E = lds + P * P;
  double *T = lds + 2 * P * P;
  double *row = lds + 3 * P * P;
  double *shr = row + P;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  const bool col = lane < dim;
  for (int e = tid; e < P * P; e += kPowerThreads) {
    const int a = e % P, b = e / P;
    Kp[e] = a < dim && b < dim ? K[a + b * dim] : 0.0;  // Kp[k + j*P] = K[k + j*dim]
    E[e] = a < dim && b < dim ? init[b + a * dim] : 0.0;  // E[k*P + i] = init[k + i*dim]
  }
  __syncthreads();
  double x[P];
#pragma unroll
  for (int k = 0; k < P; k++) x[k] = lane < P ? E[k * P + lane] : 0.0;
  // ParallelComponents.cpp:67-72
  double ev;
  {
    double sq = 0;
#pragma unroll
    for (int k = 0; k < P; k++) {
      const double p = x[k] * x[k];
      sq = sq + p;
    }
    ev = col ? sqrt(sq) : 1.0;
#pragma unroll
    for (int k = 0; k < P; k++) x[k] = x[k] / ev;
  }
  const double EPSILON = 1 - 1e-3;
  double r = 0;
  int iters = 0, st = 0;
  __syncthreads();  // E is rewritten below
  while (r < EPSILON) {
    iters++;
    if (isnan(r) || isinf(r)) {
      st = 1;
      break;
    }
    if (iters > max_iter) {
      st = 2;
      break;
    }
    // tmpOld = eVecs (:83-90)
    if (wid == 0 && lane < P) {
#pragma unroll
      for (int k = 0; k < P; k++) T[k * P + lane] = x[k];
    }
    // eVecs[k + i*p] = sum_j K[k + j*p] * tmpOld[j + i*p], j ascending
    // (:91-100): wave w takes rows [w*P/W, (w+1)*P/W), two at a time
    constexpr int kRows = P / kPowerWaves;
#pragma unroll
    for (int g = 0; g < kRows; g += 2) {
      const int k0 = wid * kRows + g;
      double y0 = 0, y1 = 0;
#pragma unroll
      for (int j = 0; j < P; j++) {
        const double2 ka = *reinterpret_cast<const double2 *>(Kp + k0 + j * P);
        const double xj = x[j];
        const double p0 = ka.x * xj, p1 = ka.y * xj;
        y0 = y0 + p0;
        y1 = y1 + p1;
      }
      if (lane < P) {
        E[(k0 + 0) * P + lane] = y0;
        E[(k0 + 1) * P + lane] = y1;
      }
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
      for (int k = 0; k < P; k++) x[k] = lane < P ? E[k * P + lane] : 0.0;
      // deflation (:102-112): for each j, lane t >= j sums E_j . E_t in
      // index order (t == j: the squared norm of the final column j, the
      // denominator, recomputed by the reference for every i); lanes t > j
      // subtract fac E_j, fac = sum / denominator
      double dj_own = 1;
      for (int j = 0; j < dim; j++) {
        if (lane == j) {
#pragma unroll
          for (int k = 0; k < P; k += 2) *reinterpret_cast<double2 *>(row + k) = make_double2(x[k], x[k + 1]);
        }
        wave_lds_sync();
        double sum = 0;
#pragma unroll
        for (int k = 0; k < P; k += 2) {
          const double2 e = *reinterpret_cast<const double2 *>(row + k);
          const double p0 = e.x * x[k];
          sum = sum + p0;
          const double p1 = e.y * x[k + 1];
          sum = sum + p1;
        }
        const double dj = __shfl(sum, j);
        if (lane == j) dj_own = sum;
        if (lane > j) {
          const double fac = sum / dj;
#pragma unroll
          for (int k = 0; k < P; k += 2) {
            const double2 e = *reinterpret_cast<const double2 *>(row + k);
            const double q0 = fac * e.x, q1 = fac * e.y;
            x[k] = x[k] - q0;
            x[k + 1] = x[k + 1] - q1;
          }
        }
        wave_lds_sync();
      }
      // normalisation (:113-120): normL2 sums the same squares in the same
      // order as the denominator above
      ev = col ? sqrt(dj_own) : 1.0;
#pragma unroll
      for (int k = 0; k < P; k++) x[k] = x[k] / ev;
      // convergence test (:122-130): r = min_i |<e_i, t_i>|
      double c = 0;
#pragma unroll
      for (int k = 0; k < P; k++) {
        const double p = x[k] * (lane < P ? T[k * P + lane] : 0.0);
        c = c + p;
      }
      if (c < 0) c *= -1;
      double rr = col && c < 1 ? c : 1.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double v = __shfl_xor(rr, o);
        rr = v < rr ? v : rr;
      }
      r = rr;
      // the new iterate and r to the other waves
      if (lane < P) {
#pragma unroll
        for (int k = 0; k < P; k++) E[k * P + lane] = x[k];
      }
      if (lane == 0) shr[0] = rr;
    }
    __syncthreads();
    if (wid != 0) {
#pragma unroll
      for (int k = 0; k < P; k++) x[k] = lane < P ? E[k * P + lane] : 0.0;
      r = shr[0];
    }
    __syncthreads();  // T and E are rewritten at the next iteration's start
  }
  if (st == 0 && wid == 0) {
    // eigenvalue reversal (:132-140); column i goes to column dim-1-i,
    // negated when the first entry of the iterate's column dim-1-i is
    // negative (:142-154)
    const int dst = dim - 1 - lane;
    const double first = __shfl(x[0], col ? dst : 0);
    if (col) {
      evals[dst] = ev;
      const bool flip = first < 0;
#pragma unroll
      for (int k = 0; k < P; k++)
        if (k < dim) K[k + dst * dim] = flip ? -x[k] : x[k];
    }
  }
  if (tid == 0) {
    status[0] = st;
    status[1] = iters;
  }
}

// ------------------------------------------------------------ exact solves

__device__ __forceinline__ void rr_pair(int p, int r, int m, int *a, int *b) {
  // circle method: position 0 fixed, positions 1..m-1 rotate by r
  auto player = [&](int pos) { return pos == 0 ? 0 : ((pos - 1 + r) % (m - 1)) + 1; };
  *a = player(p);
  *b = player(m - 1 - p);
}

// A (dim x dim col-major), B optional SPD.  W: global work of 3*dim*dim.
// Output: evals ascending, evecs column-major (B-orthonormal when B given),
// each column's largest-magnitude entry positive.
__global__ __launch_bounds__(kEigThreads) void sym_eig_kernel(const double *A, const double *B,
                                                             int dim, double *evals,
                                                             double *evecs, double *W,
                                                             int *status) {
  extern __shared__ double lds[];
  double *C = lds;                 // dim*dim, row index fast: C[i + j*dim]
  double *cs = lds + dim * dim;    // 64 cos
  double *sn = cs + 64;            // 64 sin
  double *red = sn + 64;           // 64 reduction slots + 2
  double *V = W;                   // dim*dim
  double *L = W + dim * dim;       // dim*dim (Cholesky factor, lower)
  double *M = W + 2 * dim * dim;   // dim*dim scratch
  __shared__ int perm[64];
  __shared__ int fail;
  const int tid = threadIdx.x;
  const int dd = dim * dim;
  if (tid == 0) fail = 0;
  for (int e = tid; e < dd; e += kEigThreads) {
    const int i = e % dim, j = e / dim;
    C[e] = 0.5 * (A[i + j * dim] + A[j + i * dim]);
  }
  __syncthreads();
  if (B) {
    // Cholesky of the symmetric part of B (right-looking, column k at a time)
    for (int e = tid; e < dd; e += kEigThreads) {
      const int i = e % dim, j = e / dim;
      L[e] = 0.5 * (B[i + j * dim] + B[j + i * dim]);
    }
    __syncthreads();
    for (int k = 0; k < dim; k++) {
      if (tid == 0) {
        const double d = L[k + k * dim];
        if (!(d > 0)) fail = 1;
        L[k + k * dim] = sqrt(d > 0 ? d : 1.0);
      }
      __syncthreads();
      const double lkk = L[k + k * dim];
      for (int i = k + 1 + tid; i < dim; i += kEigThreads) L[i + k * dim] /= lkk;
      __syncthreads();
      for (int e = tid; e < dd; e += kEigThreads) {
        const int i = e % dim, j = e / dim;
        if (j > k && i >= j) L[i + j * dim] -= L[i + k * dim] * L[j + k * dim];
      }
      __syncthreads();
    }
    for (int e = tid; e < dd; e += kEigThreads) {
      const int i = e % dim, j = e / dim;
      if (i < j) L[e] = 0;
    }
    __syncthreads();
    // M = L^{-1} C (forward substitution, one column per thread)
    for (int j = tid; j < dim; j += kEigThreads)
      for (int i = 0; i < dim; i++) {
        double v = C[i + j * dim];
        for (int p = 0; p < i; p++) v -= L[i + p * dim] * M[p + j * dim];
        M[i + j * dim] = v / L[i + i * dim];
      }
    __syncthreads();
    // C = L^{-1} M^T
    for (int j = tid; j < dim; j += kEigThreads)
      for (int i = 0; i < dim; i++) {
        double v = M[j + i * dim];
        for (int p = 0; p < i; p++) v -= L[i + p * dim] * C[p + j * dim];
        C[i + j * dim] = v / L[i + i * dim];
      }
    __syncthreads();
    for (int e = tid; e < dd; e += kEigThreads) {
      const int i = e % dim, j = e / dim;
      if (i < j) {
        const double v = 0.5 * (C[i + j * dim] + C[j + i * dim]);
        C[i + j * dim] = v;
        C[j + i * dim] = v;
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < dd; e += kEigThreads) V[e] = (e % dim == e / dim) ? 1.0 : 0.0;
  __syncthreads();
  const int m = dim + (dim & 1);
  const int npairs = m / 2;
  for (int sweep = 0; sweep < 60; sweep++) {
    // off-diagonal mass
    double off = 0, diag = 0;
    for (int e = tid; e < dd; e += kEigThreads) {
      const int i = e % dim, j = e / dim;
      if (i != j)
        off += C[e] * C[e];
      else
        diag += C[e] * C[e];
    }
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o);
      diag += __shfl_xor(diag, o);
    }
    if (tid % kWave == 0) {
      red[tid / kWave] = off;
      red[16 + tid / kWave] = diag;
    }
    __syncthreads();
    if (tid == 0) {
      double a = 0, b = 0;
      for (int w = 0; w < kEigWaves; w++) {
        a += red[w];
        b += red[16 + w];
      }
      red[62] = a;
      red[63] = b;
    }
    __syncthreads();
    const double offs = red[62], diags = red[63];
    __syncthreads();
    if (offs <= 1e-30 * diags || offs == 0) break;
    for (int r = 0; r < m - 1; r++) {
      if (tid < npairs) {
        int p, q;
        rr_pair(tid, r, m, &p, &q);
        double c = 1, s = 0;
        if (p < dim && q < dim) {
          const double apq = C[p + q * dim];
          if (apq != 0) {
            const double app = C[p + p * dim], aqq = C[q + q * dim];
            const double tau = (aqq - app) / (2 * apq);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
            c = 1 / sqrt(1 + t * t);
            s = t * c;
          }
        }
        cs[tid] = c;
        sn[tid] = s;
      }
      __syncthreads();
      // rows: A <- J^T A
      for (int e = tid; e < npairs * dim; e += kEigThreads) {
        const int pi = e / dim, k = e % dim;
        int p, q;
        rr_pair(pi, r, m, &p, &q);
        if (p >= dim || q >= dim) continue;
        const double c = cs[pi], s = sn[pi];
        const double apk = C[p + k * dim], aqk = C[q + k * dim];
        C[p + k * dim] = c * apk - s * aqk;
        C[q + k * dim] = s * apk + c * aqk;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int e = tid; e < npairs * dim; e += kEigThreads) {
        const int pi = e / dim, k = e % dim;
        int p, q;
        rr_pair(pi, r, m, &p, &q);
        if (p >= dim || q >= dim) continue;
        const double c = cs[pi], s = sn[pi];
        const double akp = C[k + p * dim], akq = C[k + q * dim];
        C[k + p * dim] = c * akp - s * akq;
        C[k + q * dim] = s * akp + c * akq;
        const double vkp = V[k + p * dim], vkq = V[k + q * dim];
        V[k + p * dim] = c * vkp - s * vkq;
        V[k + q * dim] = s * vkp + c * vkq;
      }
      __syncthreads();
    }
  }
  // back-transform for the generalised problem: X = L^{-T} Y
  if (B) {
    for (int j = tid; j < dim; j += kEigThreads)
      for (int i = dim - 1; i >= 0; i--) {
        double v = V[i + j * dim];
        for (int p = i + 1; p < dim; p++) v -= L[p + i * dim] * M[p + j * dim];
        M[i + j * dim] = v / L[i + i * dim];
      }
    __syncthreads();
    for (int e = tid; e < dd; e += kEigThreads) V[e] = M[e];
    __syncthreads();
  }
  if (tid == 0) {
    for (int i = 0; i < dim; i++) perm[i] = i;
    for (int i = 0; i < dim; i++) {
      int best = i;
      for (int j = i + 1; j < dim; j++)
        if (C[perm[j] + perm[j] * dim] < C[perm[best] + perm[best] * dim]) best = j;
      const int t = perm[i];
      perm[i] = perm[best];
      perm[best] = t;
    }
  }
  __syncthreads();
  for (int i = tid; i < dim; i += kEigThreads) {
    const int src = perm[i];
    evals[i] = C[src + src * dim];
    int arg = 0;
    double big = -1;
    for (int k = 0; k < dim; k++) {
      const double a = fabs(V[k + src * dim]);
      if (a > big) {
        big = a;
        arg = k;
      }
    }
    const double sg = V[arg + src * dim] < 0 ? -1.0 : 1.0;
    for (int k = 0; k < dim; k++) evecs[k + i * dim] = sg * V[k + src * dim];
  }
  if (tid == 0) status[0] = fail;
}

// ---------------------------------------------------------------- host side

int power_eig_dev(phde_ctx *c, double *K_dev, const double *init_dev, double *evals_dev, int dim,
                  int max_iter, int *status_dev) {
  if (dim < 1 || dim > 64) {
    set_error("power method: dim " + std::to_string(dim) + " outside 1..64");
    return PHDE_ERR_ARG;
  }
  // rows padded to a multiple of 8, at least 16
  const int P = std::max(16, (dim + 7) / 8 * 8);
  void (*fn)(double *, const double *, double *, int, int, int *) =
      P == 16 ? power_eig_kernel<16> : P == 24 ? power_eig_kernel<24> : P == 32 ? power_eig_kernel<32>
      : P == 40 ? power_eig_kernel<40> : P == 48 ? power_eig_kernel<48> : P == 56 ? power_eig_kernel<56>
                                                                                    : power_eig_kernel<64>;
  const size_t shmem = power_lds_bytes(P);
  if (shmem > 65536)
    PHDE_HIP(hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_EIG, &ev);
  fn<<<1, kPowerThreads, shmem, c->stream>>>(K_dev, init_dev, evals_dev, dim, max_iter, status_dev);
  c->timer.end(c->stream, PHDE_KERNEL_EIG, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

int sym_eig_dev(phde_ctx *c, const double *A_dev, const double *B_dev, int dim, double *evals_dev,
                double *evecs_dev, double *work_dev, int *status_dev) {
  const size_t shmem = sizeof(double) * (dim * dim + 3 * 64);
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_EIG, &ev);
  sym_eig_kernel<<<1, kEigThreads, shmem, c->stream>>>(A_dev, B_dev, dim, evals_dev, evecs_dev,
                                                      work_dev, status_dev);
  c->timer.end(c->stream, PHDE_KERNEL_EIG, ev);
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

// Start vectors of the power method: srand(1) then rand()/RAND_MAX, drawn in
// column-major order (ParallelComponents.cpp:57-64).  The C library's
// rand() is restated from its published additive-feedback definition
// (glibc TYPE_3: r[i] = 16807 r[i-1] mod (2^31-1) for i < 31, r[i] = r[i-31]
// for 31 <= i < 34, r[i] = r[i-3] + r[i-31] mod 2^32 afterwards; the k-th
// output is r[k+344] >> 1), so no global generator state is touched.
void glibc_rand_sequence(unsigned seed, int count, int32_t *out) {
  std::vector<uint32_t> r(344 + count);
  int64_t word = seed == 0 ? 1 : (int64_t)seed;
  r[0] = (uint32_t)word;
  for (int i = 1; i < 31; i++) {
    const int64_t hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r[i] = (uint32_t)word;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 34; i < 344 + count; i++) r[i] = r[i - 31] + r[i - 3];
  for (int k = 0; k < count; k++) out[k] = (int32_t)(r[k + 344] >> 1);
}

// ---------------------------------------------------------------- no allocation
// phde_power_eig_noalloc runs the power method above without allocating:
// its scratch — the dim*dim start values (the power method keeps its
// matrices in registers and LDS) — is this file-scope device array
// (32 KiB), which every device the library runs on holds until the process
// ends.  Host code never passes `scratch` itself
// (in host code the name is a host-side stand-in, not a device address):
// each device's address comes from hipGetSymbolAddress, once per device.
constexpr int kNoAllocMaxDim = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocMaxDim * kNoAllocMaxDim];

static std::mutex g_scratch_mu;
static double *g_scratch_dev[kMaxDevices] = {};

static int scratch_address(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("phde_power_eig_noalloc: device index out of range");
    return PHDE_ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    int cur = 0;
    PHDE_HIP(hipGetDevice(&cur));
    if (cur != device) PHDE_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) PHDE_HIP(hipSetDevice(cur));
    if (e != hipSuccess || !p) {
      set_error(std::string("phde_power_eig_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return PHDE_ERR_DEVICE;
    }
    g_scratch_dev[device] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[device];
  return PHDE_OK;
}

// The power method's start values on the device: srand(1), then
// rand() / RAND_MAX for `count` draws (the sequence glibc_rand_sequence
// gives the host), kept in a 34-entry ring: r[i] needs r[i-31] and r[i-3].
__global__ void power_start_kernel(double *init, int count) {
  if (threadIdx.x != 0) return;
  uint32_t ring[34];
  int64_t word = 1;
  ring[0] = 1u;
  for (int i = 1; i < 31; i++) {
    const int64_t hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    ring[i] = (uint32_t)word;
  }
  for (int i = 31; i < 34; i++) ring[i] = ring[i - 31];
  for (int i = 34; i < 344 + count; i++) {
    const uint32_t v = ring[(i - 31) % 34] + ring[(i - 3) % 34];
    ring[i % 34] = v;
    if (i >= 344) init[i - 344] = (double)(int32_t)(v >> 1) / 2147483647;
  }
}

}  // namespace phde

using namespace phde;

extern "C" int phde_power_eig_noalloc(phde_ctx *c, double *K_dev, double *evals_dev, int dim,
                                      int32_t *status_dev) {
  if (!c || !K_dev || !evals_dev || !status_dev) {
    set_error("phde_power_eig_noalloc: null argument");
    return PHDE_ERR_ARG;
  }
  if (dim < 1 || dim > kNoAllocMaxDim) {
    set_error("phde_power_eig_noalloc: dim " + std::to_string(dim) + " outside 1.." +
              std::to_string(kNoAllocMaxDim) + " (the size of its static scratch)");
    return PHDE_ERR_ARG;
  }
  double *scr = nullptr;
  PHDE_TRY_RC(scratch_address(c->device, &scr));
  double *init = scr;
  power_start_kernel<<<1, 64, 0, c->stream>>>(init, dim * dim);
  PHDE_HIP(hipGetLastError());
  PHDE_TRY_RC(power_eig_dev(c, K_dev, init, evals_dev, dim, 100000, reinterpret_cast<int *>(status_dev)));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}
