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

// Sum of p over lanes 0..dim-1 in lane order, starting from 0 (the
// one-thread order of the reference).  The wave's values go through its own
// LDS row (65 doubles): lane 0 reads them 16 at a time, all 16 loads in
// flight together, adds them in order and publishes the result.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double seq_sum(double p, int dim, double *wbuf) {
  const int lane = __lane_id();
  wbuf[lane] = p;
  wave_lds_sync();
  if (lane == 0) {
    double s = 0;
#pragma unroll
    for (int c = 0; c < 64; c += 16) {
      if (c >= dim) break;
      double v[16];
#pragma unroll
      for (int k = 0; k < 16; k += 2) {
        const double2 t = *reinterpret_cast<const double2 *>(wbuf + c + k);
        v[k] = t.x;
        v[k + 1] = t.y;
      }
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (c + k < dim) s += v[k];
    }
    wbuf[64] = s;
  }
  wave_lds_sync();
  return wbuf[64];
}

constexpr int kSumRow = 72;  // doubles per wave in the summation area

__global__ __launch_bounds__(kEigThreads) void power_eig_kernel(double *K, const double *init,
                                                               double *evals, int dim,
                                                               int max_iter, double *Tglob,
                                                               int *status) {
  extern __shared__ __attribute__((aligned(16))) double lds[];
  const int dd = dim * dim;
  // LDS: E, T (or T in global memory when dim is large), a copy of K, the
  // per-wave summation rows (16-byte aligned), eigenvalue / denominator /
  // convergence slots
  double *E = lds;
  double *T = Tglob ? Tglob : lds + dd;
  double *Kl = Tglob ? lds + dd : lds + 2 * dd;
  double *sums = lds + (((Tglob ? 2 : 3) * dd + 1) & ~1);
  double *small = sums + kEigWaves * kSumRow;
  double *ev = small;
  double *djj = small + 64;
  double *conv = small + 128;
  int *flags = (int *)(small + 192);
  __shared__ double r_sh;
  const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
  const bool on = lane < dim;
  double *wsum = sums + wid * kSumRow;

  for (int e = tid; e < dd; e += kEigThreads) {
    E[e] = init[e];
    Kl[e] = K[e];
  }
  __syncthreads();
  // ParallelComponents.cpp:67-72
  for (int i = wid; i < dim; i += kEigWaves) {
    const double x = on ? E[lane + i * dim] : 0.0;
    const double nrm = sqrt(seq_sum(x * x, dim, wsum));
    if (lane == 0) ev[i] = nrm;
    if (on) E[lane + i * dim] = x / nrm;
  }
  __syncthreads();

  const double EPSILON = 1 - 1e-3;
  double r = 0;
  int iters = 0, st = 0;
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
    // tmpOld = eVecs (ParallelComponents.cpp:83-90)
    for (int e = tid; e < dd; e += kEigThreads) T[e] = E[e];
    if (tid < 128) flags[tid] = 0;
    __syncthreads();
    // eVecs[k + i*p] += K[k + j*p] * tmpOld[j + i*p], j ascending (:91-100)
    for (int e = tid; e < dd; e += kEigThreads) {
      const int k = e % dim, i = e / dim;
      double acc = 0;
      for (int j = 0; j < dim; j++) acc += Kl[k + j * dim] * T[j + i * dim];
      E[e] = acc;
    }
    __syncthreads();
    // deflation (:102-112).  Column j is published in two steps: its
    // entries (flags[j]) as soon as they are final, then its squared norm
    // (flags[64 + j]), so that a later column's dot product with it runs
    // while its norm is still being summed; the arithmetic is unchanged.
    for (int i = wid; i < dim; i += kEigWaves) {
      double x = on ? E[lane + i * dim] : 0.0;
      for (int j = 0; j < i; j++) {
        while (__hip_atomic_load(&flags[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
          __builtin_amdgcn_s_sleep(1);
        const double ej = on ? E[lane + j * dim] : 0.0;
        const double dot = seq_sum(ej * x, dim, wsum);
        while (__hip_atomic_load(&flags[64 + j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
          __builtin_amdgcn_s_sleep(1);
        const double fac = dot / djj[j];
        x = x - fac * ej;
      }
      if (on) E[lane + i * dim] = x;
      wave_lds_sync();
      if (lane == 0) __hip_atomic_store(&flags[i], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      const double d = seq_sum(x * x, dim, wsum);
      if (lane == 0) {
        djj[i] = d;
        __hip_atomic_store(&flags[64 + i], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    __syncthreads();
    // normalisation (:113-120)
    for (int i = wid; i < dim; i += kEigWaves) {
      const double x = on ? E[lane + i * dim] : 0.0;
      const double nrm = sqrt(seq_sum(x * x, dim, wsum));
      if (lane == 0) ev[i] = nrm;
      if (on) E[lane + i * dim] = x / nrm;
    }
    __syncthreads();
    // convergence test (:122-130)
    for (int i = wid; i < dim; i += kEigWaves) {
      const double x = on ? E[lane + i * dim] : 0.0;
      const double y = on ? T[lane + i * dim] : 0.0;
      const double d = seq_sum(x * y, dim, wsum);
      if (lane == 0) conv[i] = d;
    }
    __syncthreads();
    if (tid == 0) {
      double rr = 1;
      for (int i = 0; i < dim; i++) {
        double tmp = conv[i];
        if (tmp < 0) tmp *= -1;
        if (tmp < rr) rr = tmp;
      }
      r_sh = rr;
    }
    __syncthreads();
    r = r_sh;
  }
  if (st == 0) {
    // eigenvalue reversal (:132-140) and sign rule (:142-154)
    for (int i = tid; i < dim; i += kEigThreads) evals[i] = ev[dim - 1 - i];
    for (int e = tid; e < dd; e += kEigThreads) {
      const int j = e % dim, i = e / dim;
      const bool flip = E[(dim - i - 1) * dim] < 0;
      const double v = E[j + i * dim];
      K[j + (dim - i - 1) * dim] = flip ? -v : v;
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
                  int max_iter, double *work_dev, int *status_dev) {
  int maxsh = 0;
  PHDE_HIP(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
  const size_t small = (193 + kEigWaves * kSumRow) * sizeof(double) + 128 * sizeof(int);
  size_t shmem = 3 * sizeof(double) * dim * dim + small;
  double *Tglob = nullptr;
  if (shmem > (size_t)maxsh) {
    shmem = 2 * sizeof(double) * dim * dim + small;
    Tglob = work_dev;
  }
  if (shmem > 65536)
    PHDE_HIP(hipFuncSetAttribute((const void *)power_eig_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_EIG, &ev);
  power_eig_kernel<<<1, kEigThreads, shmem, c->stream>>>(K_dev, init_dev, evals_dev, dim, max_iter,
                                                        Tglob, status_dev);
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
// its scratch — the dim*dim start values and, when the power method's
// matrices do not all fit in LDS, its global T (dim*dim) — is this
// file-scope device array (64 KiB), which every device the library runs on
// holds until the process ends.  Host code never passes `scratch` itself
// (in host code the name is a host-side stand-in, not a device address):
// each device's address comes from hipGetSymbolAddress, once per device.
constexpr int kNoAllocMaxDim = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[2 * kNoAllocMaxDim * kNoAllocMaxDim];

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
  double *init = scr, *work = scr + (size_t)dim * dim;
  power_start_kernel<<<1, 64, 0, c->stream>>>(init, dim * dim);
  PHDE_HIP(hipGetLastError());
  PHDE_TRY_RC(power_eig_dev(c, K_dev, init, evals_dev, dim, 100000, work, reinterpret_cast<int *>(status_dev)));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}
