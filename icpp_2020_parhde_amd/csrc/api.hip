// api.hip — the C ABI of libparhde.so (include/parhde.h) and the layout
// drivers that sequence the device stages.
//
// The drivers follow the control flow of the reference's run() methods:
// ParHDE::run (variants/ParHDE.cpp:45-352), WParHDE::run
// (variants/WParHDE.cpp:40-291), PivotMDS::run (variants/PivotMDS.cpp:68-147)
// and PHDE::run (variants/PHDE.cpp:50-136), followed by getCoordinates
// (variants/DrawingBase.cpp:196-205).  All n-sized work runs on the GPU; the
// host keeps the pivot bookkeeping (the `sources` set of the dedupe rule),
// the accept/discard decision per candidate and the start values of the
// power method.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "device_util.h"
#include "eig.h"
#include "learned.h"
#include "mds.h"
#include "orth.h"
#include "pbfs.h"
#include "phde_common.h"
#include "spmm.h"
#include "traverse.h"

namespace phde {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

hipEvent_t Timer::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
void Timer::begin(hipStream_t s, int cls, hipEvent_t *a) {
  *a = nullptr;
  launches[cls]++;
  if (!enabled) return;
  *a = get();
  (void)hipEventRecord(*a, s);
}
void Timer::end(hipStream_t s, int cls, hipEvent_t a) {
  if (!enabled || !a) return;
  hipEvent_t b = get();
  (void)hipEventRecord(b, s);
  open.push_back({cls, a, b});
}
int Timer::resolve() {
  for (auto &sp : open) {
    float ms = 0;
    if (hipEventSynchronize(sp.b) == hipSuccess && hipEventElapsedTime(&ms, sp.a, sp.b) == hipSuccess)
      this->ms[sp.cls] += ms;
    pool.push_back(sp.a);
    pool.push_back(sp.b);
  }
  open.clear();
  return PHDE_OK;
}
void Timer::reset() {
  resolve();
  for (double &m : ms) m = 0;
  for (long long &l : launches) l = 0;
}
Timer::~Timer() {
  for (auto &sp : open) {
    (void)hipEventDestroy(sp.a);
    (void)hipEventDestroy(sp.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}

__global__ void degree_kernel(const uint32_t *rowptr, const int32_t *w, int64_t n, int32_t *deg32,
                              double *degw) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t b = rowptr[i], e = rowptr[i + 1];
    deg32[i] = (int32_t)(e - b);
    if (w) {
      // WParHDE.cpp:53-58: weights summed in adjacency order
      double d = 0;
      for (uint32_t k = b; k < e; k++) d += (double)w[k];
      degw[i] = d;
    } else {
      degw[i] = (double)(e - b);  // ParHDE.cpp:58-59
    }
  }
}

__global__ void gather_cols_kernel(const double *K, int dim, int c, bool largest, double *out) {
  // eigenvector selection of DrawingBase.cpp:70-72 (Largest) and
  // :100-102 (First)
  for (int e = threadIdx.x; e < dim * c; e += blockDim.x) {
    const int i = e % dim, j = e / dim;
    out[e] = largest ? K[i + (dim - j - 1) * dim] : K[i + j * dim];
  }
}

// Default delta-stepping bucket width: a few times the largest weight, so a
// bucket spans several hops (tunable with PHDE_SSSP_DELTA).
static int32_t pick_delta(const int32_t *w, int64_t m) {
  if (const char *e = getenv("PHDE_SSSP_DELTA")) return std::max(1, atoi(e));
  int32_t mx = 1;
  for (int64_t i = 0; i < m; i++) mx = w[i] > mx ? w[i] : mx;
  return (int32_t)std::min<int64_t>((int64_t)mx * 4, INT32_MAX / 4);
}

static int check_graph_host(const uint32_t *rowptr, const uint32_t *adj, int64_t n, int64_t m) {
  if (n < 1 || m < 0 || n > (int64_t)UINT32_MAX || m > (int64_t)UINT32_MAX) return PHDE_ERR_ARG;
  if (rowptr[0] != 0 || (int64_t)rowptr[n] != m) return PHDE_ERR_ARG;
  for (int64_t i = 0; i < n; i++)
    if (rowptr[i + 1] < rowptr[i]) return PHDE_ERR_ARG;
  uint32_t mx = 0;
  for (int64_t e = 0; e < m; e++) mx = adj[e] > mx ? adj[e] : mx;
  if (m > 0 && (int64_t)mx >= n) return PHDE_ERR_ARG;
  return PHDE_OK;
}

static int ctx_init(phde_ctx *c, int64_t n, int64_t m, bool weighted) {
  PHDE_HIP(hipGetDevice(&c->device));
  PHDE_HIP(hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->device));
  PHDE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  PHDE_HIP(hipMalloc(&c->d_rowptr, sizeof(uint32_t) * (n + 1)));
  PHDE_HIP(hipMalloc(&c->d_adj, sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  if (weighted) PHDE_HIP(hipMalloc(&c->d_w, sizeof(int32_t) * std::max<int64_t>(m, 1)));
  PHDE_HIP(hipMalloc(&c->d_deg32, sizeof(int32_t) * n));
  PHDE_HIP(hipMalloc(&c->d_degw, sizeof(double) * n));
  PHDE_HIP(hipHostMalloc(&c->h_small, sizeof(double) * 64, hipHostMallocDefault));
  c->g.n = n;
  c->g.m = m;
  c->g.rowptr = c->d_rowptr;
  c->g.adj = c->d_adj;
  c->g.w = c->d_w;
  c->g.deg32 = c->d_deg32;
  c->g.degw = c->d_degw;
  c->g.orig = nullptr;
  c->g.perm = nullptr;
  return PHDE_OK;
}

static int ctx_finish(phde_ctx *c) {
  const int64_t n = c->g.n;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  degree_kernel<<<blocks, 256, 0, c->stream>>>(c->d_rowptr, c->d_w, n, c->d_deg32, c->d_degw);
  PHDE_HIP(hipGetLastError());
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

// Internal numbering of the graph copy (relabel.hip); degrees follow it;
// then the SpMM's list of long rows (spmm.hip).
static int ctx_renumber(phde_ctx *c) {
  bool done = false;
  PHDE_TRY_RC(relabel_graph(c, &done));
  if (done) PHDE_TRY_RC(ctx_finish(c));
  // unweighted graphs with a deep BFS: the partitioned traversal
  if (c->bfs_hops > 1 && !c->g.w) PHDE_TRY_RC(pbfs_setup(c));
  return spmm_hubs(c);
}

// ---------------------------------------------------------------- temporaries
static thread_local phde_ctx *t_ctx = nullptr;
CtxScope::CtxScope(phde_ctx *c) : prev(t_ctx) { t_ctx = c; }
CtxScope::~CtxScope() { t_ctx = prev; }
// Bytes of temporaries a context keeps for its later calls, per kind.
constexpr size_t kTmpCacheCap = (size_t)32 << 30;

static void cache_drop(std::multimap<size_t, void *> &m, size_t &bytes, bool host) {
  for (auto &kv : m) (void)(host ? hipHostFree(kv.second) : hipFree(kv.second));
  m.clear();
  bytes = 0;
}

void tmp_release(phde_ctx *c) {
  cache_drop(c->dev_cache, c->dev_cached, false);
  cache_drop(c->host_cache, c->host_cached, true);
}

static int cache_take(std::multimap<size_t, void *> &m, size_t &cached, size_t bytes, void **p) {
  auto it = m.find(bytes);
  if (it == m.end()) return 0;
  *p = it->second;
  m.erase(it);
  cached -= bytes;
  return 1;
}

int tmp_malloc(void **p, size_t bytes) {
  phde_ctx *c = t_ctx;
  if (c && cache_take(c->dev_cache, c->dev_cached, bytes, p)) return PHDE_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && c && !c->dev_cache.empty()) {
    (void)hipGetLastError();
    cache_drop(c->dev_cache, c->dev_cached, false);  // the cache held the memory: give it back, try again
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("device allocation of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    return PHDE_ERR_NOMEM;
  }
  return PHDE_OK;
}

void tmp_free(void *p, size_t bytes) {
  phde_ctx *c = t_ctx;
  if (c && c->dev_cached + bytes <= kTmpCacheCap) {
    c->dev_cache.emplace(bytes, p);
    c->dev_cached += bytes;
    return;
  }
  (void)hipFree(p);
}

int tmp_host_malloc(void **p, size_t bytes) {
  phde_ctx *c = t_ctx;
  if (c && cache_take(c->host_cache, c->host_cached, bytes, p)) return PHDE_OK;
  const hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("pinned allocation of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    return PHDE_ERR_NOMEM;
  }
  return PHDE_OK;
}

void tmp_host_free(void *p, size_t bytes) {
  phde_ctx *c = t_ctx;
  if (c && c->host_cached + bytes <= kTmpCacheCap) {
    c->host_cache.emplace(bytes, p);
    c->host_cached += bytes;
    return;
  }
  (void)hipHostFree(p);
}

static void ctx_free(phde_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  tmp_release(c);
  void *dev[] = {c->d_rowptr, c->d_adj, c->d_w, c->d_deg32, c->d_degw, c->d_dist, c->d_q[0],
                 c->d_q[1], c->d_hq[0], c->d_hq[1], c->d_bm[0], c->d_bm[1], c->d_slots, c->d_kc,
                 c->d_part};
  for (void *p : dev)
    if (p) (void)hipFree(p);
  if (c->h_kc) (void)hipHostFree(c->h_kc);
  for (void *p : {(void *)c->sp_hub_row, (void *)c->sp_hub_c0, (void *)c->sp_chunk_row, (void *)c->sp_chunk_beg})
    if (p) (void)hipFree(p);
  if (c->d_orig) (void)hipFree(c->d_orig);
  if (c->d_esrc) (void)hipFree(c->d_esrc);
  if (c->d_perm) (void)hipFree(c->d_perm);
  if (c->d_bar) (void)hipFree(c->d_bar);
  if (c->d_dummy) (void)hipFree(c->d_dummy);
  if (c->h_bar) (void)hipHostFree(c->h_bar);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (hipEvent_t e : c->stage_events) (void)hipEventDestroy(e);
  if (c->h_small) (void)hipHostFree(c->h_small);
  traverse_release(c);
  learned_release(c);
  pbfs_release(c);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

#define PHDE_TRY(x)     \
  do {                  \
    int rc_ = (x);      \
    if (rc_) return rc_; \
  } while (0)

using clk = std::chrono::high_resolution_clock;
static double ms_since(clk::time_point a) {
  return std::chrono::duration<double, std::milli>(clk::now() - a).count();
}

// k-centers bookkeeping shared by every algorithm: the first maximiser of
// the running min distance becomes the next source, with the reference's
// dedupe rule (ParHDE.cpp:142-155; DrawingBase.cpp:157-168).
struct KCenters {
  std::set<int64_t> sources;
  int64_t next = 0;
  // record_old: ParHDE / WParHDE insert the argmax (ParHDE.cpp:143,155);
  // distancePhaseComputation inserts the current source (DrawingBase.cpp:138,169).
  int step(phde_ctx *c, const int32_t *min_dist_dev, const KcResult &kc, bool check_dup,
           bool record_argmax, int64_t current) {
    const int64_t maxval = (int64_t)(kc.key >> 32);
    int64_t idx = (int64_t)(0xffffffffull - (kc.key & 0xffffffffull));
    const int64_t old_idx = record_argmax ? idx : current;
    if (check_dup && sources.count(idx)) {
      std::vector<int32_t> md(c->g.n);
      PHDE_HIP(hipMemcpyAsync(md.data(), min_dist_dev, sizeof(int32_t) * c->g.n,
                              hipMemcpyDeviceToHost, c->stream));
      PHDE_HIP(hipStreamSynchronize(c->stream));
      // the last such vertex in the caller's numbering
      int64_t last = -1;
      for (int64_t i = 0; i < c->g.n; i++) {
        const int64_t v = to_original(c, i);
        if (md[i] == maxval && !sources.count(v)) last = std::max(last, v);
      }
      if (last >= 0) idx = last;
    }
    sources.insert(old_idx);
    next = idx;
    return PHDE_OK;
  }
};

// S^T D S (generalised eigenproblem's right-hand side) via the Gram kernel
// on D^{1/2} S.
int weighted_gram(phde_ctx *c, const void *S_dev, bool fp32, int s, double *out_dev);

struct LayoutOut {
  phde_stats *st = nullptr;
  double t_dist = 0, t_orth = 0, t_mmult = 0, t_eig = 0, t_coords = 0, t_select = 0;
  double energy = 0;  // Hall energy of the layout, when requested
  int ntrav = 0, nacc = 0, ndisc = 0, iters = 0, nexam = 0, ngram = 0;
  long long levels = 0;
  void pivot(int64_t p) {
    if (st && ntrav < 192) st->pivots[ntrav] = p;
    ntrav++;
  }
  // ||a|| of the next examined candidate after its projection
  void examined(double normdist) {
    if (st && nexam < 192) st->normdist[nexam] = normdist;
    nexam++;
  }
};

// Eigen step + coordinates shared by all algorithms.  K_dev: s*s matrix in
// the reference's storage (consumed).  For the generalised mode B_dev holds
// S^T D S.
// A large result (coordinates, distance columns) to the caller's memory.
// The runtime's own staging of a copy into pageable memory measured ≈9.6
// GB/s on MI355X (268 MB of RGG-16.8M coordinates: 28 ms); instead the
// result goes to a pinned buffer of the context in 8 MB chunks, each
// followed by an event, and host threads copy every chunk on to the caller
// as soon as its event completes, so that the host copies overlap the
// transfers.
static int result_to_host(phde_ctx *c, void *host, const void *dev, size_t bytes) {
  auto direct = [&]() -> int {
    PHDE_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
    return PHDE_OK;
  };
  if (bytes < ((size_t)8 << 20)) return direct();
  if (c->h_stage_bytes < bytes) {
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    c->h_stage_bytes = 0;
    if (hipHostMalloc(&c->h_stage, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();  // no pinned memory to spare: the runtime's own staging
      c->h_stage = nullptr;
      return direct();
    }
    c->h_stage_bytes = bytes;
  }
  constexpr size_t kChunk = (size_t)8 << 20;
  const size_t nch = (bytes + kChunk - 1) / kChunk;
  while (c->stage_events.size() < nch) {
    hipEvent_t e;
    PHDE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    c->stage_events.push_back(e);
  }
  const char *src = static_cast<const char *>(c->h_stage);
  char *dst = static_cast<char *>(host);
  for (size_t i = 0; i < nch; i++) {
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    PHDE_HIP(hipMemcpyAsync(static_cast<char *>(c->h_stage) + off, static_cast<const char *>(dev) + off, len,
                            hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipEventRecord(c->stage_events[i], c->stream));
  }
  // thread t copies chunks t, t + nt, ... after their transfers
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min(8u, hw), nch);
  std::vector<int> rc(nt, PHDE_OK);
  auto work = [&](size_t t) {
    for (size_t i = t; i < nch; i += nt) {
      if (hipEventSynchronize(c->stage_events[i]) != hipSuccess) {
        rc[t] = PHDE_ERR_DEVICE;
        return;
      }
      const size_t off = i * kChunk;
      std::memcpy(dst + off, src + off, std::min(kChunk, bytes - off));
    }
  };
  std::vector<std::thread> th;
  size_t started = 1;  // thread 0's share is the calling thread's
  try {
    for (; started < nt; started++) th.emplace_back(work, started);
  } catch (...) {
    // no thread to spare: the calling thread takes the remaining shares
  }
  work(0);
  for (size_t t = started; t < nt; t++) work(t);
  for (auto &t : th) t.join();
  for (int r : rc)
    if (r != PHDE_OK) {
      set_error("copying a result to the host failed");
      return r;
    }
  return PHDE_OK;
}

// S_dev's rows are in the context's internal numbering unless
// caller_rows (then the coordinates need no reordering; no Hall energy).
static int finish_layout(phde_ctx *c, const phde_options *o, const void *S_dev, bool fp32, int s,
                         double *K_dev, const double *B_dev, bool largest, double *coords,
                         LayoutOut &out, bool want_energy = false, bool caller_rows = false,
                         Orth *gram = nullptr) {
  const int nc = o->num_coord;
  auto t0 = clk::now();
  DBuf<double> init, evals, work, V, evecs;
  DBuf<int> status;
  PHDE_TRY(evals.alloc(s));
  PHDE_TRY(work.alloc(3 * (size_t)s * s));
  PHDE_TRY(V.alloc((size_t)s * nc));
  PHDE_TRY(status.alloc(4));
  int h_status[4] = {0, 0, 0, 0};
  if (o->eig_mode == PHDE_EIG_POWER) {
    std::vector<int32_t> r(s * s);
    glibc_rand_sequence(1, s * s, r.data());
    std::vector<double> iv(s * s);
    for (int i = 0; i < s * s; i++) iv[i] = ((double)r[i]) / 2147483647;
    PHDE_TRY(init.alloc((size_t)s * s));
    PHDE_HIP(hipMemcpyAsync(init.p, iv.data(), sizeof(double) * s * s, hipMemcpyHostToDevice,
                            c->stream));
    PHDE_TRY(power_eig_dev(c, K_dev, init.p, evals.p, s, 100000, status.p));
    gather_cols_kernel<<<1, 256, 0, c->stream>>>(K_dev, s, nc, largest, V.p);
  } else {
    PHDE_TRY(evecs.alloc((size_t)s * s));
    PHDE_TRY(sym_eig_dev(c, K_dev, o->eig_mode == PHDE_EIG_GENERALIZED ? B_dev : nullptr, s,
                         evals.p, evecs.p, work.p, status.p));
    // ascending order: First takes the leading columns, Largest the trailing
    // ones in descending order
    gather_cols_kernel<<<1, 256, 0, c->stream>>>(evecs.p, s, nc, largest, V.p);
  }
  PHDE_HIP(hipGetLastError());
  PHDE_HIP(hipMemcpyAsync(h_status, status.p, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  out.t_eig += ms_since(t0);
  if (h_status[0] != 0) {
    if (o->verbose) printf("The algorithm failed to compute eigenvalues.\n");
    set_error("eigensolver failed (status " + std::to_string(h_status[0]) + ")");
    return PHDE_ERR_EIG;
  }
  out.iters = o->eig_mode == PHDE_EIG_POWER ? h_status[1] : 0;
  if (out.st) {
    PHDE_HIP(hipMemcpyAsync(out.st->eigenvalues, evals.p, sizeof(double) * s, hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  }
  auto t1 = clk::now();
  DBuf<double> X, part, inv;
  PHDE_TRY(X.alloc((size_t)c->g.n * nc));
  PHDE_TRY(part.alloc((size_t)nc * c->num_cus * 8 + 64));
  PHDE_TRY(inv.alloc(nc));
  // S in coefficient form (Gram-space CGS): X = C (M V)
  if (gram && gram->gram_pending)
    PHDE_TRY(gram->gram_coords(V.p, nc, X.p, part.p, inv.p));
  else
    PHDE_TRY(coords_compute(c, S_dev, fp32, s, nc, V.p, X.p, part.p, inv.p));
  DBuf<double> Xo;  // rows in the caller's numbering
  if (renumbered(c) && !caller_rows) {
    PHDE_TRY(Xo.alloc((size_t)c->g.n * nc));
    PHDE_TRY(rows_to_original(c, X.p, Xo.p, nc, sizeof(double)));
  }
  PHDE_TRY(result_to_host(c, coords, Xo.p ? Xo.p : X.p, sizeof(double) * c->g.n * nc));
  out.t_coords += ms_since(t1);
  if (want_energy) {
    DBuf<double> hp;
    PHDE_TRY(hp.alloc(2 * (size_t)c->num_cus * 4 + 16));
    double h[2];
    PHDE_TRY(hall_energy_dev(c, X.p, nc, hp.p, h));
    out.energy = h[0] / h[1];
  }
  return PHDE_OK;
}


// max_ij |B_ij / sqrt(B_ii B_jj) - delta_ij| of B = [1 S]^T D [1 S]: how far
// the accepted columns and the ones vector are from D-orthogonal (the
// reference's columns are D-orthogonal and of unit L2 norm, so B is
// diagonal, not the identity, and is normalised by its diagonal first).
static double dorth_from_gram(const std::vector<double> &B, int d) {
  double e = 0;
  for (int i = 0; i < d; i++)
    for (int k = 0; k < d; k++) {
      const double v = B[(size_t)i * d + k] / std::sqrt(B[(size_t)i * d + i] * B[(size_t)k * d + k]);
      e = std::max(e, std::fabs(v - (i == k ? 1.0 : 0.0)));
    }
  return e;
}

static int dorth_error(phde_ctx *c, Orth &orth, bool fp32, int s, double *err) {
  DBuf<double> G;
  PHDE_TRY(G.alloc((size_t)s * s));
  PHDE_TRY(weighted_gram(c, orth.S, fp32, s, G.p));
  std::vector<double> g((size_t)s * s), ones(s + 1);
  PHDE_HIP(hipMemcpyAsync(g.data(), G.p, sizeof(double) * s * s, hipMemcpyDeviceToHost, c->stream));
  PHDE_TRY(orth.ones_dots(ones.data()));  // synchronises
  const int d = s + 1;
  std::vector<double> B((size_t)d * d);
  B[0] = ones[s];
  for (int k = 0; k < s; k++) B[k + 1] = B[(size_t)(k + 1) * d] = ones[k];
  for (int i = 0; i < s; i++)
    for (int k = 0; k < s; k++) B[(size_t)(i + 1) * d + k + 1] = g[(size_t)i * s + k];
  *err = dorth_from_gram(B, d);
  return PHDE_OK;
}

// One candidate column through CGS (ParHDE) or MGS (WParHDE), with the
// reference's column normalisation 1/sqrt(sum d^2) (ParHDE.cpp:176-185) and
// discard rule ||a|| < 0.001 (ParHDE.cpp:269-283); `accepted` counts the
// columns of S so far.
static int orth_candidate(Orth &orth, const phde_options *o, bool weighted, bool fp32,
                          const int32_t *col, double sumsq, int &accepted,
                          LayoutOut &out) {
  const double scale = 1 / std::sqrt(sumsq);
  PHDE_TRY(orth.load_candidate(col, scale));
  const int passes = fp32 ? 2 : 1;
  for (int p = 0; p < passes; p++) {
    if (weighted)
      PHDE_TRY(orth.mgs_pass(accepted));
    else
      PHDE_TRY(orth.cgs_pass(accepted));
  }
  double nd = 0;
  PHDE_TRY(orth.read_norm(&nd));
  out.examined(nd);
  if (nd < 0.001) {
    if (o->verbose) std::printf("discarding vec %d, normdist %g\n", accepted + 1, nd);
    out.ndisc++;
  } else {
    PHDE_TRY(orth.accept(accepted));
    accepted++;
  }
  return PHDE_OK;
}

// Candidates cols[0..count) in order through orth_candidate's steps until
// s columns are accepted; *consumed gets how many were taken up.  ParHDE
// takes them kCgsBlock at a time (Orth::cgs_block: FP64 the same columns,
// bit for bit, with one pass over the earlier columns per block; FP32 CGS2
// with its two passes over the earlier columns shared by the block);
// PHDE_CGS_BLOCK=0 takes them one at a time.
static bool cgs_blocked(bool weighted, bool fp32) {
  static const bool env = [] {
    const char *e = getenv("PHDE_CGS_BLOCK");
    return !(e && atoi(e) == 0);
  }();
  (void)fp32;
  return !weighted && env;
}

// FP64 ParHDE takes its first candidates (up to s, at most kGramMax) through
// the Gram-space CGS (Orth::gram_cgs) when no column is accepted yet;
// PHDE_CGS_GRAM=0 leaves them to the blocked CGS.
static bool cgs_gram(bool weighted, bool fp32) {
  static const bool env = [] {
    const char *e = getenv("PHDE_CGS_GRAM");
    return !(e && atoi(e) == 0);
  }();
  return !weighted && !fp32 && env && cgs_blocked(weighted, fp32);
}

static int orth_candidates(Orth &orth, const phde_options *o, bool weighted, bool fp32,
                           const int32_t *const *cols, const double *sumsq, int count, int &accepted,
                           LayoutOut &out, int *consumed = nullptr) {
  int i = 0;
  // the verdicts of candidates [i, i + looked): counts, norms, verbose lines
  auto record = [&](const std::vector<bool> &taken, const std::vector<double> &norms, int looked, int before) {
    int acc = before;
    for (int b = 0; b < looked; b++) {
      out.examined(norms[b]);
      if (taken[b]) {
        acc++;
      } else {
        if (o->verbose) std::printf("discarding vec %d, normdist %g\n", acc + 1, norms[b]);
        out.ndisc++;
      }
    }
  };
  if (cgs_gram(weighted, fp32) && accepted == 0 && count > 0) {
    const int nc = std::min({count, o->s, kGramMax});
    std::vector<bool> taken;
    std::vector<double> norms;
    int looked = 0;
    PHDE_TRY(orth.gram_cgs(cols, sumsq, nc, o->s, accepted, taken, norms, looked));
    record(taken, norms, looked, 0);
    out.ngram += looked;
    i = looked;
  }
  if (cgs_blocked(weighted, fp32)) {
    while (i < count && accepted < o->s) {
      const int nc = std::min(kCgsBlock, count - i);
      double scales[kCgsBlock];
      for (int b = 0; b < nc; b++) scales[b] = 1 / std::sqrt(sumsq[i + b]);
      std::vector<bool> taken;
      std::vector<double> norms;
      const int before = accepted;
      PHDE_TRY(orth.cgs_block(cols + i, scales, nc, o->s, accepted, taken, norms));
      // candidates looked at: up to the one that completed s (or all)
      int looked = 0, acc = before;
      for (int b = 0; b < nc && acc < o->s; b++) {
        looked++;
        if (taken[b]) acc++;
      }
      record(taken, norms, looked, before);
      i += looked;
    }
  } else {
    for (; i < count && accepted < o->s; i++)
      PHDE_TRY(orth_candidate(orth, o, weighted, fp32, cols[i], sumsq[i], accepted, out));
  }
  if (consumed) *consumed = i;
  return PHDE_OK;
}

static int not_enough_pivots(const phde_options *o, int accepted, int limit) {
  if (o->verbose)
    std::fprintf(stderr, "Number of Pivots obtained: %d inadequate, increase the Pivots Limits: %d\n",
                 accepted + 1, limit);
  set_error("only " + std::to_string(accepted) + " of " + std::to_string(o->s) +
            " pivots are independent");
  return PHDE_ERR_PIVOTS;
}

// L*S, S^T(LS) (+ S^T D S for the generalised mode), eigensolve, coordinates
// (ParHDE.cpp:288-352).
static int hde_tail(phde_ctx *c, const phde_options *o, Orth &orth, bool fp32, double *coords,
                    LayoutOut &out, bool want_energy) {
  const int s = o->s;
  // PHDE_GRAM_TAIL=0: form S and take the vector path below
  static const bool gram_tail = [] {
    const char *e = getenv("PHDE_GRAM_TAIL");
    return !(e && atoi(e) == 0);
  }();
  if (orth.gram_pending && !gram_tail) PHDE_TRY(orth.form_S());
  if (orth.gram_pending) {
    // S in coefficient form: S^T L S (and S^T D S) from the exact C^T L C
    auto t2 = clk::now();
    DBuf<double> K, B;
    PHDE_TRY(K.alloc((size_t)s * s));
    const bool gen = o->eig_mode == PHDE_EIG_GENERALIZED;
    if (gen) PHDE_TRY(B.alloc((size_t)s * s));
    std::vector<double> Bf;
    if (o->diagnostics && out.st) Bf.resize((size_t)(s + 1) * (s + 1));
    bool ok = false;
    PHDE_TRY(orth.gram_lap(K.p, B.p, Bf.empty() ? nullptr : Bf.data(), ok));
    if (ok) {
      PHDE_HIP(hipStreamSynchronize(c->stream));
      out.t_mmult += ms_since(t2);
      if (!Bf.empty()) out.st->dorth_err = dorth_from_gram(Bf, s + 1);
      return finish_layout(c, o, nullptr, fp32, s, K.p, B.p, false, coords, out, want_energy, false, &orth);
    }
    PHDE_TRY(orth.form_S());
  }
  if (o->diagnostics && out.st) PHDE_TRY(dorth_error(c, orth, fp32, s, &out.st->dorth_err));
  auto t2 = clk::now();
  Lap lap;
  PHDE_TRY(lap.setup(c, c->g.n, s, fp32));
  PHDE_TRY(lap.run(orth.S));
  DBuf<double> B;
  if (o->eig_mode == PHDE_EIG_GENERALIZED) {
    PHDE_TRY(B.alloc((size_t)s * s));
    PHDE_TRY(weighted_gram(c, orth.S, fp32, s, B.p));
  }
  PHDE_HIP(hipStreamSynchronize(c->stream));
  out.t_mmult += ms_since(t2);
  return finish_layout(c, o, orth.S, fp32, s, lap.SLS, B.p, false, coords, out, want_energy);
}

static int check_weighting(phde_ctx *c, bool weighted) {
  if (weighted != (c->g.w != nullptr)) {
    set_error(weighted ? "Weighted ParHDE code requires the weighted edges"
                       : "ParHDE code doesn't requires the weighted edges");
    return PHDE_ERR_WEIGHTS;
  }
  return PHDE_OK;
}

// ParHDE / WParHDE with k-centers pivots.
static int layout_hde(phde_ctx *c, const phde_options *o, double *coords, LayoutOut &out,
                      bool want_energy = false) {
  const bool weighted = o->algorithm == PHDE_ALGO_WPARHDE;
  PHDE_TRY(check_weighting(c, weighted));
  const int64_t n = c->g.n;
  const int s = o->s;
  const bool fp32 = o->precision == PHDE_PREC_FP32;
  const int limit = 3 * s;
  Orth orth;
  PHDE_TRY(orth.setup(c, n, s, fp32));
  // ones vector U[:,0] = 1 / ||1|| (ParHDE.cpp:94-108)
  const double u0 = 1.0 * (1 / std::sqrt((double)n));
  PHDE_TRY(orth.init_ones(c->d_degw, u0));
  DBuf<int32_t> min_dist, cols;
  PHDE_TRY(min_dist.alloc(n));
  PHDE_TRY(fill_i32(c, min_dist.p, n, INT_MAX));
  const bool eager = o->eager && !weighted;
  // lazy FP64 ParHDE traverses its first s candidates before taking them
  // through blocked CGS (the k-centers chain does not depend on them)
  const bool lazy_blocked = !eager && cgs_blocked(weighted, fp32);
  PHDE_TRY(cols.alloc(eager ? (size_t)n * limit : lazy_blocked ? (size_t)n * std::min(s, limit) : (size_t)n));
  std::vector<double> sumsq(limit + 1, 0);
  KCenters kc;
  kc.next = o->first_pivot;
  int accepted = 0;

  auto traverse = [&](int run_count, int32_t *col) -> int {
    const int64_t src = kc.next;
    out.pivot(src);
    if (weighted)
      PHDE_TRY(sssp_run(c, to_internal(c, src), o->delta, col, min_dist.p, true));
    else
      PHDE_TRY(bfs_run(c, to_internal(c, src), o->alpha, o->beta, col, min_dist.p, true));
    const KcResult r = *c->h_kc;
    if (r.unreached) {
      set_error("graph is not connected: " + std::to_string(r.unreached) +
                " vertices unreached from " + std::to_string(src));
      return PHDE_ERR_UNREACHED;
    }
    out.levels += r.depth;
    sumsq[run_count] = u128_to_double(r.sumsq);
    const bool first = weighted ? accepted == 0 : run_count == 1;
    return kc.step(c, min_dist.p, r, !first, true, src);
  };

  if (eager) {
    auto t0 = clk::now();
    for (int rc = 1; rc <= limit; rc++) PHDE_TRY(traverse(rc, cols.p + (size_t)(rc - 1) * n));
    out.t_dist += ms_since(t0);
    auto t1 = clk::now();
    std::vector<const int32_t *> cp(limit);
    for (int rc = 1; rc <= limit; rc++) cp[rc - 1] = cols.p + (size_t)(rc - 1) * n;
    PHDE_TRY(orth_candidates(orth, o, weighted, fp32, cp.data(), sumsq.data() + 1, limit, accepted, out));
    out.t_orth += ms_since(t1);
  } else {
    int rc = 1;
    if (lazy_blocked) {
      const int first = std::min(s, limit);
      auto t0 = clk::now();
      for (; rc <= first; rc++) PHDE_TRY(traverse(rc, cols.p + (size_t)(rc - 1) * n));
      out.t_dist += ms_since(t0);
      auto t1 = clk::now();
      std::vector<const int32_t *> cp(first);
      for (int i = 0; i < first; i++) cp[i] = cols.p + (size_t)i * n;
      PHDE_TRY(orth_candidates(orth, o, weighted, fp32, cp.data(), sumsq.data() + 1, first, accepted, out));
      out.t_orth += ms_since(t1);
    }
    for (; rc <= limit && accepted < s; rc++) {
      auto t0 = clk::now();
      PHDE_TRY(traverse(rc, cols.p));
      auto t1 = clk::now();
      PHDE_TRY(orth_candidate(orth, o, weighted, fp32, cols.p, sumsq[rc], accepted, out));
      out.t_dist += std::chrono::duration<double, std::milli>(t1 - t0).count();
      out.t_orth += ms_since(t1);
    }
  }
  out.nacc = accepted;
  if (accepted < s) return not_enough_pivots(o, accepted, limit);
  return hde_tail(c, o, orth, fp32, coords, out, want_energy);
}

// Sources of the random strategies: srand(seed), then rand() % n per source
// (ParHDE.cpp:160-162, DrawingBase.cpp:175-177; the reference seeds with
// time(NULL), ParHDE.cpp:113).  RandFine and RandCoarse differ in the
// reference only in how each column is traversed (DOBFS vs sequential BFS,
// ParHDE.cpp:163-171); both give the same levels, here from the multi-source
// BFS.
static std::vector<int64_t> random_sources(const phde_options *o, int64_t n, int count) {
  std::vector<int32_t> r(count);
  glibc_rand_sequence((unsigned)o->seed, count, r.data());
  std::vector<int64_t> src(count);
  for (int i = 0; i < count; i++) src[i] = (int64_t)r[i] % n;
  return src;
}

// Distance columns of given sources, 64 per multi-source sweep (SSSP one at
// a time for weighted graphs).  D must hold n * k entries; sumsq gets the
// per-column sums of squares.
static int columns_of(phde_ctx *c, const phde_options *o, const int64_t *src, int k, int32_t *D,
                      double *sumsq, LayoutOut &out) {
  const int64_t n = c->g.n;
  std::vector<int64_t> isrc(k);
  for (int i = 0; i < k; i++) {
    out.pivot(src[i]);
    isrc[i] = to_internal(c, src[i]);
  }
  if (c->g.w) {
    for (int i = 0; i < k; i++) {
      PHDE_TRY(sssp_run(c, isrc[i], o->delta, D + (size_t)i * n, nullptr, false));
      unsigned long long unreached = 0;
      PHDE_TRY(column_sumsq(c, D + (size_t)i * n, 1, sumsq + i, &unreached));
      if (unreached) {
        set_error("graph is not connected");
        return PHDE_ERR_UNREACHED;
      }
    }
    return PHDE_OK;
  }
  for (int b = 0; b < k; b += 64) {
    const int kb = std::min(64, k - b);
    PHDE_TRY(msbfs_run(c, isrc.data() + b, kb, D + (size_t)b * n));
    unsigned long long unreached = 0;
    PHDE_TRY(column_sumsq(c, D + (size_t)b * n, kb, sumsq + b, &unreached));
    if (unreached) {
      set_error("graph is not connected: " + std::to_string(unreached) + " unreached entries");
      return PHDE_ERR_UNREACHED;
    }
  }
  return PHDE_OK;
}

// ParHDE / WParHDE with random pivots: 3s candidate sources, traversed in
// batches of 64 until s columns are accepted.
static int layout_hde_random(phde_ctx *c, const phde_options *o, double *coords, LayoutOut &out,
                             bool want_energy) {
  const bool weighted = o->algorithm == PHDE_ALGO_WPARHDE;
  PHDE_TRY(check_weighting(c, weighted));
  const int64_t n = c->g.n;
  const int s = o->s;
  const bool fp32 = o->precision == PHDE_PREC_FP32;
  const int limit = 3 * s;
  const std::vector<int64_t> src = random_sources(o, n, limit);
  Orth orth;
  PHDE_TRY(orth.setup(c, n, s, fp32));
  PHDE_TRY(orth.init_ones(c->d_degw, 1.0 * (1 / std::sqrt((double)n))));
  const int batch = weighted ? 1 : 64;
  DBuf<int32_t> D;
  PHDE_TRY(D.alloc((size_t)n * std::min(batch, limit)));
  std::vector<double> sumsq(64);
  int accepted = 0;
  for (int b = 0; b < limit && accepted < s; b += batch) {
    const int kb = std::min(batch, limit - b);
    auto t0 = clk::now();
    PHDE_TRY(columns_of(c, o, src.data() + b, kb, D.p, sumsq.data(), out));
    auto t1 = clk::now();
    out.t_dist += std::chrono::duration<double, std::milli>(t1 - t0).count();
    std::vector<const int32_t *> cp(kb);
    for (int i = 0; i < kb; i++) cp[i] = D.p + (size_t)i * n;
    PHDE_TRY(orth_candidates(orth, o, weighted, fp32, cp.data(), sumsq.data(), kb, accepted, out));
    out.t_orth += ms_since(t1);
  }
  out.nacc = accepted;
  if (accepted < s) return not_enough_pivots(o, accepted, limit);
  return hde_tail(c, o, orth, fp32, coords, out, want_energy);
}

// ParHDE with learned pivots: the first two k-centers traversals give the
// features and the first two columns, the network scores the vertices, the
// greedy suppression proposes pivots in batches (up to 64 in all) and one
// multi-source BFS per batch traverses them; CGS takes the columns in
// proposal order.
static int layout_learned(phde_ctx *c, const phde_options *o, double *coords, LayoutOut &out) {
  if (o->algorithm != PHDE_ALGO_PARHDE) {
    set_error("learned pivots are available for ParHDE only");
    return PHDE_ERR_UNSUPPORTED;
  }
  PHDE_TRY(check_weighting(c, false));
  if (!c->net) {
    set_error("no pivot network attached (phde_set_pivot_net)");
    return PHDE_ERR_ARG;
  }
  const int64_t n = c->g.n;
  const int s = o->s;
  const bool fp32 = o->precision == PHDE_PREC_FP32;
  const int kmax = (int)std::min<int64_t>(64, n);
  auto t0 = clk::now();
  int64_t p2 = 0;
  int e1 = 0, e2 = 0;
  PHDE_TRY(pivot_features(c, o->first_pivot, &p2, &e1, &e2));
  PHDE_TRY(net_score(c, c->net));
  const double rho = o->nms_rho > 0 ? o->nms_rho : 1.0;
  const int r = (int)std::max(1L, std::min(2046L, std::lround(rho * e1 / std::sqrt((double)s))));
  std::vector<int64_t> picks(kmax);
  picks[0] = o->first_pivot;
  picks[1] = p2;
  out.t_select += ms_since(t0);
  out.pivot(picks[0]);
  out.pivot(picks[1]);
  out.levels += e1 + e2;
  std::vector<double> sumsq(kmax + 1);
  PHDE_TRY(column_sumsq(c, c->lp_d12, 2, sumsq.data(), nullptr));
  DBuf<int32_t> D;
  PHDE_TRY(D.alloc((size_t)n * std::max(1, kmax - 2)));
  Orth orth;
  PHDE_TRY(orth.setup(c, n, s, fp32));
  PHDE_TRY(orth.init_ones(c->d_degw, 1.0 * (1 / std::sqrt((double)n))));
  // Proposals in batches: s + 2 first (two spares for discarded columns),
  // then as many more as columns are still missing (+2), up to 64.  Each
  // batch continues the selection (the distance-to-pivots field persists)
  // and traverses only its new pivots.
  int have = 2, accepted = 0, next = 0;
  int target = std::min(kmax, s + 2);
  for (;;) {
    if (target > have) {
      auto ts = clk::now();
      PHDE_TRY(greedy_select(c, c->lp_score, have, target, r, picks.data()));
      out.t_select += ms_since(ts);
      auto td = clk::now();
      PHDE_TRY(columns_of(c, o, picks.data() + have, target - have, D.p + (size_t)(have - 2) * n,
                          sumsq.data() + have, out));
      out.t_dist += ms_since(td);
      have = target;
    }
    auto to = clk::now();
    if (next < have && accepted < s) {
      std::vector<const int32_t *> cp(have - next);
      for (int i = next; i < have; i++) cp[i - next] = i < 2 ? c->lp_d12 + (size_t)i * n : D.p + (size_t)(i - 2) * n;
      int used = 0;
      PHDE_TRY(orth_candidates(orth, o, false, fp32, cp.data(), sumsq.data() + next, have - next, accepted, out, &used));
      next += used;
    }
    out.t_orth += ms_since(to);
    if (accepted >= s || have >= kmax) break;
    target = std::min(kmax, have + (s - accepted) + 2);
  }
  out.nacc = accepted;
  if (c->timer.enabled && out.st) out.st->n_candidates = have;
  if (accepted < s) return not_enough_pivots(o, accepted, have);
  return hde_tail(c, o, orth, fp32, coords, out, true);
}

// PivotMDS / PHDE driver: s k-centers traversals (DrawingBase.cpp:125-172),
// centring, C^T C, eig(Largest), C * V, all in the caller's numbering.
static int layout_mds(phde_ctx *c, const phde_options *o, double *coords, LayoutOut &out) {
  const int64_t n = c->g.n;
  const int s = o->s;
  const bool dbl = o->algorithm == PHDE_ALGO_PIVOTMDS;
  DBuf<int32_t> min_dist, D;
  PHDE_TRY(min_dist.alloc(n));
  PHDE_TRY(fill_i32(c, min_dist.p, n, INT_MAX));
  PHDE_TRY(D.alloc((size_t)n * s));
  KCenters kc;
  kc.next = o->first_pivot;
  auto t0 = clk::now();
  if (o->strategy == PHDE_PIVOTS_RANDFINE || o->strategy == PHDE_PIVOTS_RANDCOARSE) {
    // DrawingBase.cpp:173-190: s random sources
    const std::vector<int64_t> src = random_sources(o, n, s);
    std::vector<double> sq(s);
    PHDE_TRY(columns_of(c, o, src.data(), s, D.p, sq.data(), out));
  }
  for (int rc = 0; rc < s && o->strategy == PHDE_PIVOTS_KCENTERS; rc++) {
    const int64_t src = kc.next;
    out.pivot(src);
    int32_t *col = D.p + (size_t)rc * n;
    if (c->g.w)
      PHDE_TRY(sssp_run(c, to_internal(c, src), o->delta, col, min_dist.p, true));
    else
      PHDE_TRY(bfs_run(c, to_internal(c, src), o->alpha, o->beta, col, min_dist.p, true));
    const KcResult r = *c->h_kc;
    if (r.unreached) {
      set_error("graph is not connected");
      return PHDE_ERR_UNREACHED;
    }
    out.levels += r.depth;
    PHDE_TRY(kc.step(c, min_dist.p, r, rc != 0, false, src));
  }
  out.t_dist += ms_since(t0);
  auto t1 = clk::now();
  DBuf<int32_t> Dc;  // the columns in the caller's numbering
  if (renumbered(c)) {
    PHDE_TRY(Dc.alloc((size_t)n * s));
    PHDE_TRY(rows_to_original(c, D.p, Dc.p, s, sizeof(int32_t)));
  }
  const int32_t *Dv = Dc.p ? Dc.p : D.p;
  DBuf<double> rowterm, C, colterm_d;
  DBuf<unsigned long long> acc;
  PHDE_TRY(rowterm.alloc(n));
  PHDE_TRY(acc.alloc(2 * (size_t)s));
  PHDE_TRY(C.alloc((size_t)n * s));
  PHDE_TRY(colterm_d.alloc(s));
  PHDE_HIP(hipMemsetAsync(acc.p, 0, sizeof(unsigned long long) * 2 * s, c->stream));
  PHDE_TRY(mds_center_sums(c, Dv, n, s, dbl, rowterm.p, acc.p));
  std::vector<unsigned long long> ha(2 * (size_t)s);
  PHDE_HIP(hipMemcpyAsync(ha.data(), acc.p, sizeof(unsigned long long) * 2 * s, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  std::vector<double> colsum(s), colterm(s);
  bool inexact = false;
  for (int i = 0; i < s; i++) {
    colsum[i] = u128_to_double(&ha[2 * (size_t)i]);
    inexact |= ha[2 * (size_t)i + 1] != 0 || ha[2 * (size_t)i] >= (1ull << 53);
  }
  if (inexact) {
    // a total at or above 2^53: the reference's sequential sums round
    PHDE_TRY(mds_column_sums_seq(c, Dv, n, s, dbl, colterm_d.p));
    PHDE_HIP(hipMemcpyAsync(colsum.data(), colterm_d.p, sizeof(double) * s, hipMemcpyDeviceToHost, c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  }
  double total = 0;
  for (int i = 0; i < s; i++) {
    colterm[i] = dbl ? colsum[i] : colsum[i] / (double)n;  // PHDE.cpp:40: columnSum /= n
    total += colsum[i];                                     // PivotMDS.cpp:49
  }
  PHDE_HIP(hipMemcpyAsync(colterm_d.p, colterm.data(), sizeof(double) * s, hipMemcpyHostToDevice, c->stream));
  const double rN = (double)1 / (double)n, rK = (double)1 / (double)s;
  PHDE_TRY(mds_center(c, Dv, n, s, colterm_d.p, rowterm.p, rN, rK, total, dbl, C.p));
  out.t_orth += ms_since(t1);
  auto t2 = clk::now();
  DBuf<double> K;
  PHDE_TRY(K.alloc((size_t)s * s));
  hipEvent_t ev;
  c->timer.begin(c->stream, PHDE_KERNEL_GRAM, &ev);
  PHDE_TRY(mds_ctc(c, C.p, n, s, K.p));
  c->timer.end(c->stream, PHDE_KERNEL_GRAM, ev);
  PHDE_HIP(hipStreamSynchronize(c->stream));
  out.t_mmult += ms_since(t2);
  out.nacc = s;
  return finish_layout(c, o, C.p, false, s, K.p, nullptr, true, coords, out, false, true);
}

}  // namespace phde

using namespace phde;

extern "C" {

void phde_default_options(phde_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->algorithm = PHDE_ALGO_PARHDE;
  o->s = 10;
  o->num_coord = 2;
  o->strategy = PHDE_PIVOTS_KCENTERS;
  o->precision = PHDE_PREC_FP64;
  o->eig_mode = PHDE_EIG_POWER;
  o->eager = 0;
  o->delta = 0;
  o->alpha = 15;
  o->beta = 18;
  o->verbose = 0;
  o->first_pivot = 0;
  o->seed = 1;
  o->judge_tol = 0.02;
  o->nms_rho = 1.0;
  o->judge = 0;
}

const char *phde_version(void) { return "parhde-mi355x 0.1 (gfx950)"; }
const char *phde_last_error(void) { return g_err.c_str(); }

int phde_device_count(int *count) {
  if (!count) return PHDE_ERR_ARG;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return PHDE_OK;
}

int phde_create(const uint32_t *rowptr, const uint32_t *adj, const int32_t *weights, int64_t n,
                int64_t m, phde_ctx **out) {
  if (!rowptr || !out || (m > 0 && !adj)) return PHDE_ERR_ARG;
  *out = nullptr;
  if (check_graph_host(rowptr, adj, n, m)) {
    set_error("malformed CSR (row pointers or ids out of range)");
    return PHDE_ERR_ARG;
  }
  phde_ctx *c = new phde_ctx();
  int rc = ctx_init(c, n, m, weights != nullptr);
  if (rc == PHDE_OK) {
    if (hipMemcpy(c->d_rowptr, rowptr, sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (m > 0 && hipMemcpy(c->d_adj, adj, sizeof(uint32_t) * m, hipMemcpyHostToDevice) != hipSuccess) ||
        (weights && m > 0 &&
         hipMemcpy(c->d_w, weights, sizeof(int32_t) * m, hipMemcpyHostToDevice) != hipSuccess)) {
      set_error("graph upload failed");
      rc = PHDE_ERR_DEVICE;
    }
  }
  if (rc == PHDE_OK) rc = ctx_finish(c);
  if (rc != PHDE_OK) {
    ctx_free(c);
    delete c;
    return rc;
  }
  if (weights) {
    for (int64_t i = 0; i < m; i++)
      if (weights[i] < 0) {
        ctx_free(c);
        delete c;
        set_error("negative edge weight");
        return PHDE_ERR_ARG;
      }
    c->default_delta = pick_delta(weights, m);
  }
  rc = ctx_renumber(c);
  if (rc != PHDE_OK) {
    ctx_free(c);
    delete c;
    return rc;
  }
  *out = c;
  return PHDE_OK;
}

int phde_create_device(const uint32_t *rowptr_dev, const uint32_t *adj_dev,
                       const int32_t *weights_dev, int64_t n, int64_t m, phde_ctx **out) {
  if (!rowptr_dev || !out || (m > 0 && !adj_dev)) return PHDE_ERR_ARG;
  // validate through a host copy: the kernels trust these arrays
  std::vector<uint32_t> rp(n + 1), ad(std::max<int64_t>(m, 1));
  if (hipMemcpy(rp.data(), rowptr_dev, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToHost) !=
          hipSuccess ||
      (m > 0 && hipMemcpy(ad.data(), adj_dev, sizeof(uint32_t) * m, hipMemcpyDeviceToHost) != hipSuccess))
    return PHDE_ERR_DEVICE;
  if (check_graph_host(rp.data(), ad.data(), n, m)) return PHDE_ERR_ARG;
  std::vector<int32_t> wh;
  if (weights_dev && m > 0) {
    wh.resize(m);
    if (hipMemcpy(wh.data(), weights_dev, sizeof(int32_t) * m, hipMemcpyDeviceToHost) != hipSuccess)
      return PHDE_ERR_DEVICE;
    for (int64_t i = 0; i < m; i++)
      if (wh[i] < 0) {
        set_error("negative edge weight");
        return PHDE_ERR_ARG;
      }
  }
  phde_ctx *c = new phde_ctx();
  int rc = ctx_init(c, n, m, weights_dev != nullptr);
  if (rc == PHDE_OK) {
    if (hipMemcpy(c->d_rowptr, rowptr_dev, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToDevice) !=
            hipSuccess ||
        (m > 0 && hipMemcpy(c->d_adj, adj_dev, sizeof(uint32_t) * m, hipMemcpyDeviceToDevice) != hipSuccess) ||
        (weights_dev && m > 0 &&
         hipMemcpy(c->d_w, weights_dev, sizeof(int32_t) * m, hipMemcpyDeviceToDevice) != hipSuccess))
      rc = PHDE_ERR_DEVICE;
  }
  if (rc == PHDE_OK) rc = ctx_finish(c);
  if (rc == PHDE_OK && !wh.empty()) c->default_delta = pick_delta(wh.data(), m);
  if (rc == PHDE_OK) rc = ctx_renumber(c);
  if (rc != PHDE_OK) {
    ctx_free(c);
    delete c;
    return rc;
  }
  *out = c;
  return PHDE_OK;
}

int phde_destroy(phde_ctx *ctx) {
  if (!ctx) return PHDE_ERR_ARG;
  ctx_free(ctx);
  delete ctx;
  return PHDE_OK;
}

int64_t phde_num_vertices(const phde_ctx *ctx) { return ctx ? ctx->g.n : -1; }

int phde_vertex_order(const phde_ctx *c, uint32_t *order, int *is_renumbered) {
  if (!c) return PHDE_ERR_ARG;
  if (is_renumbered) *is_renumbered = renumbered(c) ? 1 : 0;
  if (order)
    for (int64_t k = 0; k < c->g.n; k++) order[k] = (uint32_t)to_original(c, k);
  return PHDE_OK;
}

int phde_num_parts(const phde_ctx *c, int64_t *parts) {
  if (!c || !parts) return PHDE_ERR_ARG;
  *parts = pbfs_num_parts(c);
  return PHDE_OK;
}

int phde_num_boundary(const phde_ctx *c, int64_t *boundary) {
  if (!c || !boundary) return PHDE_ERR_ARG;
  *boundary = pbfs_num_boundary(c);
  return PHDE_OK;
}

int phde_bfs_dev(phde_ctx *c, int64_t source, int alpha, int beta, int32_t *dist_dev) {
  CtxScope scope_(c);
  if (!c || !dist_dev || source < 0 || source >= c->g.n) return PHDE_ERR_ARG;
  PHDE_TRY(bfs_run(c, to_internal(c, source), alpha, beta, c->d_dist, nullptr, false));
  if (renumbered(c)) {
    PHDE_TRY(dist_finalize(c, c->d_dist, c->d_dist, kUnvisited, -1));
    PHDE_TRY(rows_to_original(c, c->d_dist, dist_dev, 1, sizeof(int32_t)));
  } else {
    PHDE_TRY(dist_finalize(c, c->d_dist, dist_dev, kUnvisited, -1));
  }
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

// Copies n x k int32 columns of internal rows to the caller's host buffer
// in the caller's numbering.
static int columns_to_host(phde_ctx *c, const int32_t *D, int k, int32_t *host) {
  DBuf<int32_t> T;
  const int32_t *src = D;
  if (renumbered(c)) {
    PHDE_TRY(T.alloc((size_t)c->g.n * k));
    PHDE_TRY(rows_to_original(c, D, T.p, k, sizeof(int32_t)));
    src = T.p;
  }
  return result_to_host(c, host, src, sizeof(int32_t) * c->g.n * k);
}

int phde_bfs(phde_ctx *c, int64_t source, int alpha, int beta, int32_t *dist) {
  CtxScope scope_(c);
  if (!c || !dist || source < 0 || source >= c->g.n) return PHDE_ERR_ARG;
  PHDE_TRY(bfs_run(c, to_internal(c, source), alpha, beta, c->d_dist, nullptr, false));
  PHDE_TRY(dist_finalize(c, c->d_dist, c->d_dist, kUnvisited, -1));
  return columns_to_host(c, c->d_dist, 1, dist);
}

int phde_sssp(phde_ctx *c, int64_t source, int32_t delta, int32_t *dist) {
  CtxScope scope_(c);
  if (!c || !dist || source < 0 || source >= c->g.n) return PHDE_ERR_ARG;
  if (!c->g.w) return PHDE_ERR_WEIGHTS;
  PHDE_TRY(bfs_setup(c));
  PHDE_TRY(sssp_run(c, to_internal(c, source), delta, c->d_dist, nullptr, false));
  return columns_to_host(c, c->d_dist, 1, dist);
}

int phde_msbfs(phde_ctx *c, const int64_t *sources, int k, int32_t *dist) {
  CtxScope scope_(c);
  if (!c || !sources || !dist || k < 1 || k > 64) return PHDE_ERR_ARG;
  std::vector<int64_t> isrc(k);
  for (int i = 0; i < k; i++) {
    if (sources[i] < 0 || sources[i] >= c->g.n) return PHDE_ERR_ARG;
    isrc[i] = to_internal(c, sources[i]);
  }
  DBuf<int32_t> D;
  PHDE_TRY(D.alloc((size_t)c->g.n * k));
  PHDE_TRY(msbfs_run(c, isrc.data(), k, D.p));
  return columns_to_host(c, D.p, k, dist);
}

int phde_layout(phde_ctx *c, const phde_options *opt, double *coords, phde_stats *stats) {
  CtxScope scope_(c);
  if (!c || !opt || !coords) return PHDE_ERR_ARG;
  const phde_options &o = *opt;
  if (o.s < 1 || o.s > 64 || o.num_coord < 1 || o.num_coord > 8 || o.num_coord > o.s ||
      o.first_pivot < 0 || o.first_pivot >= c->g.n || o.algorithm < 0 || o.algorithm > 3 ||
      o.precision < 0 || o.precision > 1 || o.eig_mode < 0 || o.eig_mode > 2) {
    set_error("invalid layout options");
    return PHDE_ERR_ARG;
  }
  if (o.strategy < PHDE_PIVOTS_RANDCOARSE || o.strategy > PHDE_PIVOTS_LEARNED) {
    set_error("unknown pivot strategy");
    return PHDE_ERR_ARG;
  }
  if (c->g.n < 2) {
    set_error("graph too small");
    return PHDE_ERR_ARG;
  }
  PHDE_TRY(bfs_setup(c));
  if (stats) std::memset(stats, 0, sizeof(*stats));
  c->timer.reset();
  c->timer.enabled = stats != nullptr;
  LayoutOut out;
  out.st = stats;
  auto t0 = clk::now();
  int rc;
  const bool hde = o.algorithm == PHDE_ALGO_PARHDE || o.algorithm == PHDE_ALGO_WPARHDE;
  const bool random = o.strategy == PHDE_PIVOTS_RANDFINE || o.strategy == PHDE_PIVOTS_RANDCOARSE;
  int fallback = 0;
  double e_kc = 0;
  if (o.strategy == PHDE_PIVOTS_LEARNED) {
    // k-centers is the fallback (too few independent proposals) and, with
    // o.judge, the judge: its layout replaces the learned one when the
    // learned Hall energy is more than judge_tol above it.
    phde_options ok = o;
    ok.strategy = PHDE_PIVOTS_KCENTERS;
    const double tol = o.judge_tol;
    rc = layout_learned(c, &o, coords, out);
    if (rc == PHDE_ERR_PIVOTS || rc == kSelectUnavailable) {
      fallback = 1;
      LayoutOut ko;
      ko.st = stats;
      rc = layout_hde(c, &ok, coords, ko, true);
      out.t_dist += ko.t_dist, out.t_orth += ko.t_orth, out.t_mmult += ko.t_mmult;
      out.t_eig += ko.t_eig, out.t_coords += ko.t_coords, out.nacc = ko.nacc, out.energy = ko.energy;
      e_kc = ko.energy;
    } else if (rc == PHDE_OK && o.judge) {
      std::vector<double> Xk((size_t)c->g.n * o.num_coord);
      LayoutOut ko;
      rc = layout_hde(c, &ok, Xk.data(), ko, true);
      e_kc = ko.energy;
      if (rc == PHDE_OK && out.energy > (1 + tol) * e_kc) {
        fallback = 1;
        std::memcpy(coords, Xk.data(), sizeof(double) * Xk.size());
        out.energy = e_kc;
      }
    }
  } else if (hde) {
    rc = random ? layout_hde_random(c, &o, coords, out, false) : layout_hde(c, &o, coords, out, false);
  } else {
    rc = layout_mds(c, &o, coords, out);
  }
  const double total = ms_since(t0);
  c->timer.resolve();
  c->timer.enabled = false;
  if (stats) {
    stats->ms_total = total;
    stats->ms_distance = out.t_dist;
    stats->ms_orth = out.t_orth;
    stats->ms_mmult = out.t_mmult;
    stats->ms_eig = out.t_eig;
    stats->ms_coords = out.t_coords;
    stats->n_traversals = out.ntrav;
    stats->n_accepted = out.nacc;
    stats->n_discarded = out.ndisc;
    stats->eig_iters = out.iters;
    stats->levels_total = out.levels;
    stats->ms_select = out.t_select;
    stats->hall_energy = out.energy;
    stats->hall_energy_kcenters = e_kc;
    stats->fallback = fallback;
    stats->n_examined = out.nexam;
    stats->n_gram = out.ngram;
    for (int k = 0; k < 16; k++) {
      stats->kernel_ms[k] = c->timer.ms[k];
      stats->kernel_launches[k] = c->timer.launches[k];
    }
  }
  return rc;
}

int phde_kcenters(phde_ctx *c, int s, int64_t first, int64_t *sources, int32_t *min_dist) {
  CtxScope scope_(c);
  if (!c || s < 1 || first < 0 || first >= c->g.n) return PHDE_ERR_ARG;
  const int64_t n = c->g.n;
  PHDE_TRY(bfs_setup(c));
  DBuf<int32_t> md;
  PHDE_TRY(md.alloc(n));
  PHDE_TRY(fill_i32(c, md.p, n, INT_MAX));
  KCenters kc;
  kc.next = first;
  for (int i = 0; i < s; i++) {
    const int64_t src = kc.next;
    if (sources) sources[i] = src;
    if (c->g.w)
      PHDE_TRY(sssp_run(c, to_internal(c, src), 0, c->d_dist, md.p, true));
    else
      PHDE_TRY(bfs_run(c, to_internal(c, src), 15, 18, c->d_dist, md.p, true));
    const KcResult r = *c->h_kc;
    if (r.unreached) {
      set_error("graph is not connected");
      return PHDE_ERR_UNREACHED;
    }
    PHDE_TRY(kc.step(c, md.p, r, i != 0, true, src));
  }
  if (min_dist) return columns_to_host(c, md.p, 1, min_dist);
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return PHDE_OK;
}

int phde_power_eig(phde_ctx *c, double *K, double *evals, int dim, int32_t *iters) {
  CtxScope scope_(c);
  if (!c || !K || !evals || dim < 1 || dim > 64) return PHDE_ERR_ARG;
  DBuf<double> Kd, init, ev;
  DBuf<int> status;
  PHDE_TRY(Kd.alloc((size_t)dim * dim));
  PHDE_TRY(init.alloc((size_t)dim * dim));
  PHDE_TRY(ev.alloc(dim));
  PHDE_TRY(status.alloc(4));
  std::vector<int32_t> r(dim * dim);
  glibc_rand_sequence(1, dim * dim, r.data());
  std::vector<double> iv(dim * dim);
  for (int i = 0; i < dim * dim; i++) iv[i] = ((double)r[i]) / 2147483647;
  PHDE_HIP(hipMemcpyAsync(Kd.p, K, sizeof(double) * dim * dim, hipMemcpyHostToDevice, c->stream));
  PHDE_HIP(hipMemcpyAsync(init.p, iv.data(), sizeof(double) * dim * dim, hipMemcpyHostToDevice,
                          c->stream));
  PHDE_TRY(power_eig_dev(c, Kd.p, init.p, ev.p, dim, 100000, status.p));
  int hs[2] = {0, 0};
  PHDE_HIP(hipMemcpyAsync(hs, status.p, sizeof(int) * 2, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(K, Kd.p, sizeof(double) * dim * dim, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(evals, ev.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  if (iters) *iters = hs[1];
  return hs[0] == 0 ? PHDE_OK : PHDE_ERR_EIG;
}

int phde_sym_eig(phde_ctx *c, const double *A, const double *B, int dim, double *evals,
                 double *evecs) {
  CtxScope scope_(c);
  if (!c || !A || !evals || !evecs || dim < 1 || dim > 64) return PHDE_ERR_ARG;
  DBuf<double> Ad, Bd, ev, V, work;
  DBuf<int> status;
  PHDE_TRY(Ad.alloc((size_t)dim * dim));
  PHDE_TRY(ev.alloc(dim));
  PHDE_TRY(V.alloc((size_t)dim * dim));
  PHDE_TRY(work.alloc(3 * (size_t)dim * dim));
  PHDE_TRY(status.alloc(4));
  PHDE_HIP(hipMemcpyAsync(Ad.p, A, sizeof(double) * dim * dim, hipMemcpyHostToDevice, c->stream));
  if (B) {
    PHDE_TRY(Bd.alloc((size_t)dim * dim));
    PHDE_HIP(hipMemcpyAsync(Bd.p, B, sizeof(double) * dim * dim, hipMemcpyHostToDevice, c->stream));
  }
  PHDE_TRY(sym_eig_dev(c, Ad.p, B ? Bd.p : nullptr, dim, ev.p, V.p, work.p, status.p));
  int hs[2] = {0, 0};
  PHDE_HIP(hipMemcpyAsync(hs, status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(evals, ev.p, sizeof(double) * dim, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(evecs, V.p, sizeof(double) * dim * dim, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  return hs[0] == 0 ? PHDE_OK : PHDE_ERR_EIG;
}

int phde_hall_energy(phde_ctx *c, const double *Y, int nc, double *energy) {
  CtxScope scope_(c);
  if (!c || !Y || !energy || nc < 1) return PHDE_ERR_ARG;
  DBuf<double> Yd, part;
  PHDE_TRY(Yd.alloc((size_t)c->g.n * nc));
  PHDE_TRY(part.alloc(2 * (size_t)c->num_cus * 4 + 16));
  PHDE_HIP(hipMemcpyAsync(Yd.p, Y, sizeof(double) * c->g.n * nc, hipMemcpyHostToDevice, c->stream));
  DBuf<double> Yi;  // rows in the internal numbering
  if (renumbered(c)) {
    PHDE_TRY(Yi.alloc((size_t)c->g.n * nc));
    PHDE_TRY(rows_to_internal(c, Yd.p, Yi.p, nc, sizeof(double)));
  }
  double h[2];
  PHDE_TRY(hall_energy_dev(c, Yi.p ? Yi.p : Yd.p, nc, part.p, h));
  *energy = h[0] / h[1];
  return PHDE_OK;
}

}  // extern "C"
