// phde_common.h — internal declarations shared by the HIP sources of
// libparhde.so.  Not installed; the public boundary is include/parhde.h.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <map>
#include <vector>

#include "parhde.h"

namespace phde {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
#define PHDE_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::phde::set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return PHDE_ERR_DEVICE;                                                       \
    }                                                                               \
  } while (0)

constexpr int32_t kUnvisited = 0x7fffffff;  // BFS level of an unreached vertex
constexpr int kWave = 64;

// Per-level frontier counters of the persistent BFS (three rotating slots).
struct BfsSlot {
  // (frontier vertices << 32) | the sum of their degrees: one atomic adds
  // both (a frontier is a set of vertices, so the sum is at most m < 2^32
  // and never carries into the count)
  unsigned long long qs;
  unsigned int huge;  // of which degree >= kHugeDegree
  unsigned int pad[5];
};

// The exact 128-bit sum of squares (low, high word) rounded once to a
// double; equal to the reference's sequential double sum of squares
// (ParHDE.cpp:176-185) whenever that sum stays below 2^53.
inline double u128_to_double(const unsigned long long w[2]) {
  return (double)(((unsigned __int128)w[1] << 64) | (unsigned __int128)w[0]);
}

// Result of the fused k-centers epilogue of one traversal.
struct KcResult {
  unsigned long long key;        // (min_dist << 32) | (0xffffffff - vertex), maximised
  unsigned long long sumsq[2];   // sum of dist^2, exact 128-bit integer (low, high word)
  unsigned long long unreached;  // vertices left at kUnvisited
  long long depth;               // eccentricity of the source (BFS levels)
};

struct DevGraph {
  int64_t n, m;
  const uint32_t *rowptr;
  const uint32_t *adj;
  const int32_t *w;      // nullptr when unweighted
  const int32_t *deg32;  // vertex degree (count of neighbours)
  const double *degw;    // degree as used by the Laplacian (count or weight sum)
  // internal numbering (relabel.hip): caller's id of internal vertex k and
  // internal id of caller's vertex v; both nullptr when the numbers agree
  const uint32_t *orig;
  const uint32_t *perm;
};

// Caller's id of internal vertex v (tie-breaks use the caller's ids).
__device__ __forceinline__ uint32_t original_id(const DevGraph &g, uint32_t v) {
  return g.orig ? g.orig[v] : v;
}
// Internal id of caller's vertex v.
__device__ __forceinline__ uint32_t internal_id(const DevGraph &g, uint32_t v) {
  return g.perm ? g.perm[v] : v;
}

struct PartGraph;  // pbfs.hip: the parts of a deep graph

// Timing of kernel classes with HIP events on the context stream.
struct Timer {
  bool enabled = false;
  struct Span {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Span> open;
  std::vector<hipEvent_t> pool;
  double ms[16] = {0};
  long long launches[16] = {0};
  hipEvent_t get();
  void begin(hipStream_t s, int cls, hipEvent_t *a);
  void end(hipStream_t s, int cls, hipEvent_t a);
  int resolve();  // waits for recorded spans, accumulates ms
  void reset();
  ~Timer();
};

}  // namespace phde

struct phde_ctx {
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  phde::DevGraph g{};
  uint32_t *d_rowptr = nullptr, *d_adj = nullptr;
  // internal numbering (relabel.hip); empty / nullptr when not renumbered
  uint32_t *d_orig = nullptr, *d_perm = nullptr;
  std::vector<uint32_t> h_orig, h_perm;
  int32_t *d_w = nullptr, *d_deg32 = nullptr;
  double *d_degw = nullptr;
  // traversal scratch
  int32_t *d_dist = nullptr;
  uint32_t *d_q[2] = {nullptr, nullptr};
  uint32_t *d_hq[2] = {nullptr, nullptr};
  uint64_t *d_bm[2] = {nullptr, nullptr};
  phde::BfsSlot *d_slots = nullptr;
  phde::KcResult *d_kc = nullptr;
  phde::KcResult *h_kc = nullptr;  // pinned
  int bfs_grid = 0;
  bool bfs_persistent = true;
  int bar_backoff = 1;
  int bfs_stride = 1;  // persistent BFS: every bfs_stride-th block works
  int bfs_hops = 0;    // persistent BFS: hops per grid barrier (multi-hop traversal; 0: level by level)
  uint32_t bfs_qcap = 0;  // multi-hop traversal: queue entries usable (0: all n + 1)
  int bfs_hop_batch = 16;  // multi-hop traversal: queue entries a wave starts a round with
  int bfs_vpw = 1;     // narrow top-down levels: fewest queue entries per wave
  int bfs_epw = 1024;   // narrow top-down levels: edges per wave aimed at
  unsigned int *d_bar = nullptr, *h_bar = nullptr;  // grid-barrier counter + error flag
  void *h_stage = nullptr;  // pinned staging of large results copied to the caller
  size_t h_stage_bytes = 0;
  std::vector<hipEvent_t> stage_events;  // one per 8 MB chunk of a staged copy
  int32_t *d_dummy = nullptr;                       // per-thread scratch of the BFS claims
  // mapped host flags polled by the level driver (levels.h) and their
  // device-side aliases; d_done is a device-memory copy for the kernels
  int *h_flags = nullptr, *d_flags = nullptr, *d_done = nullptr;
  // generic pinned staging for small read-backs
  double *h_small = nullptr;
  // reduction scratch
  double *d_part = nullptr;
  size_t part_elems = 0;
  // delta-stepping scratch (traverse.hip)
  uint32_t *ts_far[2] = {nullptr, nullptr};
  int32_t *ts_nstamp = nullptr, *ts_fstamp = nullptr;
  unsigned int *ts_cnt = nullptr;
  int32_t default_delta = 1;
  // multi-source BFS scratch (traverse.hip)
  unsigned long long *ms_visit = nullptr, *ms_fr[3] = {nullptr, nullptr, nullptr};
  uint32_t *ms_q3 = nullptr;
  int64_t *ms_src = nullptr;
  unsigned long long *ms_dummy = nullptr;
  // SpMM hub rows (spmm.hip): rows longer than one chunk, cut into chunks;
  // sp_nhubs < 0 until built
  int64_t sp_nhubs = -1, sp_nchunks = 0;
  uint32_t *sp_hub_row = nullptr, *sp_hub_c0 = nullptr, *sp_chunk_row = nullptr, *sp_chunk_beg = nullptr;
  // learned pivot proposal (learned.hip): features (n float4), the two
  // feature traversals (2n), distance-to-pivots field, scores, training target
  float *lp_feat = nullptr;
  int32_t *lp_d12 = nullptr, *lp_sd = nullptr;
  float *lp_score = nullptr, *lp_target = nullptr;
  phde_mpnn *net = nullptr;  // borrowed (phde_set_pivot_net)
  phde::PartGraph *pg = nullptr;  // partitioned BFS (pbfs.hip); null when not in use
  uint32_t *d_esrc = nullptr;     // source vertex of each directed edge (built on first use)
  // Temporaries of a call that ran on this context, freed back here for its
  // later calls (size -> buffers): hipFree waits for the device and costs
  // ≈250 us, hipMalloc up to a few hundred; every temporary of a context is
  // used on its one stream, so stream order makes the reuse safe.
  std::multimap<size_t, void *> dev_cache, host_cache;
  size_t dev_cached = 0, host_cached = 0;
  phde::Timer timer;
};

#define PHDE_TRY_RC(x) \
  do {                 \
    int rc__ = (x);    \
    if (rc__) return rc__; \
  } while (0)

namespace phde {

// Device and pinned-host temporaries.  Inside a CtxScope (the API call
// that runs on a context), a buffer freed goes to the context's cache and
// an allocation of the same size takes it from there; outside, they are
// plain hipMalloc / hipFree.
int tmp_malloc(void **p, size_t bytes);
void tmp_free(void *p, size_t bytes);
int tmp_host_malloc(void **p, size_t bytes);
void tmp_host_free(void *p, size_t bytes);
// Makes c the calling thread's current context for the temporaries above
// (restores the previous one at the end of the scope).
struct CtxScope {
  explicit CtxScope(phde_ctx *c);
  ~CtxScope();
  phde_ctx *prev;
};
// Frees a context's cached temporaries.
void tmp_release(phde_ctx *c);

// RAII device buffer for driver temporaries.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t bytes = 0;
  int alloc(size_t count) {
    bytes = sizeof(T) * (count > 0 ? count : 1);
    return tmp_malloc(reinterpret_cast<void **>(&p), bytes);
  }
  ~DBuf() {
    if (p) tmp_free(p, bytes);
  }
};

// relabel.hip — the context's internal numbering.  relabel_graph renumbers
// the context's graph copy (at creation; sets *relabeled when it did).
// rows_to_original: out[v] = in[perm[v]]; rows_to_internal: out[k] =
// in[orig[k]]; `cols` columns of n rows, `esize`-byte elements (4, 8, 16),
// out != in (a plain copy when the numbers agree).
int relabel_graph(phde_ctx *c, bool *relabeled);
int rows_to_original(phde_ctx *c, const void *in, void *out, int cols, int esize);
int rows_to_internal(phde_ctx *c, const void *in, void *out, int cols, int esize);
inline int64_t to_internal(const phde_ctx *c, int64_t v) { return c->h_perm.empty() ? v : c->h_perm[v]; }
inline int64_t to_original(const phde_ctx *c, int64_t v) { return c->h_orig.empty() ? v : c->h_orig[v]; }
inline bool renumbered(const phde_ctx *c) { return c->d_orig != nullptr; }

// bfs.hip
int bfs_setup(phde_ctx *c);
// Called once a graph is known to have a deep BFS: its traversals then take
// the multi-hop kernel (several hops per grid barrier, see bfs.hip).
void bfs_deep_graph(phde_ctx *c);
// Runs one traversal from `src`, writing levels into dist_out (device,
// kUnvisited for unreached).  When min_dist != nullptr the k-centers update
// and argmax are fused; result in c->d_kc (device) and, after the call
// returns, c->h_kc (host, synchronised).
int bfs_run(phde_ctx *c, int64_t src, int alpha, int beta, int32_t *dist_out, int32_t *min_dist,
            bool sync_result);
// Unreached -> -1 conversion for API output.
int dist_finalize(phde_ctx *c, const int32_t *dist_in, int32_t *dist_out, int32_t unreached_val,
                  int32_t unreached_out);

// Launch of a persistent kernel (blocks that wait on one another through a
// grid barrier or a work queue).  Its grid is sized from the occupancy query
// so that every block is resident at once on an otherwise idle GPU, and
// every wait in it is bounded: a grid that other kernels keep from becoming
// resident ends with a timeout flag, and the caller repeats the work
// without a persistent kernel.  This is an ordinary launch:
// hipLaunchCooperativeKernel places the blocks no differently here, and
// after one the HIP runtime's exit handler faults inside libhsa-runtime64
// whenever the process runs under rocprofv3.
inline hipError_t launch_resident(const void *fn, int grid, int block, void **args, size_t shmem,
                                  hipStream_t stream) {
  return hipLaunchKernel(fn, dim3(grid), dim3(block), args, shmem, stream);
}

// orth.hip — D-orthogonalisation and reductions
int fill_i32(phde_ctx *c, int32_t *p, int64_t n, int32_t v);
int fill_u32(phde_ctx *c, unsigned int *p, int64_t n, unsigned int v);
int fill_f64(phde_ctx *c, double *p, int64_t n, double v);

}  // namespace phde
