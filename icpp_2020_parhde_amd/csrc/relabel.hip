// relabel.hip — the context's internal vertex numbering.
//
// On a high-diameter graph a level-synchronous traversal is bound by the
// latency of each level's dependent chain (queue entry -> row -> neighbour
// level -> claim -> append).  Numbering the vertices so that a level's
// vertices and their neighbours sit close together in memory shortens that
// chain: fewer distinct cache lines per wave, fewer transactions in flight.
// The context therefore renumbers its copy of the graph once, at creation,
// in the order of (level from vertex 0, level from the vertex farthest from
// 0, caller's id) — a stable sort of two BFS level arrays — and keeps both
// maps:
//   orig[k] : caller's id of internal vertex k   (DevGraph::orig)
//   perm[v] : internal id of caller's vertex v   (DevGraph::perm)
// Each row keeps the caller's neighbour order (ids mapped), so per-row sums
// (degrees, Laplacian rows, the pivot network's neighbour means) are the
// same numbers as without renumbering.  Everything a caller passes in or
// gets back (sources, distance columns, coordinates, features, scores,
// pivots) is in the caller's numbering; ties between vertices (the
// k-centers argmax, the pivot selection) are broken on the caller's ids.
//
// Measured on MI355X, k-centers layouts: RGG n=4.2M s=50 1,325 -> 1,220 ms
// (BFS 25.2 -> 23.0 ms per traversal), RGG n=16.8M s=50 3,084 -> 2,717 ms,
// weighted grid 160^3 s=20 264 -> 254 ms; but a renumbered copy of RMAT
// scale 22 went 207 -> 226 ms (a low-diameter graph's levels are wide and
// the renumbering scatters its hubs).  So by default only graphs whose BFS
// from vertex 0 has at least kMinDepth levels are renumbered;
// PHDE_RELABEL=0 / =1 turns it off / on for every graph.  Sorting each
// renumbered row by neighbour id measured no faster (23.01 vs 22.96 ms).
#include <hipcub/hipcub.hpp>

#include <cstdlib>

#include "phde_common.h"

namespace phde {

namespace {

constexpr int64_t kMinVertices = 1 << 16;
constexpr long long kMinDepth = 64;

__global__ void relabel_keys_kernel(const int32_t *d1, const int32_t *d2, int64_t n,
                                    unsigned long long *keys, uint32_t *vals) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    keys[i] = ((unsigned long long)(uint32_t)d1[i] << 32) | (uint32_t)d2[i];
    vals[i] = (uint32_t)i;
  }
}

__global__ void relabel_inverse_kernel(const uint32_t *orig, int64_t n, uint32_t *perm) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    perm[orig[k]] = (uint32_t)k;
}

// deg[k] = degree of the caller's vertex orig[k]; deg[n] = 0 (for the scan)
__global__ void relabel_degree_kernel(const uint32_t *rowptr, const uint32_t *orig, int64_t n,
                                      uint32_t *deg) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= n;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k == n) {
      deg[n] = 0;
    } else {
      const uint32_t o = orig[k];
      deg[k] = rowptr[o + 1] - rowptr[o];
    }
  }
}

// One wave per internal row: the caller's row orig[k], ids mapped by perm,
// neighbour order kept.
__global__ __launch_bounds__(256) void relabel_rows_kernel(const uint32_t *rowptr, const uint32_t *adj,
                                                           const int32_t *w, const uint32_t *rp_new,
                                                           const uint32_t *orig, const uint32_t *perm,
                                                           int64_t n, uint32_t *adj_new, int32_t *w_new) {
  const int lane = threadIdx.x % kWave;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t k = wave; k < n; k += nwaves) {
    const uint32_t o = orig[k];
    const uint32_t b = rowptr[o], e = rowptr[o + 1], nb = rp_new[k];
    for (uint32_t j = b + lane; j < e; j += kWave) {
      adj_new[nb + (j - b)] = perm[adj[j]];
      if (w) w_new[nb + (j - b)] = w[j];
    }
  }
}

// out[c * n + i] = in[c * n + idx[i]]
template <typename T>
__global__ void gather_rows_kernel(const T *in, const uint32_t *idx, int64_t n, int cols, T *out) {
  const int64_t total = n * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = e / n, i = e - col * n;
    out[e] = in[col * n + idx[i]];
  }
}

int grid_for(const phde_ctx *c, int64_t work) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, (int64_t)c->num_cus * 16));
}

int gather_rows(phde_ctx *c, const void *in, void *out, int cols, int esize, const uint32_t *idx) {
  const int64_t n = c->g.n;
  if (!idx) {
    if (in != out)
      PHDE_HIP(hipMemcpyAsync(out, in, (size_t)esize * n * cols, hipMemcpyDeviceToDevice, c->stream));
    return PHDE_OK;
  }
  if (in == out) {
    set_error("internal: renumbering gather in place");
    return PHDE_ERR_ARG;
  }
  const int nb = grid_for(c, n * cols);
  switch (esize) {
    case 4:
      gather_rows_kernel<<<nb, 256, 0, c->stream>>>((const uint32_t *)in, idx, n, cols, (uint32_t *)out);
      break;
    case 8:
      gather_rows_kernel<<<nb, 256, 0, c->stream>>>((const unsigned long long *)in, idx, n, cols,
                                                    (unsigned long long *)out);
      break;
    case 16:
      gather_rows_kernel<<<nb, 256, 0, c->stream>>>((const uint4 *)in, idx, n, cols, (uint4 *)out);
      break;
    default:
      set_error("internal: unsupported element size");
      return PHDE_ERR_ARG;
  }
  PHDE_HIP(hipGetLastError());
  return PHDE_OK;
}

}  // namespace

int rows_to_original(phde_ctx *c, const void *in, void *out, int cols, int esize) {
  return gather_rows(c, in, out, cols, esize, c->d_perm);
}

int rows_to_internal(phde_ctx *c, const void *in, void *out, int cols, int esize) {
  return gather_rows(c, in, out, cols, esize, c->d_orig);
}

int relabel_graph(phde_ctx *c, bool *relabeled) {
  *relabeled = false;
  const int64_t n = c->g.n, m = c->g.m;
  int mode = -1;  // auto
  if (const char *e = getenv("PHDE_RELABEL")) mode = atoi(e) ? 1 : 0;
  if (mode == 0 || n < 2 || n > INT32_MAX || (mode < 0 && n < kMinVertices)) return PHDE_OK;
  PHDE_TRY_RC(bfs_setup(c));
  DBuf<int32_t> d1, d2, md;
  PHDE_TRY_RC(d1.alloc(n));
  PHDE_TRY_RC(d2.alloc(n));
  PHDE_TRY_RC(md.alloc(n));
  PHDE_TRY_RC(fill_i32(c, md.p, n, INT32_MAX));
  PHDE_TRY_RC(bfs_run(c, 0, 15, 18, d1.p, md.p, true));
  const KcResult r1 = *c->h_kc;
  if (r1.depth >= kMinDepth) bfs_deep_graph(c);
  if (mode < 0 && r1.depth < kMinDepth) return PHDE_OK;
  // the farthest vertex from 0 (lowest id on ties; an unreached one when
  // the graph is not connected — any order is valid, this one is still
  // deterministic)
  const int64_t far = (int64_t)(0xffffffffull - (r1.key & 0xffffffffull));
  PHDE_TRY_RC(bfs_run(c, far, 15, 18, d2.p, md.p, true));

  DBuf<unsigned long long> k0, k1;
  DBuf<uint32_t> v0, deg, rp;
  PHDE_TRY_RC(k0.alloc(n));
  PHDE_TRY_RC(k1.alloc(n));
  PHDE_TRY_RC(v0.alloc(n));
  PHDE_TRY_RC(deg.alloc(n + 1));
  PHDE_TRY_RC(rp.alloc(n + 1));
  uint32_t *orig = nullptr, *perm = nullptr, *adj_new = nullptr;
  int32_t *w_new = nullptr;
  PHDE_HIP(hipMalloc(&orig, sizeof(uint32_t) * n));
  c->d_orig = orig;
  PHDE_HIP(hipMalloc(&perm, sizeof(uint32_t) * n));
  c->d_perm = perm;
  const int nb = grid_for(c, n);
  relabel_keys_kernel<<<nb, 256, 0, c->stream>>>(d1.p, d2.p, n, k0.p, v0.p);
  PHDE_HIP(hipGetLastError());
  // stable radix sort: ties keep the caller's order
  size_t tmp_bytes = 0;
  PHDE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, v0.p, orig, (int)n, 0, 64,
                                              c->stream));
  {
    DBuf<unsigned char> tmp;
    PHDE_TRY_RC(tmp.alloc(tmp_bytes));
    PHDE_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, k0.p, k1.p, v0.p, orig, (int)n, 0, 64,
                                                c->stream));
    relabel_inverse_kernel<<<nb, 256, 0, c->stream>>>(orig, n, perm);
    relabel_degree_kernel<<<nb, 256, 0, c->stream>>>(c->d_rowptr, orig, n, deg.p);
    PHDE_HIP(hipGetLastError());
    size_t scan_bytes = 0;
    PHDE_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, deg.p, rp.p, (int)(n + 1), c->stream));
    DBuf<unsigned char> stmp;
    PHDE_TRY_RC(stmp.alloc(scan_bytes));
    PHDE_HIP(hipcub::DeviceScan::ExclusiveSum(stmp.p, scan_bytes, deg.p, rp.p, (int)(n + 1), c->stream));
    PHDE_HIP(hipStreamSynchronize(c->stream));
  }
  PHDE_HIP(hipMalloc(&adj_new, sizeof(uint32_t) * std::max<int64_t>(m, 1)));
  if (c->d_w && hipMalloc(&w_new, sizeof(int32_t) * std::max<int64_t>(m, 1)) != hipSuccess) {
    (void)hipFree(adj_new);
    set_error("out of device memory renumbering the graph");
    return PHDE_ERR_DEVICE;
  }
  relabel_rows_kernel<<<grid_for(c, n * kWave), 256, 0, c->stream>>>(c->d_rowptr, c->d_adj, c->d_w, rp.p,
                                                                       orig, perm, n, adj_new, w_new);
  const hipError_t le = hipGetLastError();
  const hipError_t se = hipStreamSynchronize(c->stream);
  if (le != hipSuccess || se != hipSuccess) {
    (void)hipFree(adj_new);
    if (w_new) (void)hipFree(w_new);
    set_error(std::string("renumbering the graph: ") + hipGetErrorString(le != hipSuccess ? le : se));
    return PHDE_ERR_DEVICE;
  }
  // the renumbered CSR replaces the caller-order copy
  PHDE_HIP(hipMemcpyAsync(c->d_rowptr, rp.p, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToDevice, c->stream));
  (void)hipFree(c->d_adj);
  c->d_adj = adj_new;
  if (c->d_w) {
    (void)hipFree(c->d_w);
    c->d_w = w_new;
  }
  c->g.adj = c->d_adj;
  c->g.w = c->d_w;
  c->g.orig = orig;
  c->g.perm = perm;
  c->h_orig.resize(n);
  c->h_perm.resize(n);
  PHDE_HIP(hipMemcpyAsync(c->h_orig.data(), orig, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipMemcpyAsync(c->h_perm.data(), perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
  PHDE_HIP(hipStreamSynchronize(c->stream));
  *relabeled = true;
  return PHDE_OK;
}

}  // namespace phde
