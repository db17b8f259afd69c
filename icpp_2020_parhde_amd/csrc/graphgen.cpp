// graphgen.cpp — seeded synthetic graphs, largest-component extraction and
// binary CSR I/O for the ParHDE layout path (host code, no GPU).
//
// The binary CSR format is the one spectralDrawing.cpp:136-164 reads and
// mtx2csr.cpp:281-296 writes; component extraction follows the relabelling
// rule of mtx2csr.cpp:166-269 (largest component, first found on ties, new
// ids assigned in increasing old id).  The generators themselves have no
// counterpart in the reference: vertex numbering decides k-centers ties, so
// every generator below is a pure function of its arguments, independent of
// the number of OpenMP threads.
#include "phde_graph.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

inline uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t mix64(uint64_t x) {
  uint64_t s = x;
  return splitmix64(s);
}

struct Xoshiro256ss {
  uint64_t s[4];
  explicit Xoshiro256ss(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) s[i] = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  inline uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // uniform double in [0, 1) with 53 random bits
  inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

int alloc_csr(int64_t n, int64_t m, bool weighted, pg_csr *g) {
  g->n = n;
  g->m = m;
  g->rowptr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
  g->adj = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m > 0 ? m : 1));
  g->w = weighted ? (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1)) : nullptr;
  if (!g->rowptr || !g->adj || (weighted && !g->w)) {
    pg_free(g);
    return PG_ERR_NOMEM;
  }
  return PG_OK;
}

// Prefix sum of per-vertex counts into a uint32 row-pointer array.
int counts_to_rowptr(const std::vector<int64_t> &cnt, int64_t n, uint32_t *rowptr, int64_t *m_out) {
  int64_t acc = 0;
  rowptr[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    acc += cnt[i];
    if (acc > (int64_t)UINT32_MAX) return PG_ERR_RANGE;
    rowptr[i + 1] = (uint32_t)acc;
  }
  *m_out = acc;
  return PG_OK;
}

}  // namespace

extern "C" {

void pg_free(pg_csr *g) {
  if (!g) return;
  free(g->rowptr);
  free(g->adj);
  free(g->w);
  g->rowptr = nullptr;
  g->adj = nullptr;
  g->w = nullptr;
  g->n = g->m = 0;
}

int pg_mesh(int64_t rows, int64_t cols, pg_csr *out) {
  if (!out || rows < 1 || cols < 1) return PG_ERR_ARG;
  const int64_t n = rows * cols;
  std::vector<int64_t> cnt(n);
  auto deg = [&](int64_t i, int64_t j) {
    int d = 0;
    if (i > 0 && j > 0) d++;
    if (i > 0) d++;
    if (j > 0) d++;
    if (j + 1 < cols) d++;
    if (i + 1 < rows) d++;
    if (i + 1 < rows && j + 1 < cols) d++;
    return d;
  };
  for (int64_t i = 0; i < rows; i++)
    for (int64_t j = 0; j < cols; j++) cnt[i * cols + j] = deg(i, j);
  std::vector<uint32_t> rp(n + 1);
  int64_t m;
  int rc = counts_to_rowptr(cnt, n, rp.data(), &m);
  if (rc) return rc;
  rc = alloc_csr(n, m, false, out);
  if (rc) return rc;
  memcpy(out->rowptr, rp.data(), sizeof(uint32_t) * (n + 1));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; i++) {
    for (int64_t j = 0; j < cols; j++) {
      const int64_t id = i * cols + j;
      uint32_t *p = out->adj + out->rowptr[id];
      if (i > 0 && j > 0) *p++ = (uint32_t)(id - cols - 1);
      if (i > 0) *p++ = (uint32_t)(id - cols);
      if (j > 0) *p++ = (uint32_t)(id - 1);
      if (j + 1 < cols) *p++ = (uint32_t)(id + 1);
      if (i + 1 < rows) *p++ = (uint32_t)(id + cols);
      if (i + 1 < rows && j + 1 < cols) *p++ = (uint32_t)(id + cols + 1);
    }
  }
  return PG_OK;
}

int pg_grid3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, int32_t wmax, pg_csr *out) {
  if (!out || nx < 1 || ny < 1 || nz < 1) return PG_ERR_ARG;
  const int64_t n = nx * ny * nz;
  std::vector<int64_t> cnt(n);
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < n; id++) {
    const int64_t x = id / (ny * nz), y = (id / nz) % ny, z = id % nz;
    cnt[id] = (x > 0) + (y > 0) + (z > 0) + (z + 1 < nz) + (y + 1 < ny) + (x + 1 < nx);
  }
  std::vector<uint32_t> rp(n + 1);
  int64_t m;
  int rc = counts_to_rowptr(cnt, n, rp.data(), &m);
  if (rc) return rc;
  rc = alloc_csr(n, m, wmax > 0, out);
  if (rc) return rc;
  memcpy(out->rowptr, rp.data(), sizeof(uint32_t) * (n + 1));
  const int64_t sx = ny * nz, sy = nz;
  auto weight = [&](int64_t a, int64_t b) -> int32_t {
    const uint64_t lo = (uint64_t)std::min(a, b), hi = (uint64_t)std::max(a, b);
    const uint64_t h = mix64(seed ^ mix64(lo * 0x9E3779B97F4A7C15ull + hi));
    return (int32_t)(1 + h % (uint64_t)wmax);
  };
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < n; id++) {
    const int64_t x = id / sx, y = (id / sy) % ny, z = id % nz;
    int64_t nb[6];
    int k = 0;
    if (x > 0) nb[k++] = id - sx;
    if (y > 0) nb[k++] = id - sy;
    if (z > 0) nb[k++] = id - 1;
    if (z + 1 < nz) nb[k++] = id + 1;
    if (y + 1 < ny) nb[k++] = id + sy;
    if (x + 1 < nx) nb[k++] = id + sx;
    const int64_t base = out->rowptr[id];
    for (int t = 0; t < k; t++) {
      out->adj[base + t] = (uint32_t)nb[t];
      if (out->w) out->w[base + t] = weight(id, nb[t]);
    }
  }
  return PG_OK;
}

int pg_from_edges(int64_t n, int64_t ne, const uint32_t *u, const uint32_t *v, const int32_t *wt,
                  pg_csr *out) {
  if (!out || n < 1 || ne < 0 || (ne > 0 && (!u || !v))) return PG_ERR_ARG;
  for (int64_t e = 0; e < ne; e++)
    if ((int64_t)u[e] >= n || (int64_t)v[e] >= n) return PG_ERR_ARG;
  // Directed copies in both directions, then per-row sort + dedupe.
  std::vector<std::atomic<int64_t>> cnt(n);
  for (int64_t i = 0; i < n; i++) cnt[i].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; e++) {
    if (u[e] == v[e]) continue;
    cnt[u[e]].fetch_add(1, std::memory_order_relaxed);
    cnt[v[e]].fetch_add(1, std::memory_order_relaxed);
  }
  std::vector<int64_t> off(n + 1);
  off[0] = 0;
  for (int64_t i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i].load(std::memory_order_relaxed);
  const int64_t md = off[n];
  std::vector<uint64_t> tmp(md);  // (neighbour << 32) | (weight as uint32)
  std::vector<std::atomic<int64_t>> cur(n);
  for (int64_t i = 0; i < n; i++) cur[i].store(off[i], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < ne; e++) {
    if (u[e] == v[e]) continue;
    const uint32_t w = wt ? (uint32_t)wt[e] : 0u;
    int64_t p = cur[u[e]].fetch_add(1, std::memory_order_relaxed);
    tmp[p] = ((uint64_t)v[e] << 32) | w;
    p = cur[v[e]].fetch_add(1, std::memory_order_relaxed);
    tmp[p] = ((uint64_t)u[e] << 32) | w;
  }
  std::vector<int64_t> ucnt(n);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; i++) {
    uint64_t *b = tmp.data() + off[i], *e = tmp.data() + off[i + 1];
    std::sort(b, e);
    int64_t k = 0;
    for (uint64_t *p = b; p < e; p++)
      if (k == 0 || (b[k - 1] >> 32) != (*p >> 32)) b[k++] = *p;
    ucnt[i] = k;
  }
  std::vector<uint32_t> rp(n + 1);
  int64_t m;
  int rc = counts_to_rowptr(ucnt, n, rp.data(), &m);
  if (rc) return rc;
  rc = alloc_csr(n, m, wt != nullptr, out);
  if (rc) return rc;
  memcpy(out->rowptr, rp.data(), sizeof(uint32_t) * (n + 1));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const uint64_t *b = tmp.data() + off[i];
    for (int64_t k = 0; k < ucnt[i]; k++) {
      out->adj[rp[i] + k] = (uint32_t)(b[k] >> 32);
      if (out->w) out->w[rp[i] + k] = (int32_t)(uint32_t)(b[k] & 0xffffffffu);
    }
  }
  return PG_OK;
}

}  // extern "C"

namespace {
int lcc_impl(const pg_csr *in, pg_csr *out, int64_t *n_components, std::vector<uint32_t> *kept) {
  if (!in || !out || in->n < 1) return PG_ERR_ARG;
  const int64_t n = in->n;
  std::vector<int64_t> comp(n, 0);
  std::vector<uint32_t> queue(n);
  int64_t ncomp = 0, best = 0, best_size = 0;
  for (int64_t i = 0; i < n; i++) {
    if (comp[i]) continue;
    ncomp++;
    comp[i] = ncomp;
    int64_t head = 0, tail = 0;
    queue[tail++] = (uint32_t)i;
    while (head < tail) {
      const uint32_t x = queue[head++];
      for (uint32_t e = in->rowptr[x]; e < in->rowptr[x + 1]; e++) {
        const uint32_t y = in->adj[e];
        if (!comp[y]) {
          comp[y] = ncomp;
          queue[tail++] = y;
        }
      }
    }
    if (tail > best_size) {
      best_size = tail;
      best = ncomp;
    }
  }
  if (n_components) *n_components = ncomp;
  std::vector<int64_t> newid(n, -1);
  int64_t nn = 0, mm = 0;
  for (int64_t i = 0; i < n; i++)
    if (comp[i] == best) {
      newid[i] = nn++;
      mm += in->rowptr[i + 1] - in->rowptr[i];
      if (kept) kept->push_back((uint32_t)i);
    }
  int rc = alloc_csr(nn, mm, in->w != nullptr, out);
  if (rc) return rc;
  out->rowptr[0] = 0;
  for (int64_t i = 0, k = 0; i < n; i++)
    if (newid[i] >= 0) {
      out->rowptr[k + 1] = out->rowptr[k] + (in->rowptr[i + 1] - in->rowptr[i]);
      k++;
    }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    if (newid[i] < 0) continue;
    const int64_t dst = out->rowptr[newid[i]];
    for (uint32_t e = in->rowptr[i], k = 0; e < in->rowptr[i + 1]; e++, k++) {
      out->adj[dst + k] = (uint32_t)newid[in->adj[e]];
      if (out->w) out->w[dst + k] = in->w[e];
    }
  }
  return PG_OK;
}
}  // namespace

extern "C" {

int pg_lcc(const pg_csr *in, pg_csr *out, int64_t *n_components) {
  return lcc_impl(in, out, n_components, nullptr);
}

int pg_rgg(int64_t n, double avg_degree, uint64_t seed, int lcc, pg_csr *out, double **xy_out) {
  if (!out || n < 2 || !(avg_degree > 0)) return PG_ERR_ARG;
  const double r = std::sqrt(avg_degree / (M_PI * (double)n));
  const double r2 = r * r;
  std::vector<double> xy(2 * n);
  {
    Xoshiro256ss rng(seed);
    for (int64_t i = 0; i < n; i++) {
      xy[2 * i] = rng.uniform();
      xy[2 * i + 1] = rng.uniform();
    }
  }
  int64_t G = (int64_t)std::floor(1.0 / r);
  if (G < 1) G = 1;
  if (G > 65536) G = 65536;
  auto cell_of = [&](int64_t i) -> int64_t {
    int64_t cx = (int64_t)(xy[2 * i] * G), cy = (int64_t)(xy[2 * i + 1] * G);
    if (cx >= G) cx = G - 1;
    if (cy >= G) cy = G - 1;
    return cx * G + cy;
  };
  const int64_t ncell = G * G;
  std::vector<int64_t> cstart(ncell + 1, 0);
  std::vector<int64_t> cell(n);
  for (int64_t i = 0; i < n; i++) {
    cell[i] = cell_of(i);
    cstart[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  std::vector<uint32_t> members(n);
  {
    std::vector<int64_t> fill(cstart.begin(), cstart.end() - 1);
    for (int64_t i = 0; i < n; i++) members[fill[cell[i]]++] = (uint32_t)i;  // ascending ids per cell
  }
  auto for_neighbours = [&](int64_t i, auto &&fn) {
    const int64_t cx = cell[i] / G, cy = cell[i] % G;
    const double x = xy[2 * i], y = xy[2 * i + 1];
    for (int64_t dx = -1; dx <= 1; dx++) {
      const int64_t ax = cx + dx;
      if (ax < 0 || ax >= G) continue;
      for (int64_t dy = -1; dy <= 1; dy++) {
        const int64_t ay = cy + dy;
        if (ay < 0 || ay >= G) continue;
        const int64_t c = ax * G + ay;
        for (int64_t k = cstart[c]; k < cstart[c + 1]; k++) {
          const uint32_t j = members[k];
          if ((int64_t)j == i) continue;
          const double ddx = xy[2 * j] - x, ddy = xy[2 * j + 1] - y;
          if (ddx * ddx + ddy * ddy < r2) fn(j);
        }
      }
    }
  };
  std::vector<int64_t> cnt(n);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    int64_t c = 0;
    for_neighbours(i, [&](uint32_t) { c++; });
    cnt[i] = c;
  }
  pg_csr full{};
  std::vector<uint32_t> rp(n + 1);
  int64_t m;
  int rc = counts_to_rowptr(cnt, n, rp.data(), &m);
  if (rc) return rc;
  rc = alloc_csr(n, m, false, &full);
  if (rc) return rc;
  memcpy(full.rowptr, rp.data(), sizeof(uint32_t) * (n + 1));
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    uint32_t *p = full.adj + full.rowptr[i];
    uint32_t *b = p;
    for_neighbours(i, [&](uint32_t j) { *p++ = j; });
    std::sort(b, p);
  }
  if (!lcc) {
    *out = full;
    if (xy_out) {
      *xy_out = (double *)malloc(sizeof(double) * 2 * n);
      if (*xy_out) memcpy(*xy_out, xy.data(), sizeof(double) * 2 * n);
    }
    return PG_OK;
  }
  std::vector<uint32_t> kept;
  rc = lcc_impl(&full, out, nullptr, xy_out ? &kept : nullptr);
  if (rc == PG_OK && xy_out) {
    *xy_out = (double *)malloc(sizeof(double) * 2 * out->n);
    if (*xy_out)
      for (int64_t k = 0; k < out->n; k++) {
        (*xy_out)[2 * k] = xy[2 * kept[k]];
        (*xy_out)[2 * k + 1] = xy[2 * kept[k] + 1];
      }
  }
  pg_free(&full);
  return rc;
}

int pg_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed, int lcc,
            pg_csr *out) {
  if (!out || scale < 1 || scale > 31 || edgefactor < 1 || a < 0 || b < 0 || c < 0 ||
      a + b + c > 1.0)
    return PG_ERR_ARG;
  const int64_t n = (int64_t)1 << scale;
  const int64_t ne = (int64_t)edgefactor * n;
  std::vector<uint32_t> u(ne), v(ne);
  const int64_t block = 4096;
  const int64_t nblocks = (ne + block - 1) / block;
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t blk = 0; blk < nblocks; blk++) {
    Xoshiro256ss rng(mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)blk));
    const int64_t e0 = blk * block, e1 = std::min(ne, e0 + block);
    for (int64_t e = e0; e < e1; e++) {
      uint32_t uu = 0, vv = 0;
      for (int l = 0; l < scale; l++) {
        const double r = rng.uniform();
        uint32_t bu, bv;
        if (r < a) {
          bu = 0; bv = 0;
        } else if (r < ab) {
          bu = 0; bv = 1;
        } else if (r < abc) {
          bu = 1; bv = 0;
        } else {
          bu = 1; bv = 1;
        }
        uu = (uu << 1) | bu;
        vv = (vv << 1) | bv;
      }
      u[e] = uu;
      v[e] = vv;
    }
  }
  pg_csr full{};
  int rc = pg_from_edges(n, ne, u.data(), v.data(), nullptr, &full);
  std::vector<uint32_t>().swap(u);
  std::vector<uint32_t>().swap(v);
  if (rc) return rc;
  if (!lcc) {
    *out = full;
    return PG_OK;
  }
  rc = pg_lcc(&full, out, nullptr);
  pg_free(&full);
  return rc;
}

int pg_holed_grid(int64_t rows, int64_t cols, int holes, uint64_t seed, pg_csr *out) {
  if (!out || rows < 2 || cols < 2 || holes < 0) return PG_ERR_ARG;
  const int64_t n = rows * cols;
  std::vector<char> gone(n, 0);
  Xoshiro256ss rng(seed);
  for (int h = 0; h < holes; h++) {
    const int64_t hr = 1 + (int64_t)(rng.uniform() * (double)std::max<int64_t>(1, rows / 6));
    const int64_t hc = 1 + (int64_t)(rng.uniform() * (double)std::max<int64_t>(1, cols / 6));
    const int64_t r0 = (int64_t)(rng.uniform() * (double)(rows - hr));
    const int64_t c0 = (int64_t)(rng.uniform() * (double)(cols - hc));
    for (int64_t i = r0; i < r0 + hr; i++)
      for (int64_t j = c0; j < c0 + hc; j++) gone[i * cols + j] = 1;
  }
  std::vector<uint32_t> eu, ev;
  for (int64_t i = 0; i < rows; i++)
    for (int64_t j = 0; j < cols; j++) {
      const int64_t id = i * cols + j;
      if (gone[id]) continue;
      if (j + 1 < cols && !gone[id + 1]) {
        eu.push_back((uint32_t)id);
        ev.push_back((uint32_t)(id + 1));
      }
      if (i + 1 < rows && !gone[id + cols]) {
        eu.push_back((uint32_t)id);
        ev.push_back((uint32_t)(id + cols));
      }
    }
  pg_csr full{};
  int rc = pg_from_edges(n, (int64_t)eu.size(), eu.data(), ev.data(), nullptr, &full);
  if (rc) return rc;
  rc = pg_lcc(&full, out, nullptr);
  pg_free(&full);
  return rc;
}

int pg_write_csr(const char *path, const pg_csr *g) {
  if (!path || !g || !g->rowptr || (g->m > 0 && !g->adj)) return PG_ERR_ARG;
  FILE *f = fopen(path, "wb");
  if (!f) return PG_ERR_IO;
  const int64_t hdr[6] = {g->n, g->m, 1, 0, 1, 0};
  bool ok = fwrite(hdr, sizeof(int64_t), 6, f) == 6;
  ok = ok && fwrite(g->rowptr, sizeof(uint32_t), (size_t)(g->n + 1), f) == (size_t)(g->n + 1);
  ok = ok && fwrite(g->adj, sizeof(uint32_t), (size_t)g->m, f) == (size_t)g->m;
  if (g->w) ok = ok && fwrite(g->w, sizeof(int32_t), (size_t)g->m, f) == (size_t)g->m;
  ok = (fclose(f) == 0) && ok;
  return ok ? PG_OK : PG_ERR_IO;
}

int pg_read_csr(const char *path, int weighted, pg_csr *out) {
  if (!path || !out) return PG_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return PG_ERR_IO;
  int64_t hdr[6];
  if (fread(hdr, sizeof(int64_t), 6, f) != 6 || hdr[0] < 1 || hdr[1] < 0 ||
      hdr[1] > (int64_t)UINT32_MAX) {
    fclose(f);
    return PG_ERR_FORMAT;
  }
  int rc = alloc_csr(hdr[0], hdr[1], weighted != 0, out);
  if (rc) {
    fclose(f);
    return rc;
  }
  bool ok = fread(out->rowptr, sizeof(uint32_t), (size_t)(hdr[0] + 1), f) == (size_t)(hdr[0] + 1);
  ok = ok && fread(out->adj, sizeof(uint32_t), (size_t)hdr[1], f) == (size_t)hdr[1];
  if (weighted) ok = ok && fread(out->w, sizeof(int32_t), (size_t)hdr[1], f) == (size_t)hdr[1];
  fclose(f);
  if (!ok || out->rowptr[0] != 0 || (int64_t)out->rowptr[hdr[0]] != hdr[1]) {
    pg_free(out);
    return PG_ERR_FORMAT;
  }
  return PG_OK;
}

int pg_validate(const pg_csr *g) {
  if (!g || g->n < 1 || !g->rowptr) return PG_ERR_ARG;
  const int64_t n = g->n;
  if (g->rowptr[0] != 0 || (int64_t)g->rowptr[n] != g->m) return PG_ERR_FORMAT;
  std::atomic<int> bad(0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t i = 0; i < n; i++) {
    if (g->rowptr[i + 1] < g->rowptr[i]) {
      bad = 1;
      continue;
    }
    for (uint32_t e = g->rowptr[i]; e < g->rowptr[i + 1]; e++) {
      const uint32_t j = g->adj[e];
      if ((int64_t)j >= n || (int64_t)j == i || (e > g->rowptr[i] && g->adj[e - 1] >= j)) {
        bad = 1;
        break;
      }
      // symmetric partner must exist (binary search in j's sorted list)
      const uint32_t *b = g->adj + g->rowptr[j], *en = g->adj + g->rowptr[j + 1];
      const uint32_t *p = std::lower_bound(b, en, (uint32_t)i);
      if (p == en || *p != (uint32_t)i) {
        bad = 1;
        break;
      }
      if (g->w && g->w[p - g->adj] != g->w[e]) {
        bad = 1;
        break;
      }
    }
  }
  return bad ? PG_ERR_FORMAT : PG_OK;
}

}  // extern "C"
