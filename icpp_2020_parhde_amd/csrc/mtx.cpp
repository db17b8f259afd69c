// mtx.cpp — Matrix Market coordinate files into the binary CSR of the layout
// path (part of libphde_graph.so, host only).
//
// Follows the reference converter mtx2csr.cpp:
//  - header (lines starting "%%", :44-79): must say "coordinate"; "general"
//    marks a non-symmetric pattern; without "pattern", "real" weights are
//    read and dropped and "integer" weights are kept;
//  - size line N1 N2 M (:85-103): symmetric files need N1 == N2; a general
//    square matrix is read as the undirected graph of A + A^T; a general
//    rectangular one as the bipartite graph on N1 + N2 vertices, columns
//    first (column j -> vertex j-1, row i -> vertex N2+i-1);
//  - entries (:104-131): 1-based, both directions stored, self loops dropped
//    (non-bipartite case);
//  - each vertex's neighbours sorted by (neighbour, weight) with the signed
//    weight order, duplicates removed keeping the first, i.e. the smallest
//    weight (:134-157);
//  - largest component, relabelled in old-id order (:166-269; pg_lcc), and
//    the binary CSR with weights for integer files only (:271-296).
// Deviations, for malformed input only: blank lines are skipped, and an
// index of 0 or beyond the size line is an error (the reference asserts the
// first and does not check the second).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "phde_graph.h"

namespace {

struct Line {
  FILE *f;
  std::string s;
  bool next() {
    s.clear();
    int ch;
    bool any = false;
    while ((ch = fgetc(f)) != EOF) {
      any = true;
      if (ch == '\n') break;
      s.push_back((char)ch);
    }
    if (!s.empty() && s.back() == '\r') s.pop_back();
    return any;
  }
};

bool blank(const std::string &s) {
  for (char ch : s)
    if (ch != ' ' && ch != '\t') return false;
  return true;
}

}  // namespace

extern "C" {

int pg_read_mtx(const char *path, pg_csr *out, int *kind) {
  if (!path || !out) return PG_ERR_ARG;
  out->n = 0;
  out->m = 0;
  out->rowptr = nullptr;
  out->adj = nullptr;
  out->w = nullptr;
  FILE *f = fopen(path, "r");
  if (!f) return PG_ERR_IO;
  Line ln{f, {}};
  bool symmetric = true;
  int fmt = PG_MTX_PATTERN;
  bool have_size = false, coordinate = false, seen_header = false;
  long long n1 = 0, n2 = 0, nz = 0;
  int64_t n = 0;
  bool bipartite = false;
  std::vector<uint32_t> src, dst;
  std::vector<int32_t> wt;
  int rc = PG_OK;
  while (rc == PG_OK && ln.next()) {
    const std::string &s = ln.s;
    if (!s.empty() && s[0] == '%') {
      if (s.size() > 1 && s[1] == '%') {
        seen_header = true;
        if (s.find("coordinate") == std::string::npos) {
          rc = PG_ERR_FORMAT;
          break;
        }
        coordinate = true;
        if (s.find("general") != std::string::npos) symmetric = false;
        if (s.find("pattern") == std::string::npos) {
          if (s.find("real") != std::string::npos) fmt = PG_MTX_REAL;
          if (s.find("integer") != std::string::npos) fmt = PG_MTX_INTEGER;
        }
      }
      continue;
    }
    if (blank(s)) continue;
    if (!have_size) {
      if (sscanf(s.c_str(), "%lld %lld %lld", &n1, &n2, &nz) != 3 || n1 < 1 || n2 < 1 || nz < 0) {
        rc = PG_ERR_FORMAT;
        break;
      }
      if (symmetric) {
        if (n1 != n2) {
          rc = PG_ERR_FORMAT;
          break;
        }
        n = n1;
      } else if (n1 == n2) {
        symmetric = true;  // A + A^T
        n = n1;
      } else {
        bipartite = true;
        n = n1 + n2;
      }
      if (n > (int64_t)UINT32_MAX) {
        rc = PG_ERR_RANGE;
        break;
      }
      have_size = true;
      src.reserve((size_t)std::min<long long>(nz, 1LL << 28));
      dst.reserve((size_t)std::min<long long>(nz, 1LL << 28));
      continue;
    }
    unsigned long long u = 0, v = 0;
    long long w = 1;
    double wr = 0;
    int got;
    if (fmt == PG_MTX_PATTERN)
      got = sscanf(s.c_str(), "%llu %llu", &u, &v) == 2;
    else if (fmt == PG_MTX_REAL)
      got = sscanf(s.c_str(), "%llu %llu %lf", &u, &v, &wr) == 3;  // weight dropped
    else
      got = sscanf(s.c_str(), "%llu %llu %lld", &u, &v, &w) == 3;
    if (!got || u < 1 || v < 1 || (long long)u > n1 || (long long)v > n2 || w < INT32_MIN || w > INT32_MAX) {
      rc = PG_ERR_FORMAT;
      break;
    }
    if (bipartite) u += (unsigned long long)n2;
    else if (u == v)
      continue;
    src.push_back((uint32_t)(u - 1));
    dst.push_back((uint32_t)(v - 1));
    if (fmt == PG_MTX_INTEGER) wt.push_back((int32_t)w);
  }
  fclose(f);
  if (rc == PG_OK && (!seen_header || !coordinate || !have_size)) rc = PG_ERR_FORMAT;
  if (rc != PG_OK) return rc;

  // both directions, (neighbour << 32 | weight with its sign bit flipped):
  // the unsigned order of that key is the reference's (neighbour, signed
  // weight) pair order
  const int64_t ne = (int64_t)src.size();
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t e = 0; e < ne; e++) {
    off[src[e] + 1]++;
    off[dst[e] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) off[i + 1] += off[i];
  std::vector<uint64_t> key((size_t)off[n]);
  {
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (int64_t e = 0; e < ne; e++) {
      const uint32_t wb = (uint32_t)(wt.empty() ? 1 : wt[e]) ^ 0x80000000u;
      key[cur[src[e]]++] = ((uint64_t)dst[e] << 32) | wb;
      key[cur[dst[e]]++] = ((uint64_t)src[e] << 32) | wb;
    }
  }
  std::vector<int64_t> keep(n);
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++) {
    uint64_t *b = key.data() + off[i], *e = key.data() + off[i + 1];
    std::sort(b, e);
    int64_t k = 0;
    for (uint64_t *p = b; p < e; p++)
      if (k == 0 || (b[k - 1] >> 32) != (*p >> 32)) b[k++] = *p;
    keep[i] = k;
    m += k;
  }
  if (m > (int64_t)UINT32_MAX) return PG_ERR_RANGE;
  out->rowptr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1));
  out->adj = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)std::max<int64_t>(m, 1));
  if (fmt == PG_MTX_INTEGER) out->w = (int32_t *)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  if (!out->rowptr || !out->adj || (fmt == PG_MTX_INTEGER && !out->w)) {
    pg_free(out);
    return PG_ERR_NOMEM;
  }
  out->n = n;
  out->m = m;
  out->rowptr[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    const uint64_t *b = key.data() + off[i];
    const uint32_t r = out->rowptr[i];
    for (int64_t k = 0; k < keep[i]; k++) {
      out->adj[r + k] = (uint32_t)(b[k] >> 32);
      if (out->w) out->w[r + k] = (int32_t)((uint32_t)b[k] ^ 0x80000000u);
    }
    out->rowptr[i + 1] = r + (uint32_t)keep[i];
  }
  if (kind) *kind = fmt;
  return PG_OK;
}

int pg_mtx2csr(const char *mtx_path, const char *csr_path, int64_t *n, int64_t *m) {
  if (!mtx_path || !csr_path) return PG_ERR_ARG;
  pg_csr full{}, lcc{};
  int rc = pg_read_mtx(mtx_path, &full, nullptr);
  if (rc) return rc;
  rc = pg_lcc(&full, &lcc, nullptr);
  pg_free(&full);
  if (rc) return rc;
  rc = pg_write_csr(csr_path, &lcc);
  if (rc == PG_OK) {
    if (n) *n = lcc.n;
    if (m) *m = lcc.m;
  }
  pg_free(&lcc);
  return rc;
}

}  // extern "C"
