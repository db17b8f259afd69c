// oracle.cpp — CPU/OpenMP restatement of the reference ParHDE layout path.
//
// TEST INFRASTRUCTURE ONLY: linked by nothing in the product.  tests/ use it
// as the parity checker, __graft_entry__.smoke() as the checker of one small
// layout, bench.py as the timed CPU baseline ("port").
//
// Each routine follows the reference routine cited beside it, keeping its
// operation order (sequential sums where the reference sums with one
// thread, lowest-index argmax ties, the CGS/MGS update order, the power
// method's iteration and sign rule).  Where the reference's result depends
// on the OpenMP schedule (reduction order with >1 thread, the argmax tie
// combine, the `omp critical` CGS combine) this restatement fixes the
// single-thread semantics: ties go to the lowest vertex index and per-thread
// partial sums are combined in thread order.
//
// What pins each function (tests/test_oracle.py):
//   or_seq_bfs, or_dobfs  closed-form distances on meshes, grids, paths,
//                         cycles and stars; equality with each other and
//                         with scipy's shortest_path on random graphs.
//   or_deltastep          scipy Dijkstra on random weighted graphs; or_dijkstra;
//                         independence from delta.
//   or_power_eig          the reference's own ParallelComponents.cpp compiled
//                         from source into oracle/_ref (bit-exact), and
//                         fixtures written from it into tests/golden for
//                         machines without the reference tree.
//   or_parhde             the reference's pivot sequence and discard on the
//                         128x128 mesh (captured from a run of the compiled
//                         reference, tests/golden), D-orthogonality
//                         invariants, the closed-form projection for CGS,
//                         dense Laplacian algebra for L*S and S^T(LS),
//                         brute-force farthest-first pivots on small graphs.
//   or_wparhde            the same invariants with MGS; Dijkstra columns.
//   or_pivotmds, or_phde  dense numpy restatements of the centring formulas
//                         and eigen-decomposition of C^T C.
//   or_hall_energy        closed forms on paths and cycles.
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <limits>
#include <queue>
#include <set>
#include <vector>
#include <omp.h>

namespace {

using clk = std::chrono::high_resolution_clock;
inline double secs(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

inline int64_t degree(const or_graph *g, int64_t v) { return g->rowptr[v + 1] - g->rowptr[v]; }

// Bit set with the operations the direction-optimising BFS uses.
struct Bits {
  std::vector<uint64_t> w;
  explicit Bits(int64_t n) : w((n + 63) / 64, 0) {}
  void reset() { std::fill(w.begin(), w.end(), 0); }
  bool get(int64_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
  void set(int64_t i) { w[i >> 6] |= 1ull << (i & 63); }
  void set_atomic(int64_t i) { __atomic_fetch_or(&w[i >> 6], 1ull << (i & 63), __ATOMIC_RELAXED); }
  void swap(Bits &o) { w.swap(o.w); }
};

// Sequential L2 norm: the reference's normL2 with one thread
// (ParallelComponents.cpp:13-20), pow(x,2) being x*x exactly.
double norm_seq(int64_t n, const double *v) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += v[i] * v[i];
  return std::sqrt(s);
}
double dot_seq(const double *x, const double *y, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

// Parallel sum reduction combined in thread order (deterministic for a
// fixed thread count).
template <class F>
double par_sum(int64_t n, F term) {
  const int T = omp_get_max_threads();
  std::vector<double> part(T, 0.0);
#pragma omp parallel
  {
    double s = 0;
    const int t = omp_get_thread_num();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) s += term(i);
    part[t] = s;
  }
  double s = 0;
  for (int t = 0; t < T; t++) s += part[t];
  return s;
}

// k-centers step shared by ParHDE.cpp:124-140, WParHDE.cpp:117-135 and
// DrawingBase.cpp:142-156: min_dists = min(min_dists, col), then the first
// (lowest index) maximiser.
void kcenter_argmax(int64_t n, const double *col, double *min_dists, double *val, int64_t *idx) {
  const int T = omp_get_max_threads();
  std::vector<double> bv(T, -1.0);
  std::vector<int64_t> bi(T, 0);
#pragma omp parallel
  {
    double v = -1;
    int64_t k = 0;
    const int t = omp_get_thread_num();
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; i++) {
      if (col[i] < min_dists[i]) min_dists[i] = col[i];
      if (min_dists[i] > v) {
        v = min_dists[i];
        k = i;
      }
    }
    bv[t] = v;
    bi[t] = k;
  }
  double v = -1;
  int64_t k = 0;
  for (int t = 0; t < T; t++)  // static schedule: thread t owns a higher index range than t-1
    if (bv[t] > v) {
      v = bv[t];
      k = bi[t];
    }
  *val = v;
  *idx = k;
}

// Dedupe rule (ParHDE.cpp:142-155 and twins): when the argmax is already in
// `sources`, the last vertex (serial order) with the same min distance that
// is not in `sources` becomes the next source.
int64_t kcenter_dedupe(int64_t n, const double *min_dists, double maxval, int64_t start,
                       const std::set<int64_t> &sources) {
  if (sources.find(start) == sources.end()) return start;
  for (int64_t i = 0; i < n; i++)
    if (min_dists[i] == maxval && sources.find(i) == sources.end()) start = i;
  return start;
}

void distance_column(const or_graph *g, int64_t src, double *col, int32_t delta) {
  // DrawingBase::DistanceAlgorithm (DrawingBase.cpp:110-117): BFS when
  // unweighted, delta-stepping otherwise; beta is always 18.
  if (g->w == nullptr)
    or_dobfs(g, src, col, 15, 18);
  else
    or_deltastep(g, src, delta, col);
}

// matrixTransMultiplication (DrawingBase.cpp:221-244): C[k + i*mdim] =
// sum_j A[j + i*kdim] * B[j + k*kdim], j ascending.
void mat_trans_mult(const double *A, const double *B, double *C, int64_t mdim, int64_t kdim,
                    int64_t ndim) {
#pragma omp parallel for collapse(2) schedule(guided)
  for (int64_t i = 0; i < mdim; i++)
    for (int64_t k = 0; k < ndim; k++) {
      double sum = 0;
      for (int64_t j = 0; j < kdim; j++) sum += A[j + i * kdim] * B[j + k * kdim];
      C[k + i * mdim] = sum;
    }
}

// computeCoordinates (DrawingBase.cpp:207-219).
void compute_coords(int64_t n, int s, int c, const double *mat, const double *eig, double *X) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    for (int k = 0; k < c; k++) {
      double x = 0;
      for (int j = 0; j < s; j++) x += mat[i + j * n] * eig[j + k * s];
      X[i + k * n] = x;
    }
}

// getCoordinates (DrawingBase.cpp:196-205).
void normalise_coords(int64_t n, int c, double *X) {
  for (int k = 0; k < c; k++) {
    const double nrm = 1 / norm_seq(n, X + k * n);
    for (int64_t j = 0; j < n; j++) X[j + k * n] *= nrm;
  }
}

// Laplacian times a column-major n*s matrix (ParHDE.cpp:288-300 unweighted,
// WParHDE.cpp:229-241 weighted).
void laplacian_mult(const or_graph *g, const double *deg, const double *S, double *LS, int s) {
  const int64_t n = g->n;
#pragma omp parallel for collapse(2) schedule(guided)
  for (int k = 0; k < s; k++)
    for (int64_t i = 0; i < n; i++) {
      double v = deg[i] * S[i + k * n];
      for (uint32_t e = g->rowptr[i]; e < g->rowptr[i + 1]; e++) {
        if (g->w)
          v -= g->w[e] * S[g->adj[e] + k * n];
        else
          v -= S[g->adj[e] + k * n];
      }
      LS[i + k * n] = v;
    }
}

void finish_layout(const or_graph *g, const double *deg, const double *S, int s, int numCoord,
                   or_result *r, clk::time_point t_start, int *rc) {
  const int64_t n = g->n;
  auto t0 = clk::now();
  std::vector<double> LS((size_t)n * s), SLS((size_t)s * s);
  laplacian_mult(g, deg, S, LS.data(), s);
  mat_trans_mult(S, LS.data(), SLS.data(), s, n, s);
  auto t1 = clk::now();
  r->t_mmult += secs(t0, t1);
  if (r->SLS) memcpy(r->SLS, SLS.data(), sizeof(double) * s * s);
  std::vector<double> eig((size_t)s * numCoord), ev(s);
  int32_t iters = 0;
  if (or_compute_eigens(SLS.data(), s, numCoord, 0, eig.data(), ev.data(), &iters) != 0) {
    *rc = OR_ERR_EIG;
    return;
  }
  r->eig_iters = iters;
  if (r->eigvec) memcpy(r->eigvec, eig.data(), sizeof(double) * s * numCoord);
  if (r->eigval) memcpy(r->eigval, ev.data(), sizeof(double) * s);
  std::vector<double> X((size_t)n * numCoord);
  compute_coords(n, s, numCoord, S, eig.data(), X.data());
  r->t_hde = secs(t_start, clk::now());
  if (r->coords_raw) memcpy(r->coords_raw, X.data(), sizeof(double) * n * numCoord);
  normalise_coords(n, numCoord, X.data());
  if (r->coords) memcpy(r->coords, X.data(), sizeof(double) * n * numCoord);
  if (r->S) memcpy(r->S, S, sizeof(double) * n * s);
}

// One CGS D-orthogonalisation step of candidate An against U0 and the
// `nacc` accepted columns of S (ParHDE.cpp:204-266).  Returns ||An|| after
// the update (An is updated in place).
double cgs_step(int64_t n, int nacc, const double *S, const double *U0, double *An,
                const double *degrees) {
  const int T = omp_get_max_threads();
  std::vector<double> part_num((size_t)T * 65, 0.0), part_den((size_t)T * 65, 0.0);
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
    double num_l[65], den_l[65];
    for (int k = 0; k < 65; k++) num_l[k] = den_l[k] = 0;
#pragma omp for schedule(static)
    for (int64_t ii = 0; ii < n; ii++) {
      const double a = An[ii];
      for (int k = 0; k < nacc; k++) {
        num_l[k] += a * S[ii + n * k] * degrees[ii];
        den_l[k] += S[ii + n * k] * S[ii + n * k] * degrees[ii];
      }
      num_l[64] += a * U0[ii] * degrees[ii];
      den_l[64] += U0[ii] * U0[ii] * degrees[ii];
    }
    for (int k = 0; k < 65; k++) {
      part_num[(size_t)t * 65 + k] = num_l[k];
      part_den[(size_t)t * 65 + k] = den_l[k];
    }
  }
  double num[65] = {0}, den[65] = {0}, coef[65];
  for (int t = 0; t < T; t++)
    for (int k = 0; k < 65; k++) {
      num[k] += part_num[(size_t)t * 65 + k];
      den[k] += part_den[(size_t)t * 65 + k];
    }
  for (int k = 0; k < nacc; k++) coef[k] = num[k] * (1 / den[k]);
  coef[64] = num[64] * (1 / den[64]);
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < n; ii++) {
    double acc = 0;
    for (int k = 0; k < nacc; k++) acc += coef[k] * S[ii + n * k];
    acc += coef[64] * U0[ii];
    An[ii] -= acc;
  }
  return std::sqrt(par_sum(n, [&](int64_t i) { return An[i] * An[i]; }));
}

}  // namespace

extern "C" {

int or_num_threads(void) { return omp_get_max_threads(); }

void or_seq_bfs(const or_graph *g, int64_t src, int32_t *dist) {
  // DrawingBase::seqBfs (DrawingBase.cpp:246-271) with int levels.
  const int64_t n = g->n;
  for (int64_t i = 0; i < n; i++) dist[i] = -1;
  std::vector<uint32_t> q(n);
  int64_t head = 0, tail = 0;
  q[tail++] = (uint32_t)src;
  dist[src] = 0;
  while (head < tail) {
    const uint32_t h = q[head++];
    for (uint32_t e = g->rowptr[h]; e < g->rowptr[h + 1]; e++) {
      const uint32_t v = g->adj[e];
      if (dist[v] < 0) {
        dist[v] = dist[h] + 1;
        q[tail++] = v;
      }
    }
  }
}

void or_dobfs(const or_graph *g, int64_t src, double *dists, int alpha, int beta) {
  // DOBFS (gapbs/bfs.cc:144-184) with InitParent (:134-141), TDStep
  // (:76-106), BUStep (:48-73), QueueToBitmap / BitmapToQueue (:109-132).
  const int64_t n = g->n;
  std::vector<int64_t> parent(n);
#pragma omp parallel for
  for (int64_t v = 0; v < n; v++) parent[v] = degree(g, v) != 0 ? -degree(g, v) : -1;
  parent[src] = src;
  // sliding queue: window [qb, qe), pending [qe, qt)
  std::vector<int64_t> queue(n + 1);
  int64_t qb = 0, qe = 0, qt = 0;
  queue[qt++] = src;
  dists[src] = 0;
  qb = qe;
  qe = qt;
  Bits curr(n), front(n);
  int64_t edges_to_check = g->m;
  int64_t scout_count = degree(g, src);
  while (qb != qe) {
    if (scout_count > edges_to_check / alpha) {
      for (int64_t i = qb; i < qe; i++) front.set_atomic(queue[i]);
      int64_t awake_count = qe - qb, old_awake_count;
      qb = qe;
      qe = qt;
      do {
        old_awake_count = awake_count;
        awake_count = 0;
        curr.reset();
#pragma omp parallel for reduction(+ : awake_count) schedule(dynamic, 1024)
        for (int64_t u = 0; u < n; u++) {
          if (parent[u] < 0) {
            for (uint32_t e = g->rowptr[u]; e < g->rowptr[u + 1]; e++) {
              const int64_t v = g->adj[e];
              if (front.get(v)) {
                parent[u] = v;
                dists[u] = dists[v] + 1;
                awake_count++;
                curr.set(u);
                break;
              }
            }
          }
        }
        front.swap(curr);
      } while ((awake_count >= old_awake_count) || (awake_count > n / (int64_t)beta));
      for (int64_t v = 0; v < n; v++)
        if (front.get(v)) queue[qt++] = v;
      qb = qe;
      qe = qt;
      scout_count = 1;
    } else {
      edges_to_check -= scout_count;
      int64_t sc = 0;
      const double incr = dists[queue[qb]] + 1;
      const int64_t b = qb, e_ = qe;
#pragma omp parallel
      {
        std::vector<int64_t> local;
#pragma omp for reduction(+ : sc) schedule(dynamic, 64)
        for (int64_t q = b; q < e_; q++) {
          const int64_t u = queue[q];
          for (uint32_t e = g->rowptr[u]; e < g->rowptr[u + 1]; e++) {
            const int64_t v = g->adj[e];
            int64_t cur = parent[v];
            if (cur < 0 && __sync_bool_compare_and_swap(&parent[v], cur, u)) {
              local.push_back(v);
              sc += -cur;
              dists[v] = incr;
            }
          }
        }
        int64_t at;
#pragma omp atomic capture
        {
          at = qt;
          qt += (int64_t)local.size();
        }
        std::copy(local.begin(), local.end(), queue.begin() + at);
      }
      scout_count = sc;
      qb = qe;
      qe = qt;
    }
  }
}

void or_deltastep(const or_graph *g, int64_t src, int32_t delta, double *dists) {
  // DeltaStep (gapbs/sssp.cc:53-140): shared frontier, thread-local bins,
  // stale-vertex skip, CAS-min relax, vote for the smallest non-empty bin.
  typedef int32_t WeightT;
  const WeightT kDistInf = std::numeric_limits<WeightT>::max() / 2;
  const size_t kMaxBin = std::numeric_limits<size_t>::max() / 2;
  const int64_t n = g->n;
  std::vector<WeightT> dist(n, kDistInf);
  dist[src] = 0;
  std::vector<int64_t> frontier(std::max<int64_t>(g->m, 1) + n);
  size_t shared_indexes[2] = {0, kMaxBin};
  size_t frontier_tails[2] = {1, 0};
  frontier[0] = src;
#pragma omp parallel
  {
    std::vector<std::vector<int64_t>> local_bins(0);
    size_t iter = 0;
    while (shared_indexes[iter & 1] != kMaxBin) {
      size_t &curr_bin_index = shared_indexes[iter & 1];
      size_t &next_bin_index = shared_indexes[(iter + 1) & 1];
      size_t &curr_frontier_tail = frontier_tails[iter & 1];
      size_t &next_frontier_tail = frontier_tails[(iter + 1) & 1];
#pragma omp for nowait schedule(dynamic, 64)
      for (size_t i = 0; i < curr_frontier_tail; i++) {
        const int64_t u = frontier[i];
        if (dist[u] >= delta * static_cast<WeightT>(curr_bin_index)) {
          for (uint32_t e = g->rowptr[u]; e < g->rowptr[u + 1]; e++) {
            const uint32_t v = g->adj[e];
            WeightT old_dist = dist[v];
            const WeightT new_dist = dist[u] + g->w[e];
            if (new_dist < old_dist) {
              bool changed = true;
              while (!__sync_bool_compare_and_swap(&dist[v], old_dist, new_dist)) {
                old_dist = dist[v];
                if (old_dist <= new_dist) {
                  changed = false;
                  break;
                }
              }
              if (changed) {
                const size_t dest_bin = new_dist / delta;
                if (dest_bin >= local_bins.size()) local_bins.resize(dest_bin + 1);
                local_bins[dest_bin].push_back(v);
              }
            }
          }
        }
      }
      for (size_t i = curr_bin_index; i < local_bins.size(); i++) {
        if (!local_bins[i].empty()) {
#pragma omp critical
          next_bin_index = std::min(next_bin_index, i);
          break;
        }
      }
#pragma omp barrier
#pragma omp single nowait
      {
        curr_bin_index = kMaxBin;
        curr_frontier_tail = 0;
      }
      if (next_bin_index < local_bins.size()) {
        const size_t copy_start =
            __sync_fetch_and_add(&next_frontier_tail, local_bins[next_bin_index].size());
        std::copy(local_bins[next_bin_index].begin(), local_bins[next_bin_index].end(),
                  frontier.data() + copy_start);
        local_bins[next_bin_index].resize(0);
      }
      iter++;
#pragma omp barrier
    }
  }
#pragma omp parallel for
  for (int64_t p = 0; p < n; p++) dists[p] = dist[p];
}

void or_dijkstra(const or_graph *g, int64_t src, int64_t *dist) {
  const int64_t n = g->n;
  for (int64_t i = 0; i < n; i++) dist[i] = -1;
  typedef std::pair<int64_t, int64_t> P;
  std::priority_queue<P, std::vector<P>, std::greater<P>> pq;
  std::vector<char> done(n, 0);
  dist[src] = 0;
  pq.push(P(0, src));
  while (!pq.empty()) {
    const P t = pq.top();
    pq.pop();
    const int64_t u = t.second;
    if (done[u]) continue;
    done[u] = 1;
    for (uint32_t e = g->rowptr[u]; e < g->rowptr[u + 1]; e++) {
      const int64_t v = g->adj[e];
      const int64_t nd = t.first + (g->w ? g->w[e] : 1);
      if (dist[v] < 0 || nd < dist[v]) {
        dist[v] = nd;
        pq.push(P(nd, v));
      }
    }
  }
}

int or_power_eig(double *K, double *eValues, int64_t n, int32_t *iters) {
  // eigenValuePowerMethod (ParallelComponents.cpp:48-158) with one-thread
  // sums; start vectors from the C library's rand() after srand(1).
  const int64_t s = n, p = n;
  const double EPSILON = 1 - 1e-3;
  std::vector<double> eVecs(n * s), tmpOld(n * s);
  srand(1);
  for (int64_t i = 0; i < s; i++)
    for (int64_t j = 0; j < n; j++) eVecs[j + i * n] = ((double)rand()) / RAND_MAX;
  double r = 0;
  for (int64_t i = 0; i < s; i++) {
    eValues[i] = norm_seq(n, eVecs.data() + i * n);
    for (int64_t j = 0; j < n; j++) eVecs[j + i * n] /= eValues[i];
  }
  int32_t count = 0;
  while (r < EPSILON) {
    count++;
    if (std::isnan(r) || std::isinf(r)) {
      if (iters) *iters = count;
      return 1;
    }
    for (int64_t i = 0; i < s; i++)
      for (int64_t j = 0; j < p; j++) {
        tmpOld[j + i * n] = eVecs[j + i * n];
        eVecs[j + i * n] = 0;
      }
    for (int64_t i = 0; i < s; i++)
      for (int64_t j = 0; j < p; j++)
        for (int64_t k = 0; k < p; k++) eVecs[k + i * p] += K[k + j * p] * tmpOld[j + i * p];
    for (int64_t i = 0; i < s; i++)
      for (int64_t j = 0; j < i; j++) {
        const double fac = dot_seq(eVecs.data() + j * p, eVecs.data() + i * p, p) /
                           dot_seq(eVecs.data() + j * p, eVecs.data() + j * p, p);
        for (int64_t k = 0; k < p; k++) eVecs[k + i * p] -= fac * eVecs[k + j * p];
      }
    for (int64_t i = 0; i < s; i++) {
      eValues[i] = norm_seq(p, eVecs.data() + i * n);
      for (int64_t j = 0; j < n; j++) eVecs[j + i * n] /= eValues[i];
    }
    r = 1;
    for (int64_t i = 0; i < s; i++) {
      double tmp = dot_seq(eVecs.data() + i * n, tmpOld.data() + i * n, n);
      if (tmp < 0) tmp *= -1;
      if (tmp < r) r = tmp;
    }
  }
  for (int64_t a = 0, b = s - 1; a < b; a++, b--) std::swap(eValues[a], eValues[b]);
  for (int64_t i = 0; i < s; i++) {
    const bool flip = eVecs[(s - i - 1) * n] < 0;
    for (int64_t j = 0; j < n; j++)
      K[j + (s - i - 1) * n] = flip ? -eVecs[j + i * n] : eVecs[j + i * n];
  }
  if (iters) *iters = count;
  return 0;
}

int or_compute_eigens(double *K, int64_t dim, int numCoord, int which, double *eigvec,
                      double *eigval, int32_t *iters) {
  // compute_eigens + mkl_eigen_value_vector_comp_{first,largest}
  // (DrawingBase.cpp:39-108), non-MKL branch.
  std::vector<double> w(dim);
  if (or_power_eig(K, w.data(), dim, iters) != 0) return OR_ERR_EIG;
  for (int64_t j = 0; j < numCoord; j++)
    for (int64_t i = 0; i < dim; i++)
      eigvec[i + j * dim] = which == 1 ? K[i + (dim - j - 1) * dim] : K[i + j * dim];
  if (eigval) memcpy(eigval, w.data(), sizeof(double) * dim);
  return OR_OK;
}

// ParHDE::run (ParHDE.cpp:45-352), one try.  given == nullptr: k-centers
// sources (doRand == 2); otherwise the candidate sources in the given order
// (the random strategies, ParHDE.cpp:158-172, and the learned proposals),
// traversed lazily until s columns are accepted.
static int parhde_impl(const or_graph *g, int s, int numCoord, int lazy, const int64_t *given,
                       int ngiven, or_result *r) {
  if (!g || !r || s < 1 || s > 64 || numCoord < 1 || numCoord > s || g->w) return OR_ERR_ARG;
  const int64_t n = g->n;
  const int limit = given ? ngiven : 3 * s;
  if (given) lazy = 1;
  r->n_bfs = r->n_accepted = 0;
  r->t_hde = r->t_dist = r->t_orth = r->t_mmult = 0;
  if (r->accepted)
    for (int i = 0; i < limit; i++) r->accepted[i] = -1;
  std::vector<double> degrees(n);
  for (int64_t i = 0; i < n; i++) degrees[i] = (double)(g->rowptr[i + 1] - g->rowptr[i]);

  auto t_start = clk::now();
  // Candidate storage: eager keeps all 3s columns, lazy one at a time.
  const int64_t ncols = lazy ? 2 : (int64_t)limit + 1;
  std::vector<double> U((size_t)n * ncols);
  std::vector<double> S((size_t)n * s);
  std::vector<double> min_dists(n);
  for (int64_t p = 0; p < n; p++) {
    U[p] = 1;
    min_dists[p] = INT_MAX;
  }
  {
    const double nrm = 1 / std::sqrt(par_sum(n, [&](int64_t i) { return U[i] * U[i]; }));
    for (int64_t p = 0; p < n; p++) U[p] = U[p] * nrm;
  }
  int64_t start_idx = 0;
  std::set<int64_t> sources;
  // Per-column k-centers step + L2 normalisation of the new column.
  auto traverse = [&](int run_count, double *col) {
    if (given) start_idx = given[run_count - 1];
    if (r->pivots) r->pivots[run_count - 1] = start_idx;
    distance_column(g, start_idx, col, 0);
    r->n_bfs++;
    if (given) return;
    double maxval;
    int64_t idx;
    kcenter_argmax(n, col, min_dists.data(), &maxval, &idx);
    start_idx = idx;
    const int64_t old_idx = start_idx;
    if (run_count != 1) start_idx = kcenter_dedupe(n, min_dists.data(), maxval, start_idx, sources);
    sources.insert(old_idx);
  };
  auto normalise = [&](double *col) {
    double nd = 0;
    for (int64_t p = 0; p < n; p++) nd += col[p] * col[p];
    nd = 1 / std::sqrt(nd);
    for (int64_t p = 0; p < n; p++) col[p] = col[p] * nd;
  };
  int correct_pivots = 1;
  // CGS D-orthogonalisation of one candidate (ParHDE.cpp:193-277).
  auto cgs = [&](int run_count, double *An) {
    const int j = correct_pivots;
    double nd = cgs_step(n, j - 1, S.data(), U.data(), An, degrees.data());
    if (r->normdist) r->normdist[run_count - 1] = nd;
    if (nd < 0.001) {
      std::cout << "discarding vec " << j << ", normdist " << nd << std::endl;
      if (r->accepted) r->accepted[run_count - 1] = 0;
    } else {
      nd = 1 / nd;
#pragma omp parallel for schedule(static)
      for (int64_t p = 0; p < n; p++) S[p + (j - 1) * (int64_t)n] = An[p] * nd;
      correct_pivots++;
      if (r->accepted) r->accepted[run_count - 1] = 1;
    }
  };

  if (!lazy) {
    auto t0 = clk::now();
    for (int rc_ = 1; rc_ <= limit; rc_++) traverse(rc_, U.data() + (size_t)rc_ * n);
#pragma omp parallel for
    for (int rc_ = 1; rc_ <= limit; rc_++) normalise(U.data() + (size_t)rc_ * n);
    r->t_dist = secs(t_start, clk::now());
    (void)t0;
    auto t1 = clk::now();
    for (int rc_ = 1; rc_ <= limit; rc_++) {
      if (correct_pivots > s) break;
      cgs(rc_, U.data() + (size_t)rc_ * n);
    }
    r->t_orth = secs(t1, clk::now());
  } else {
    r->t_dist = secs(t_start, clk::now());  // ones vector set-up
    for (int rc_ = 1; rc_ <= limit; rc_++) {
      if (correct_pivots > s) break;
      auto t0 = clk::now();
      double *col = U.data() + n;
      traverse(rc_, col);
      normalise(col);
      auto t1 = clk::now();
      cgs(rc_, col);
      auto t2 = clk::now();
      r->t_dist += secs(t0, t1);
      r->t_orth += secs(t1, t2);
    }
  }
  r->n_accepted = correct_pivots - 1;
  if (correct_pivots - 1 < s) {
    std::cerr << "Number of Pivots obtained: " << correct_pivots
              << " inadequate, increase the Pivots Limits: " << limit << std::endl;
    return OR_ERR_PIVOTS;
  }
  int rc = OR_OK;
  finish_layout(g, degrees.data(), S.data(), s, numCoord, r, t_start, &rc);
  return rc;
}

int or_parhde(const or_graph *g, int s, int numCoord, int lazy, or_result *r) {
  return parhde_impl(g, s, numCoord, lazy, nullptr, 0, r);
}

int or_parhde_sources(const or_graph *g, int s, int numCoord, const int64_t *sources, int nsrc,
                      or_result *r) {
  if (!sources || nsrc < 1) return OR_ERR_ARG;
  return parhde_impl(g, s, numCoord, 1, sources, nsrc, r);
}

int or_wparhde(const or_graph *g, int s, int numCoord, int32_t delta, or_result *r) {
  // WParHDE::run (WParHDE.cpp:40-291): SSSP k-centers coupled with MGS.
  if (!g || !r || s < 1 || s > 64 || numCoord < 1 || numCoord > s) return OR_ERR_ARG;
  if (!g->w) return OR_ERR_WEIGHTS;
  if (delta < 1) delta = 10000;
  const int64_t n = g->n;
  const int limit = 3 * s;
  r->n_bfs = r->n_accepted = 0;
  r->t_hde = r->t_dist = r->t_orth = r->t_mmult = 0;
  if (r->accepted)
    for (int i = 0; i < limit; i++) r->accepted[i] = -1;
  std::vector<double> degrees(n);
  for (int64_t i = 0; i < n; i++) {
    degrees[i] = 0;
    for (uint32_t k = g->rowptr[i]; k < g->rowptr[i + 1]; k++) degrees[i] += (double)g->w[k];
  }
  auto t_start = clk::now();
  std::vector<double> U((size_t)n * (s + 1)), S((size_t)n * s), min_dists(n);
  for (int64_t p = 0; p < n; p++) {
    U[p] = 1;
    min_dists[p] = INT_MAX;
  }
  {
    const double nrm = 1 / std::sqrt(par_sum(n, [&](int64_t i) { return U[i] * U[i]; }));
    for (int64_t p = 0; p < n; p++) U[p] *= nrm;
  }
  int64_t start_idx = 0;
  std::set<int64_t> sources;
  int correct_pivots = 1;
  auto t_inner = clk::now();
  for (int run_count = 1; run_count <= limit; run_count++) {
    if (correct_pivots > s) break;
    double *col = U.data() + (size_t)correct_pivots * n;
    if (r->pivots) r->pivots[run_count - 1] = start_idx;
    distance_column(g, start_idx, col, delta);
    r->n_bfs++;
    double maxval;
    int64_t idx;
    kcenter_argmax(n, col, min_dists.data(), &maxval, &idx);
    start_idx = idx;
    const int64_t old_idx = start_idx;
    if (correct_pivots != 1) start_idx = kcenter_dedupe(n, min_dists.data(), maxval, start_idx, sources);
    sources.insert(old_idx);
    double nd = 1 / std::sqrt(par_sum(n, [&](int64_t i) { return col[i] * col[i]; }));
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) col[p] = col[p] * nd;
    auto t_mid = clk::now();
    r->t_dist += secs(t_inner, t_mid);
    t_inner = t_mid;
    // Coupled MGS against U[:,0..j-1] (WParHDE.cpp:167-190).
    int j = correct_pivots;
    double *Uj = U.data() + (size_t)j * n;
    for (int k = 0; k < j; k++) {
      const double *Uk = U.data() + (size_t)k * n;
      double den = par_sum(n, [&](int64_t p) { return Uk[p] * (Uk[p] * degrees[p]); });
      const double num = par_sum(n, [&](int64_t p) { return Uj[p] * (Uk[p] * degrees[p]); });
      den = 1 / den;
#pragma omp parallel for schedule(static)
      for (int64_t p = 0; p < n; p++) Uj[p] -= (num * Uk[p] * den);
    }
    nd = std::sqrt(par_sum(n, [&](int64_t i) { return Uj[i] * Uj[i]; }));
    if (r->normdist) r->normdist[run_count - 1] = nd;
    if (nd < 0.001) {
      std::cout << "discarding vec " << j << ", normdist " << nd << std::endl;
      if (r->accepted) r->accepted[run_count - 1] = 0;
      j--;
      correct_pivots--;
    } else {
      nd = 1 / nd;
#pragma omp parallel for schedule(static)
      for (int64_t p = 0; p < n; p++) Uj[p] = Uj[p] * nd;
      if (r->accepted) r->accepted[run_count - 1] = 1;
    }
    // WParHDE.cpp:211-213 copies column j (after the discard decrement, the
    // previous accepted column, which is already in place) into S[:, j-1].
    if (j >= 1) {
      const double *Ujj = U.data() + (size_t)j * n;
#pragma omp parallel for schedule(static)
      for (int64_t p = 0; p < n; p++) S[p + (size_t)(j - 1) * n] = Ujj[p];
    }
    correct_pivots++;
    auto t_end = clk::now();
    r->t_orth += secs(t_inner, t_end);
    t_inner = t_end;
  }
  r->n_accepted = correct_pivots - 1;
  if (correct_pivots - 1 < s) {
    std::cerr << "Number of Pivots obtained: " << correct_pivots
              << " inadequate, increase the Pivots Limits: " << limit << std::endl;
    return OR_ERR_PIVOTS;
  }
  int rc = OR_OK;
  finish_layout(g, degrees.data(), S.data(), s, numCoord, r, t_start, &rc);
  return rc;
}

namespace {
// distancePhaseComputation (DrawingBase.cpp:119-194), k-centers branch:
// exactly s traversals, columns kept as raw distances.
void distance_phase(const or_graph *g, int s, double *D, or_result *r) {
  const int64_t n = g->n;
  std::vector<double> min_dists(n, (double)INT_MAX);
  int64_t start_idx = 0;
  std::set<int64_t> sources;
  for (int run_count = 0; run_count < s; run_count++) {
    if (r->pivots) r->pivots[run_count] = start_idx;
    distance_column(g, start_idx, D + (size_t)run_count * n, 10000);
    r->n_bfs++;
    const int64_t old_idx = start_idx;
    double maxval;
    int64_t idx;
    kcenter_argmax(n, D + (size_t)run_count * n, min_dists.data(), &maxval, &idx);
    start_idx = idx;
    if (run_count) start_idx = kcenter_dedupe(n, min_dists.data(), maxval, start_idx, sources);
    sources.insert(old_idx);
  }
}

int finish_mds(const or_graph *g, int s, int numCoord, const double *C, or_result *r,
               clk::time_point t_start) {
  const int64_t n = g->n;
  auto t0 = clk::now();
  std::vector<double> CtC((size_t)s * s);
  mat_trans_mult(C, C, CtC.data(), s, n, s);
  r->t_mmult = secs(t0, clk::now());
  if (r->SLS) memcpy(r->SLS, CtC.data(), sizeof(double) * s * s);
  std::vector<double> eig((size_t)s * numCoord), ev(s);
  int32_t iters = 0;
  if (or_compute_eigens(CtC.data(), s, numCoord, 1, eig.data(), ev.data(), &iters) != 0)
    return OR_ERR_EIG;
  r->eig_iters = iters;
  if (r->eigvec) memcpy(r->eigvec, eig.data(), sizeof(double) * s * numCoord);
  if (r->eigval) memcpy(r->eigval, ev.data(), sizeof(double) * s);
  std::vector<double> X((size_t)n * numCoord);
  compute_coords(n, s, numCoord, C, eig.data(), X.data());
  r->t_hde = secs(t_start, clk::now());
  if (r->coords_raw) memcpy(r->coords_raw, X.data(), sizeof(double) * n * numCoord);
  normalise_coords(n, numCoord, X.data());
  if (r->coords) memcpy(r->coords, X.data(), sizeof(double) * n * numCoord);
  if (r->S) memcpy(r->S, C, sizeof(double) * n * s);
  return OR_OK;
}
}  // namespace

int or_pivotmds(const or_graph *g, int s, int numCoord, or_result *r) {
  // PivotMDS::run (PivotMDS.cpp:68-147) with doubleCentering (:30-66).
  if (!g || !r || s < 1 || s > 64 || numCoord < 1 || numCoord > s) return OR_ERR_ARG;
  const int64_t n = g->n;
  r->n_bfs = r->n_accepted = 0;
  r->t_hde = r->t_dist = r->t_orth = r->t_mmult = 0;
  auto t_start = clk::now();
  std::vector<double> D((size_t)n * s), C((size_t)n * s);
  distance_phase(g, s, D.data(), r);
  r->t_dist = secs(t_start, clk::now());
  auto t0 = clk::now();
  std::vector<double> colSum(s), rowSum(n, 0.0);
  double total = 0;
  for (int i = 0; i < s; i++) {
    double cs = 0;
    for (int64_t j = 0; j < n; j++) {
      const double d2 = D[j + (size_t)i * n] * D[j + (size_t)i * n];
      cs += d2;
      rowSum[j] += d2;
    }
    colSum[i] = cs;
    total += cs;
  }
  const double rN = (double)1 / (double)n, rK = (double)1 / (double)s;
#pragma omp parallel for collapse(2) schedule(guided)
  for (int i = 0; i < s; i++)
    for (int64_t j = 0; j < n; j++) {
      const double d = D[j + (size_t)i * n];
      C[j + (size_t)i * n] = -0.5 * (d * d - (rN * colSum[i]) - (rK * rowSum[j]) + (rK * rN * total));
    }
  r->t_orth = secs(t0, clk::now());
  r->n_accepted = s;
  return finish_mds(g, s, numCoord, C.data(), r, t_start);
}

int or_phde(const or_graph *g, int s, int numCoord, or_result *r) {
  // PHDE::run (PHDE.cpp:50-136) with columnCentering (:30-47).
  if (!g || !r || s < 1 || s > 64 || numCoord < 1 || numCoord > s) return OR_ERR_ARG;
  const int64_t n = g->n;
  r->n_bfs = r->n_accepted = 0;
  r->t_hde = r->t_dist = r->t_orth = r->t_mmult = 0;
  auto t_start = clk::now();
  std::vector<double> D((size_t)n * s), C((size_t)n * s);
  distance_phase(g, s, D.data(), r);
  r->t_dist = secs(t_start, clk::now());
  auto t0 = clk::now();
  for (int i = 0; i < s; i++) {
    double cs = 0;
    for (int64_t j = 0; j < n; j++) cs += D[j + (size_t)i * n];
    cs /= n;
    for (int64_t j = 0; j < n; j++) C[j + (size_t)i * n] = D[j + (size_t)i * n] - cs;
  }
  r->t_orth = secs(t0, clk::now());
  r->n_accepted = s;
  return finish_mds(g, s, numCoord, C.data(), r, t_start);
}

int or_parhde_sample(const or_graph *g, int s, int numCoord, int nbfs, or_sample *out) {
  // weighted graphs: WParHDE's traversals (delta-stepping, the reference's
  // delta 10000, WParHDE.cpp:104); its MGS step is costed as a CGS step (the
  // same dot products and updates per candidate)
  if (!g || !out || s < 2 || s > 64 || nbfs < 1) return OR_ERR_ARG;
  const int64_t n = g->n;
  auto t_begin = clk::now();
  std::vector<double> degrees(n);
  for (int64_t i = 0; i < n; i++) degrees[i] = (double)(g->rowptr[i + 1] - g->rowptr[i]);
  // traversals, exactly as or_parhde's k-centers loop
  std::vector<double> U0(n, 1.0), col(n), min_dists(n, (double)INT_MAX);
  {
    const double nrm = 1 / std::sqrt(par_sum(n, [&](int64_t i) { return U0[i] * U0[i]; }));
    for (int64_t p = 0; p < n; p++) U0[p] *= nrm;
  }
  int64_t src = 0;
  std::set<int64_t> sources;
  auto t0 = clk::now();
  for (int k = 0; k < nbfs; k++) {
    distance_column(g, src, col.data(), g->w ? 10000 : 0);
    double maxval;
    int64_t idx;
    kcenter_argmax(n, col.data(), min_dists.data(), &maxval, &idx);
    const int64_t old = idx;
    if (k) idx = kcenter_dedupe(n, min_dists.data(), maxval, idx, sources);
    sources.insert(old);
    src = idx;
    double nd = 0;
    for (int64_t p = 0; p < n; p++) nd += col[p] * col[p];
    nd = 1 / std::sqrt(nd);
    for (int64_t p = 0; p < n; p++) col[p] *= nd;
  }
  out->t_bfs_each = secs(t0, clk::now()) / nbfs;
  out->n_bfs_timed = nbfs;
  // stand-in columns: shifted copies of the last distance column (the cost
  // of the dense stages does not depend on the values)
  std::vector<double> S((size_t)n * s), An(n);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < n; p++)
    for (int k = 0; k < s; k++) S[p + (size_t)k * n] = col[(p + 7919 * (int64_t)k) % n] + 1e-3 * k;
  // CGS at three sizes, least-squares line over j, summed over j = 1..s
  const int js[3] = {1, (s + 1) / 2, s};
  double tj[3];
  for (int q = 0; q < 3; q++) {
    std::copy(col.begin(), col.end(), An.begin());
    auto a = clk::now();
    (void)cgs_step(n, js[q] - 1, S.data(), U0.data(), An.data(), degrees.data());
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; p++) An[p] = An[p] * 0.5;  // the accept-and-scale pass
    tj[q] = secs(a, clk::now());
  }
  {
    double mx = 0, my = 0;
    for (int q = 0; q < 3; q++) {
      mx += js[q];
      my += tj[q];
    }
    mx /= 3;
    my /= 3;
    double sxy = 0, sxx = 0;
    for (int q = 0; q < 3; q++) {
      sxy += (js[q] - mx) * (tj[q] - my);
      sxx += (js[q] - mx) * (js[q] - mx);
    }
    const double slope = sxx > 0 ? sxy / sxx : 0, icpt = my - slope * mx;
    double tot = 0;
    for (int j = 1; j <= s; j++) tot += std::max(0.0, icpt + slope * j);
    out->t_orth_total = tot;
  }
  auto t2 = clk::now();
  std::vector<double> LS((size_t)n * s), SLS((size_t)s * s);
  laplacian_mult(g, degrees.data(), S.data(), LS.data(), s);
  mat_trans_mult(S.data(), LS.data(), SLS.data(), s, n, s);
  out->t_mmult = secs(t2, clk::now());
  // power method on a symmetric s x s matrix with the same scale
  std::vector<double> K((size_t)s * s), ev(s);
  for (int i = 0; i < s; i++)
    for (int k = 0; k < s; k++) K[k + i * s] = 0.5 * (SLS[k + i * s] + SLS[i + k * s]);
  auto t3 = clk::now();
  int32_t iters = 0;
  std::vector<double> eig((size_t)s * numCoord);
  (void)or_compute_eigens(K.data(), s, numCoord, 0, eig.data(), ev.data(), &iters);
  out->t_eig = secs(t3, clk::now());
  auto t4 = clk::now();
  std::vector<double> X((size_t)n * numCoord);
  compute_coords(n, s, numCoord, S.data(), eig.data(), X.data());
  normalise_coords(n, numCoord, X.data());
  out->t_coords = secs(t4, clk::now());
  out->t_spent = secs(t_begin, clk::now());
  return OR_OK;
}

double or_hall_energy(const or_graph *g, const double *Y, int c) {
  const int64_t n = g->n;
  long double num = 0, den = 0;
  for (int k = 0; k < c; k++) {
    const double *y = Y + (size_t)k * n;
    for (int64_t i = 0; i < n; i++) {
      long double di = 0;
      for (uint32_t e = g->rowptr[i]; e < g->rowptr[i + 1]; e++) {
        const long double w = g->w ? g->w[e] : 1;
        const long double diff = (long double)y[i] - y[g->adj[e]];
        num += 0.5L * w * diff * diff;
        di += w;
      }
      den += di * (long double)y[i] * y[i];
    }
  }
  return (double)(num / den);
}

}  // extern "C"
