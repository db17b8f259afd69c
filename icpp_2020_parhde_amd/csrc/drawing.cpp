// drawing.cpp — the reference's C++ class interface (DrawingBase, ParHDE,
// WParHDE, PivotMDS, PHDE, ParallelComponents, DOBFS, DeltaStep) on top of
// the C ABI of libparhde.so.  Every n-sized computation goes through the
// GPU library; this file only marshals arguments, keeps the per-object
// buffers the reference exposes (eigenVectors, coordinates) and prints the
// reference's progress and timing lines.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "DrawingBase.h"
#include "PHDE.h"
#include "ParHDE.h"
#include "PivotMDS.h"
#include "WParHDE.h"

namespace {

[[noreturn]] void die(const char *what, int rc, int code) {
  std::cerr << what << ": " << phde_last_error() << " (status " << rc << ")" << std::endl;
  std::exit(code);
}

// Fingerprint of an array's contents: four independent lanes of 64-bit
// multiply-rotate mixing over 8-byte words, folded together (not
// cryptographic; a changed word changes the result with overwhelming
// probability).  About 10 GB/s on one core.
uint64_t fingerprint(const void *p, size_t bytes) {
  const unsigned char *b = static_cast<const unsigned char *>(p);
  const uint64_t K = 0x9E3779B97F4A7C15ull;
  uint64_t h[4] = {K, K ^ 1, K ^ 2, K ^ 3};
  auto mix = [&](uint64_t &x, uint64_t v) {
    x ^= v * 0xC2B2AE3D27D4EB4Full;
    x = (x << 31 | x >> 33) * K;
  };
  size_t i = 0;
  for (; i + 32 <= bytes; i += 32) {
    uint64_t v[4];
    std::memcpy(v, b + i, 32);
    for (int j = 0; j < 4; j++) mix(h[j], v[j]);
  }
  uint64_t tail = bytes;
  for (; i < bytes; i++) tail = tail * 131 + b[i];
  mix(h[0], tail);
  return h[0] ^ (h[1] << 1 | h[1] >> 63) ^ (h[2] << 2 | h[2] >> 62) ^ (h[3] << 3 | h[3] >> 61);
}

// A context for callers that hand raw arrays to the free functions: the last
// graph is kept resident so repeated DOBFS / DeltaStep calls on the same
// arrays do not re-upload it.  The reference reads the arrays it is given on
// every call, so the cache is keyed on their contents as well as their
// addresses: arrays changed in place, or freed and allocated again at the
// same addresses, are uploaded again.
struct GraphCache {
  std::mutex mu;
  const void *row = nullptr, *col = nullptr, *w = nullptr;
  long n = -1, m = -1;
  uint64_t fp = 0;
  phde_ctx *ctx = nullptr;
  static uint64_t contents(const unsigned *r, const unsigned *c, const int32_t *wt, long nn, long mm) {
    uint64_t f = fingerprint(r, sizeof(unsigned) * (size_t)(nn + 1));
    f = f * 31 + fingerprint(c, sizeof(unsigned) * (size_t)mm);
    if (wt) f = f * 31 + fingerprint(wt, sizeof(int32_t) * (size_t)mm);
    return f;
  }
  phde_ctx *get(const unsigned *r, const unsigned *c, const int32_t *wt, long nn, long mm) {
    const uint64_t f = nn >= 0 && mm >= 0 && r && (c || mm == 0) ? contents(r, c, wt, nn, mm) : 0;
    if (ctx && r == row && c == col && wt == w && nn == n && mm == m && f == fp) return ctx;
    if (ctx) phde_destroy(ctx);
    ctx = nullptr;
    int rc = phde_create(r, c, wt, nn, mm, &ctx);
    if (rc) die("graph upload", rc, 1);
    row = r;
    col = c;
    w = wt;
    n = nn;
    m = mm;
    fp = f;
    return ctx;
  }
};
GraphCache &cache() {
  static GraphCache g;
  return g;
}

// The GPU state of each layout object, keyed by its address: the objects
// keep the reference's layout (variants/DrawingBase.h:13-48), which has no
// room for it.  Allocated once and never destroyed, so that objects that
// outlive static destruction still find it.
struct DrawState {
  phde_ctx *ctx = nullptr;
  phde_options opts;
  phde_stats stats;
};
std::mutex &states_mu() {
  static std::mutex *mu = new std::mutex;
  return *mu;
}
std::unordered_map<const DrawingBase *, DrawState> &states() {
  static auto *m = new std::unordered_map<const DrawingBase *, DrawState>;
  return *m;
}
// References to map elements stay valid while other objects come and go.
DrawState &state_of(const DrawingBase *b) {
  std::lock_guard<std::mutex> lk(states_mu());
  return states().at(b);
}

// Context for the eigen helpers, which need a device but no graph.
phde_ctx *eig_ctx() {
  static phde_ctx *c = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    static const uint32_t rp[2] = {0, 0};
    int rc = phde_create(rp, nullptr, nullptr, 1, 0, &c);
    if (rc) die("device context", rc, 1);
  });
  return c;
}

}  // namespace

// ---------------------------------------------------------------- free functions

void DOBFS(unsigned int *row, unsigned int *col, long source, long numNodes, long numEdges,
           double *dists, int alpha, int beta) {
  std::lock_guard<std::mutex> lk(cache().mu);
  phde_ctx *c = cache().get(row, col, nullptr, numNodes, numEdges);
  std::vector<int32_t> d(numNodes);
  int rc = phde_bfs(c, source, alpha, beta, d.data());
  if (rc) die("DOBFS", rc, 1);
  // bfs.cc writes levels of reached vertices only
  for (long i = 0; i < numNodes; i++)
    if (d[i] >= 0) dists[i] = (double)d[i];
}

double *DeltaStep(const unsigned int *row, unsigned int *col, int32_t *weights, long source,
                  long numNodes, long numEdges, double *dist, int32_t delta) {
  std::lock_guard<std::mutex> lk(cache().mu);
  phde_ctx *c = cache().get(row, col, weights, numNodes, numEdges);
  std::vector<int32_t> d(numNodes);
  int rc = phde_sssp(c, source, delta, d.data());
  if (rc) die("DeltaStep", rc, 1);
  for (long i = 0; i < numNodes; i++) dist[i] = (double)d[i];
  return dist;
}

// ---------------------------------------------------------------- ParallelComponents

ParallelComponents::ParallelComponents() {}

double ParallelComponents::normL2(long n, double *vec) {
  double s = 0;
  for (long i = 0; i < n; i++) s += vec[i] * vec[i];
  return std::sqrt(s);
}

double ParallelComponents::dotProd(double *x, double *y, int n) {
  double s = 0;
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

int ParallelComponents::eigenValuePowerMethod(double *K, double *eValues, long n, int) {
  int rc = phde_power_eig(eig_ctx(), K, eValues, (int)n, nullptr);
  return rc == PHDE_OK ? 0 : 1;
}

// ---------------------------------------------------------------- DrawingBase

DrawingBase::DrawingBase(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m,
                         int numSources, int doRand, int numCoord)
    : rowPtrs(rowPtrs),
      adj(adj),
      weights(weights),
      n(n),
      m(m),
      numSources(numSources),
      numCoord(numCoord),
      doRand(doRand) {
  eigenVectors = (double *)malloc(sizeof(double) * numSources * numCoord);
  coordinates = (double *)malloc(sizeof(double) * n * numCoord);
  DrawState st;
  std::memset(&st.stats, 0, sizeof(st.stats));
  phde_default_options(&st.opts);
  st.opts.s = numSources;
  st.opts.num_coord = numCoord;
  // doRand as spectralDrawing.cpp:88-95 maps -p (0 RandCoarse, 1 RandFine,
  // 2 Kcenters); 3 selects the learned proposal (setPivotNet)
  st.opts.strategy = doRand == 3   ? PHDE_PIVOTS_LEARNED
                     : doRand == 2 ? PHDE_PIVOTS_KCENTERS
                                   : (doRand == 1 ? PHDE_PIVOTS_RANDFINE : PHDE_PIVOTS_RANDCOARSE);
  int rc = phde_create(rowPtrs, adj, weights, n, m, &st.ctx);
  if (rc) die("graph upload", rc, 1);
  std::lock_guard<std::mutex> lk(states_mu());
  states()[this] = st;
}

phde_options &DrawingBase::options() { return state_of(this).opts; }
const phde_stats &DrawingBase::stats() const { return state_of(this).stats; }
int DrawingBase::setPivotNet(phde_mpnn *net) { return phde_set_pivot_net(state_of(this).ctx, net); }

DrawingBase::~DrawingBase() {
  free(eigenVectors);
  free(coordinates);
  phde_ctx *ctx = nullptr;
  {
    std::lock_guard<std::mutex> lk(states_mu());
    auto it = states().find(this);
    if (it != states().end()) {
      ctx = it->second.ctx;
      states().erase(it);
    }
  }
  if (ctx) phde_destroy(ctx);
}

void DrawingBase::compute_eigens(double *matrix, int dimension, double *eigVecs, EigenChoice c) {
  // DrawingBase.cpp:39-108 (non-MKL): power method, then the first or last
  // numCoord columns.
  std::vector<double> w(dimension);
  if (phde_power_eig(state_of(this).ctx, matrix, w.data(), dimension, nullptr) != PHDE_OK) {
    printf("The algorithm failed to compute eigenvalues.\n");
    std::exit(1);
  }
  for (long j = 0; j < numCoord; j++)
    for (long i = 0; i < dimension; i++)
      eigVecs[i + j * dimension] =
          c == Largest ? matrix[i + (dimension - j - 1) * dimension] : matrix[i + j * dimension];
}

void DrawingBase::DistanceAlgorithm(unsigned int *row, unsigned int *col, long source, long numNodes,
                                    long numEdges, double *dists, int32_t *wts, int32_t delta,
                                    int alpha, int beta) {
  (void)beta;  // DrawingBase.cpp:114 always passes beta = 18
  if (wts == nullptr)
    DOBFS(row, col, source, numNodes, numEdges, dists, alpha, 18);
  else
    DeltaStep(row, col, wts, source, numNodes, numEdges, dists, delta);
}

void DrawingBase::seqBfs(unsigned int start, double *distCol) {
  std::vector<int32_t> d(n);
  int rc = phde_bfs(state_of(this).ctx, start, 15, 18, d.data());
  if (rc) die("seqBfs", rc, 1);
  for (long i = 0; i < n; i++)
    if (d[i] >= 0) distCol[i] = (double)d[i];
}

void DrawingBase::distancePhaseComputation(double **distMatrix, double *eltDistance, int trial,
                                           int alpha, int beta, int32_t delta) {
  // k-centers traversals (DrawingBase.cpp:125-172) through the library's
  // multi-column distance call; the device computes columns one source at a
  // time, the sources chosen exactly as the reference does.
  auto t0 = std::chrono::high_resolution_clock::now();
  std::vector<double> min_dists(n, (double)INT_MAX);
  long start_idx = 0;
  std::vector<char> used(n, 0);
  for (int run_count = 0; run_count < numSources; run_count++) {
    double *col = (*distMatrix) + (size_t)run_count * n;
    DistanceAlgorithm(rowPtrs, adj, start_idx, n, m, col, weights, delta, alpha, beta);
    const long old_idx = start_idx;
    double best = -1;
    long arg = 0;
    for (long i = 0; i < n; i++) {
      if (col[i] < min_dists[i]) min_dists[i] = col[i];
      if (min_dists[i] > best) {
        best = min_dists[i];
        arg = i;
      }
    }
    start_idx = arg;
    if (run_count && used[start_idx])
      for (long i = 0; i < n; i++)
        if (min_dists[i] == best && !used[i]) start_idx = i;
    used[old_idx] = 1;
  }
  eltDistance[trial] =
      std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
}

double *DrawingBase::getCoordinates() {
  // DrawingBase.cpp:196-205: normalise every column in place, on each call.
  for (int i = 0; i < numCoord; i++) {
    const double nrm = 1 / parComp.normL2(n, coordinates + (size_t)i * n);
    for (long j = 0; j < n; j++) coordinates[j + (size_t)i * n] *= nrm;
  }
  return coordinates;
}

void DrawingBase::computeCoordinates(double *mat) {
  // DrawingBase.cpp:207-219.
  for (long i = 0; i < n; i++)
    for (int k = 0; k < numCoord; k++) {
      double x = 0;
      for (int j = 0; j < numSources; j++) x += mat[i + (size_t)j * n] * eigenVectors[j + k * numSources];
      coordinates[i + (size_t)k * n] = x;
    }
}

void DrawingBase::matrixTransMultiplication(double *A, double *B, double *C, long mdim, long kdim,
                                            long ndim) {
  for (long i = 0; i < mdim; i++)
    for (long k = 0; k < ndim; k++) {
      double sum = 0;
      for (long j = 0; j < kdim; j++) sum += A[j + i * kdim] * B[j + k * kdim];
      C[k + i * mdim] = sum;
    }
}

int DrawingBase::runLayout(int algo, int tries, int alpha, int beta, int32_t delta) {
  DrawState &S = state_of(this);
  phde_options &opts = S.opts;
  phde_stats &lastStats = S.stats;
  opts.algorithm = algo;
  opts.alpha = alpha > 0 ? alpha : 15;
  opts.beta = 18;
  (void)beta;
  if (delta > 1) opts.delta = delta;
  std::vector<double> X((size_t)n * numCoord);
  double tHDE = 0, tDist = 0, tOrth = 0, tMM = 0;
  int rc = PHDE_OK;
  for (int t = 0; t < tries; t++) {
    rc = phde_layout(S.ctx, &opts, X.data(), &lastStats);
    if (rc == PHDE_ERR_WEIGHTS) {
      std::cerr << (algo == PHDE_ALGO_WPARHDE ? "Weighted ParHDE code requires the weighted edges"
                                              : "ParHDE code doesn't requires the weighted edges")
                << std::endl;
      std::exit(0);
    }
    if (rc == PHDE_ERR_PIVOTS) {
      std::cerr << "Number of Pivots obtained: " << lastStats.n_accepted + 1
                << " inadequate, increase the Pivots Limits: " << 3 * numSources << std::endl;
      std::exit(0);
    }
    if (rc == PHDE_ERR_EIG) {
      printf("The algorithm failed to compute eigenvalues.\n");
      std::exit(1);
    }
    if (rc) die("layout", rc, 1);
    tHDE += lastStats.ms_total / 1000;
    tDist += lastStats.ms_distance / 1000;
    tOrth += lastStats.ms_orth / 1000;
    tMM += lastStats.ms_mmult / 1000;
  }
  std::memcpy(coordinates, X.data(), sizeof(double) * X.size());
  const double d = tries > 0 ? tries : 1;
  const bool hde = algo == PHDE_ALGO_PARHDE || algo == PHDE_ALGO_WPARHDE;
  std::cout << "**** " << (hde ? "Means" : (algo == PHDE_ALGO_PIVOTMDS ? "PivotMDS Means" : "PHDE Means"))
            << " ******" << std::endl;
  std::cout << (hde ? "HDE mean time " : "Total mean time ") << tHDE / d << " s." << std::endl;
  std::cout << (algo == PHDE_ALGO_WPARHDE ? "SSSP mean time " : "BFS mean time ") << tDist / d << " s."
            << std::endl;
  if (hde) {
    std::cout << "DOrthogonalize mean time " << tOrth / d << " s." << std::endl;
    std::cout << "Matrix Multiply Mean time " << tMM / d << " s." << std::endl;
  } else {
    std::cout << "Matrix Multiply Mean time " << tMM / d << " s." << std::endl;
    std::cout << (algo == PHDE_ALGO_PIVOTMDS ? "Double Centering Mean time " : "ColumnCentering Mean time ")
              << tOrth / d << " s." << std::endl;
  }
  return rc;
}

// ---------------------------------------------------------------- subclasses

ParHDE::ParHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m,
               int numSources, int doRand, int numCoord)
    : DrawingBase(rowPtrs, adj, weights, n, m, numSources, doRand, numCoord),
      SLSmatrix(nullptr),
      LSmatrix(nullptr) {
  if (weights) {  // ParHDE.cpp:27-31
    std::cerr << "ParHDE code doesn't requires the weighted edges" << std::endl;
    std::exit(0);
  }
}
ParHDE::~ParHDE() {
  // null here (the products live on the GPU); freed as ParHDE.cpp:34-38 does
  free(SLSmatrix);
  free(LSmatrix);
}
int ParHDE::run(int tries, int alpha, int beta, int32_t delta) {
  printf("Running ParHDE\n");
  std::cout << "Performing ParHDE with Number of sources: " << numSources << " and ";
  std::cout << "Number of experiments: " << tries << std::endl;
  return runLayout(PHDE_ALGO_PARHDE, tries, alpha, beta, delta);
}

WParHDE::WParHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m,
                 int numSources, int numCoord)
    : DrawingBase(rowPtrs, adj, weights, n, m, numSources, 2, numCoord),
      SLSmatrix(nullptr),
      LSmatrix(nullptr) {
  if (!weights) {  // WParHDE.cpp:22-26
    std::cerr << "Weighted ParHDE code requires the weighted edges" << std::endl;
    std::exit(0);
  }
}
WParHDE::~WParHDE() {
  // null here (the products live on the GPU); freed as WParHDE.cpp:29-33 does
  free(SLSmatrix);
  free(LSmatrix);
}
int WParHDE::run(int tries, int alpha, int beta, int32_t delta) {
  printf("Running Weighted ParHDE\n");
  std::cout << "Performing parHDE with Number of sources: " << numSources << " and ";
  std::cout << "Number of experiments: " << tries << std::endl;
  return runLayout(PHDE_ALGO_WPARHDE, tries, alpha, beta, delta);
}

PivotMDS::PivotMDS(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m,
                   int numSources, int doRand, int numCoord)
    : DrawingBase(rowPtrs, adj, weights, n, m, numSources, doRand, numCoord), matrixC(nullptr) {}
PivotMDS::~PivotMDS() {
  // null here (the products live on the GPU); freed as PivotMDS.cpp:21-24 does
  free(matrixC);
}
int PivotMDS::run(int tries, int alpha, int beta, int32_t delta) {
  std::cout << "Performing PivotMDS with " << tries << " tries and pivots: " << numSources << std::endl;
  return runLayout(PHDE_ALGO_PIVOTMDS, tries, alpha, beta, delta);
}

PHDE::PHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int numSources,
           int doRand, int numCoord)
    : DrawingBase(rowPtrs, adj, weights, n, m, numSources, doRand, numCoord), matrixC(nullptr) {}
PHDE::~PHDE() {
  // null here (the products live on the GPU); freed as PHDE.cpp:21-24 does
  free(matrixC);
}
int PHDE::run(int tries, int alpha, int beta, int32_t delta) {
  std::cout << "Performing PHDE with " << tries << " tries and pivots: " << numSources << std::endl;
  return runLayout(PHDE_ALGO_PHDE, tries, alpha, beta, delta);
}
