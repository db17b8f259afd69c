/* parhde.h — C ABI of the MI355X ParHDE layout library (libparhde.so).
 *
 * This is the drop-in boundary for the reference's layout path.  The
 * reference has no C ABI of its own: its entry points are the C++ class
 * `DrawingBase` (variants/DrawingBase.h:13-48) with subclasses ParHDE,
 * WParHDE, PivotMDS and PHDE, and the two traversal functions DOBFS
 * (gapbs/bfs.cc:144) and DeltaStep (gapbs/sssp.cc:53), declared at
 * variants/DrawingBase.h:62-67.  The C++ layer shipped with this library
 * (include/DrawingBase.h) re-exposes those classes and functions on top of
 * the calls below; a foreign-language caller binds these calls directly.
 *
 * Conventions
 *  - Graphs are symmetric CSR with uint32 offsets and ids (the binary CSR
 *    format of spectralDrawing.cpp:136-164); weights are int32 or NULL.
 *  - Host pointers unless a name ends in _dev (device pointers, HBM).
 *  - Matrices are column-major, as in the reference (n*c coordinates,
 *    column k at offset k*n).
 *  - Every call returns a phde_status; 0 is success.  No call exits the
 *    process: the C++ layer maps statuses onto the reference's exit codes.
 *  - A context is single-stream and not reentrant: one thread at a time.
 */
#ifndef PARHDE_H
#define PARHDE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum phde_status {
  PHDE_OK = 0,
  PHDE_ERR_ARG = -1,       /* bad argument (size, range, NULL) */
  PHDE_ERR_DEVICE = -2,    /* HIP runtime error or no usable GPU */
  PHDE_ERR_NOMEM = -3,     /* device allocation failed */
  PHDE_ERR_PIVOTS = -4,    /* fewer than s independent pivots (ParHDE.cpp:281-285) */
  PHDE_ERR_EIG = -5,       /* eigensolver failed (ParallelComponents.cpp:77-81) */
  PHDE_ERR_WEIGHTS = -6,   /* ParHDE given weights / WParHDE without (ParHDE.cpp:27-31, WParHDE.cpp:22-26) */
  PHDE_ERR_UNREACHED = -7, /* graph not connected: a traversal left vertices unreached */
  PHDE_ERR_UNSUPPORTED = -8
} phde_status;

typedef struct phde_ctx phde_ctx;
/* Learned pivot proposal network (learned-pivot strategy). */
typedef struct phde_mpnn phde_mpnn;

/* Algorithms (spectralDrawing.cpp:101-115 `-a`, :177-186). */
enum { PHDE_ALGO_PIVOTMDS = 0, PHDE_ALGO_PHDE = 1, PHDE_ALGO_WPARHDE = 2, PHDE_ALGO_PARHDE = 3 };
/* Pivot strategies (spectralDrawing.cpp:88-95 `-p`). */
enum { PHDE_PIVOTS_RANDCOARSE = 0, PHDE_PIVOTS_RANDFINE = 1, PHDE_PIVOTS_KCENTERS = 2, PHDE_PIVOTS_LEARNED = 3 };
/* Arithmetic of the D-orthogonalisation / SpMM stages. */
enum { PHDE_PREC_FP64 = 0, PHDE_PREC_FP32 = 1 };
/* Eigensolver: the reference's power method (ParallelComponents.cpp:48-158),
 * an exact Jacobi solve of the same standard problem, or the generalised
 * problem (S^T L S) x = lambda (S^T D S) x by Cholesky + Jacobi. */
enum { PHDE_EIG_POWER = 0, PHDE_EIG_JACOBI = 1, PHDE_EIG_GENERALIZED = 2 };

typedef struct phde_options {
  int32_t algorithm;   /* PHDE_ALGO_*; default PARHDE */
  int32_t s;           /* pivots, 1..64 (reference: 1..50, ParHDE.cpp:15) */
  int32_t num_coord;   /* coordinates c, 1..s */
  int32_t strategy;    /* PHDE_PIVOTS_*; default KCENTERS */
  int32_t precision;   /* PHDE_PREC_*; default FP64 */
  int32_t eig_mode;    /* PHDE_EIG_*; default POWER */
  int32_t eager;       /* 1: run all 3s traversals like ParHDE.cpp:120; 0: only as many as consumed */
  int32_t delta;       /* delta-stepping bucket width; <= 0 picks one from the weights */
  int32_t alpha, beta; /* direction-switch parameters (bfs.cc:161,172); defaults 15, 18 */
  int32_t verbose;     /* 1: print the reference's progress lines to stdout */
  int64_t first_pivot; /* first k-centers source; reference: 0 */
  uint64_t seed;       /* random strategies: srand(seed) then rand() % n per source */
  /* learned strategy */
  double judge_tol;    /* with judge: keep k-centers when E_learned > (1 + judge_tol) E_kcenters; default 0.02 */
  double nms_rho;      /* suppression radius r = nms_rho * ecc / sqrt(s) hops; <= 0 picks 1.0 */
  int32_t judge;       /* 1: also run the k-centers layout and keep the better one (see judge_tol) */
  int32_t diagnostics; /* 1: also measure the D-orthogonality error of S (phde_stats.dorth_err) */
} phde_options;

typedef struct phde_stats {
  double ms_total;       /* device time of the layout (HDE mean time semantics) */
  double ms_distance;    /* traversals + k-centers update + column norms */
  double ms_orth;        /* D-orthogonalisation */
  double ms_mmult;       /* L*S and S^T(LS) */
  double ms_eig;         /* s*s eigensolve */
  double ms_coords;      /* S*V + normalisation */
  int32_t n_traversals;  /* BFS / SSSP runs */
  int32_t n_accepted;
  int32_t n_discarded;
  int32_t eig_iters;
  int64_t levels_total;  /* sum of BFS depths (level-synchronous steps) */
  int64_t pivots[192];   /* sources in traversal order (first n_traversals valid) */
  double  kernel_ms[16]; /* per-kernel-class device time, see PHDE_KERNEL_* */
  int64_t kernel_launches[16];
  /* learned strategy */
  double ms_select;            /* features, network scores and pivot selection */
  double hall_energy;          /* tr(Y^T L Y) / tr(Y^T D Y) of the returned layout (learned strategy) */
  double hall_energy_kcenters; /* of the k-centers layout when judged, else 0 */
  int32_t fallback;            /* 1 when the k-centers layout was returned instead */
  int32_t n_candidates;        /* pivots proposed by the selection */
  /* numerical diagnostics (ParHDE / WParHDE) */
  int32_t n_examined;          /* candidates taken through the orthogonalisation (<= 192 recorded) */
  int32_t n_gram;              /* of those, candidates taken through the Gram-space CGS (FP64 ParHDE) */
  double normdist[192];        /* per examined candidate, in order: ||a||_2 after the projection, the
                                  value of the discard test ||a|| < 0.001 (ParHDE.cpp:267-283) */
  double eigenvalues[64];      /* the s eigenvalue estimates of the s*s eigensolve, in the solver's
                                  output order (power method: after its reversal, ParallelComponents.cpp:132-140) */
  double dorth_err;            /* with options.diagnostics: max_ij |B_ij / sqrt(B_ii B_jj) - delta_ij| of
                                  B = [1 S]^T D [1 S] (D = degrees); 0 otherwise */
} phde_stats;

enum {
  PHDE_KERNEL_BFS_LEVEL = 0,
  PHDE_KERNEL_KCENTER = 1,
  PHDE_KERNEL_ORTH = 2,
  PHDE_KERNEL_SPMM = 3,
  PHDE_KERNEL_GRAM = 4,
  PHDE_KERNEL_EIG = 5,
  PHDE_KERNEL_COORDS = 6,
  PHDE_KERNEL_SSSP = 7,
  PHDE_KERNEL_MSBFS = 8,
  PHDE_KERNEL_OTHER = 9
};

/* Fill *o with the reference defaults: ParHDE, s=10, c=2, k-centers, FP64,
 * power method, lazy traversals, alpha 15, beta 18, first pivot 0. */
void phde_default_options(phde_options *o);

/* Library / device information. */
const char *phde_version(void);
int phde_device_count(int *count);

/* Create a context holding the graph in HBM.  Replaces the borrowed graph
 * arrays of the DrawingBase constructor (variants/DrawingBase.h:41,
 * DrawingBase.cpp:16-37); the host arrays may be freed afterwards. */
int phde_create(const uint32_t *rowptr, const uint32_t *adj, const int32_t *weights, int64_t n,
                int64_t m, phde_ctx **out);
/* Same, from device-resident arrays (copied device-to-device). */
int phde_create_device(const uint32_t *rowptr_dev, const uint32_t *adj_dev,
                       const int32_t *weights_dev, int64_t n, int64_t m, phde_ctx **out);
int phde_destroy(phde_ctx *ctx);
int64_t phde_num_vertices(const phde_ctx *ctx);
/* The context's internal vertex order (no reference counterpart).  A
 * context renumbers its copy of a high-diameter graph (BFS depth from
 * vertex 0 >= 64, n >= 65,536; environment PHDE_RELABEL=0/1 forces off/on)
 * in the order of (level from vertex 0, level from the vertex farthest
 * from 0, id); every argument and result of this API stays in the
 * caller's numbering.  order (n entries, may be NULL) gets the caller's
 * ids in internal order; *renumbered (may be NULL) 1 or 0. */
int phde_vertex_order(const phde_ctx *ctx, uint32_t *order, int *renumbered);

/* Parts of the context's partitioned traversal (no reference counterpart):
 * an unweighted graph with a deep BFS is cut at creation into parts that
 * each fit one compute unit's local memory, and its traversals run part by
 * part (same levels).  *parts = 0 when the context does not use it
 * (low-diameter or weighted graphs, PHDE_PBFS=0). */
int phde_num_parts(const phde_ctx *ctx, int64_t *parts);
/* Boundary vertices of those parts when the context's traversals run by
 * the boundary-distance method (no reference counterpart): each part's
 * boundary-to-boundary distances inside the part are computed at creation,
 * and a traversal propagates levels between parts over those distances
 * before filling in each part's interior (same levels).  *boundary = -1
 * when the context does not use it (as above, PHDE_OBFS=0, or an in-part
 * distance of 255 or more). */
int phde_num_boundary(const phde_ctx *ctx, int64_t *boundary);

/* Breadth-first levels from one source: DOBFS (gapbs/bfs.cc:144-184).
 * dist[v] = hop count, -1 when unreached.  alpha/beta <= 0 use 15/18. */
int phde_bfs(phde_ctx *ctx, int64_t source, int alpha, int beta, int32_t *dist);
int phde_bfs_dev(phde_ctx *ctx, int64_t source, int alpha, int beta, int32_t *dist_dev);

/* Shortest-path distances from one source: DeltaStep (gapbs/sssp.cc:53-140).
 * dist[v] = exact distance, 2^30-1 when unreached (sssp.cc:50). */
int phde_sssp(phde_ctx *ctx, int64_t source, int32_t delta, int32_t *dist);

/* Up to 64 breadth-first searches in one sweep (bit-parallel, one 64-bit
 * mask per vertex).  dist is n*k column-major; -1 when unreached.  No
 * reference counterpart: it equals k calls of phde_bfs. */
int phde_msbfs(phde_ctx *ctx, const int64_t *sources, int k, int32_t *dist);

/* One complete layout: DrawingBase::run + getCoordinates (ParHDE.cpp:45-352,
 * WParHDE.cpp:40-291, PivotMDS.cpp:68-147, PHDE.cpp:50-136,
 * DrawingBase.cpp:196-205).  coords: n*num_coord column-major, each column
 * of unit L2 norm.  stats optional. */
int phde_layout(phde_ctx *ctx, const phde_options *opt, double *coords, phde_stats *stats);

/* The reference's subspace power method (ParallelComponents.cpp:48-158) on
 * one s*s matrix, run by one GPU workgroup.  K (column-major, dim*dim) is
 * overwritten with the eigenvector matrix as the reference does; evals gets
 * the eigenvalue estimates.  iters optional. */
int phde_power_eig(phde_ctx *ctx, double *K, double *evals, int dim, int32_t *iters);

/* The power method of phde_power_eig on device-resident data, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc or
 * hipFree: its scratch (the start values, 32 KiB) is a
 * static device array of the library, one per GPU, held until the process
 * ends, so calls that overlap on one GPU race on it.  K_dev (dim*dim,
 * column-major) is overwritten with the eigenvector matrix, evals_dev (dim)
 * gets the eigenvalue estimates, status_dev (2 int32) the status (0 ok,
 * 1 NaN/Inf, 2 iteration cap) and the iteration count, all device
 * pointers.  dim <= 64, else PHDE_ERR_ARG.  Returns once the results are
 * written (synchronises the context's stream). */
int phde_power_eig_noalloc(phde_ctx *ctx, double *K_dev, double *evals_dev, int dim, int32_t *status_dev);

/* Exact symmetric eigensolve (cyclic Jacobi, one workgroup) of the
 * symmetric part of A (dim*dim).  With B != NULL, the generalised problem
 * A x = lambda B x (B symmetric positive definite, Cholesky-reduced).
 * evals ascending, evecs column-major, B-orthonormal columns. */
int phde_sym_eig(phde_ctx *ctx, const double *A, const double *B, int dim, double *evals,
                 double *evecs);

/* Hall energy tr(Y^T L Y) / tr(Y^T D Y) of an n*c column-major layout. */
int phde_hall_energy(phde_ctx *ctx, const double *Y, int c, double *energy);

/* ---------------------------------------------------------------- learned
 * pivots.  No reference counterpart: a message-passing network scores the
 * vertices from features of two traversals, a greedy graph non-maximum
 * suppression picks up to 64 spread-out pivots, and one multi-source BFS
 * gives their distance columns (phde_layout with PHDE_PIVOTS_LEARNED and
 * PHDE_ALGO_PARHDE).  The network is trained on the GPU with the calls
 * below against k-centers coverage fields. */

/* Number of FP32 parameters of the network. */
int phde_net_num_params(void);
/* New network on `device` with seeded Glorot-uniform weights. */
int phde_net_create(int device, uint64_t seed, phde_mpnn **out);
int phde_net_destroy(phde_mpnn *net);
int phde_net_get_params(phde_mpnn *net, float *params);
/* Also resets the optimiser state. */
int phde_net_set_params(phde_mpnn *net, const float *params);
/* Weight file: "PHDENET1", int32 count, FP32 parameters. */
int phde_net_save(phde_mpnn *net, const char *path);
int phde_net_load(phde_mpnn *net, const char *path);
/* Attach (borrowed) or detach (NULL) the network used by PHDE_PIVOTS_LEARNED. */
int phde_set_pivot_net(phde_ctx *ctx, phde_mpnn *net);

/* Node features (n*4, row-major): log(1+deg)/4, BFS level from `first` /
 * its eccentricity, BFS level from the vertex farthest from `first` / its
 * eccentricity, 1.  sources (optional, 2 entries) gets the two sources. */
int phde_pivot_features(phde_ctx *ctx, int64_t first, float *feat, int64_t *sources);
/* Training data of this context's graph: features (n*4) and target (n). */
int phde_set_training_data(phde_ctx *ctx, const float *feat, const float *target);
/* Network scores (n) from the context's features. */
int phde_net_scores(phde_mpnn *net, phde_ctx *ctx, float *score);
/* Mean squared error against the target and its gradient (optional). */
int phde_net_gradient(phde_mpnn *net, phde_ctx *ctx, double *loss, float *grad);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradient. */
int phde_net_adam(phde_mpnn *net, phde_ctx *ctx, float lr);
/* phde_net_gradient + phde_net_adam; loss before the step. */
int phde_net_train_step(phde_mpnn *net, phde_ctx *ctx, float lr, double *loss);
/* Greedy suppression on given scores: pivots[0] = first, pivots[1] = the
 * vertex farthest from it, then k-2 picks each the best-scoring vertex more
 * than `radius` hops from all previous picks (if none is left, the one
 * farthest from them).  2 <= k <= 64. */
int phde_select_pivots(phde_ctx *ctx, const float *score, int64_t first, int k, int radius,
                       int64_t *pivots);
/* k-centers alone (farthest-first, lowest index on ties, first source
 * `first`): s sources and the final distance-to-nearest-source field. */
int phde_kcenters(phde_ctx *ctx, int s, int64_t first, int64_t *sources, int32_t *min_dist);

/* Last error text of this thread (empty when none). */
const char *phde_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
