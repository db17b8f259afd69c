/* oracle.h — CPU restatement of the reference ParHDE layout path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * C++ class layer, the embed CLI) links or calls this code.  It is used by
 * tests/ as the parity checker, by __graft_entry__.smoke() as the checker of
 * one small layout, and by bench.py as the timed CPU baseline.
 *
 * Every function restates a reference routine (file:line given beside it)
 * with the reference's operation order, so that FP64 outputs can be
 * compared to the GPU path and to the reference.  Parity pins are listed in
 * oracle.cpp's header.
 */
#ifndef PHDE_ORACLE_H
#define PHDE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_graph {
  int64_t n, m;
  const uint32_t *rowptr;
  const uint32_t *adj;
  const int32_t *w; /* NULL when unweighted */
} or_graph;

/* Outputs of one layout.  Every pointer is optional (NULL = not wanted) and
 * caller-allocated with the size given. */
typedef struct or_result {
  double *coords;       /* n*c column-major, after getCoordinates()'s column normalisation */
  double *coords_raw;   /* n*c column-major, S*V before normalisation */
  double *S;            /* n*s column-major D-orthogonalised pivot columns */
  double *SLS;          /* s*s as handed to the eigensolver (before it is overwritten) */
  double *eigvec;       /* s*c eigenvectors used for the coordinates */
  double *eigval;       /* s eigenvalue estimates from the power method (after reversal) */
  int64_t *pivots;      /* up to 3s sources in traversal order */
  int32_t *accepted;    /* up to 3s: 1 accepted, 0 discarded, -1 not examined */
  double *normdist;     /* up to 3s: the L2 norm each candidate had after orthogonalisation */
  int32_t n_bfs;        /* traversals run */
  int32_t n_accepted;   /* accepted candidates */
  int32_t eig_iters;    /* power-method iterations */
  double t_hde, t_dist, t_orth, t_mmult; /* seconds, the reference's phase timers */
} or_result;

enum { OR_OK = 0, OR_ERR_ARG = -1, OR_ERR_PIVOTS = -2, OR_ERR_EIG = -3, OR_ERR_WEIGHTS = -4 };

/* Textbook queue BFS (DrawingBase.cpp:246-271 shape).  dist[v] = level, -1 if unreached. */
void or_seq_bfs(const or_graph *g, int64_t src, int32_t *dist);
/* Direction-optimising BFS (gapbs/bfs.cc:134-184), levels written as
 * doubles to reached vertices only. */
void or_dobfs(const or_graph *g, int64_t src, double *dists, int alpha, int beta);
/* Delta-stepping SSSP (gapbs/sssp.cc:53-140); unreached = 2^30-1. */
void or_deltastep(const or_graph *g, int64_t src, int32_t delta, double *dists);
/* Textbook binary-heap Dijkstra; unreached = -1. */
void or_dijkstra(const or_graph *g, int64_t src, int64_t *dist);

/* Subspace power method (ParallelComponents.cpp:48-158).  K is n*n
 * column-major and is overwritten with the eigenvectors; returns 0, or 1
 * when the iteration produced NaN/Inf.  iters (optional) gets the count. */
int or_power_eig(double *K, double *eValues, int64_t n, int32_t *iters);

/* The reference eigen wrappers (DrawingBase.cpp:51-108): which = 0 First,
 * 1 Largest; eigvec is dim*numCoord. */
int or_compute_eigens(double *K, int64_t dim, int numCoord, int which, double *eigvec,
                      double *eigval, int32_t *iters);

/* Full layouts.  lazy != 0 runs traversals only until s candidates are
 * accepted (bit-identical outputs, fewer traversals); lazy == 0 is the
 * reference's eager 3s traversals.  delta is used for weighted graphs. */
int or_parhde(const or_graph *g, int s, int numCoord, int lazy, or_result *r);
/* ParHDE with the candidate sources given in order (random strategies,
 * ParHDE.cpp:158-172, and learned proposals): CGS until s are accepted. */
int or_parhde_sources(const or_graph *g, int s, int numCoord, const int64_t *sources, int nsrc,
                      or_result *r);
int or_wparhde(const or_graph *g, int s, int numCoord, int32_t delta, or_result *r);
int or_pivotmds(const or_graph *g, int s, int numCoord, or_result *r);
int or_phde(const or_graph *g, int s, int numCoord, or_result *r);

/* Bounded CPU-baseline sample of one ParHDE layout on the full graph:
 * `nbfs` k-centers traversals are timed (each with its argmax and column
 * normalisation, the reference's "BFS mean time" work), the CGS step is
 * timed at three numbers of accepted columns and interpolated linearly over
 * the s candidates, and L*S + S^T(LS), the power method and the
 * coordinates are timed once at full size on stand-in columns (their cost
 * does not depend on the values).  Seconds. */
typedef struct or_sample {
  double t_bfs_each;   /* one k-centers traversal */
  double t_orth_total; /* CGS over s accepted candidates */
  double t_mmult;      /* L*S and S^T(LS) */
  double t_eig;        /* power method */
  double t_coords;     /* S*V + normalisation */
  double t_spent;      /* CPU wall time the sample took */
  int32_t n_bfs_timed;
} or_sample;
int or_parhde_sample(const or_graph *g, int s, int numCoord, int nbfs, or_sample *out);

/* Hall energy tr(Y^T L Y) / tr(Y^T D Y) of an n*c column-major layout. */
double or_hall_energy(const or_graph *g, const double *Y, int c);

/* OpenMP threads the oracle will use. */
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
