/* phde_graph.h — host-side graph formats and seeded generators (C ABI).
 *
 * These routines sit on the caller side of the ParHDE layout path: they
 * produce the binary CSR the `embed` command reads and the synthetic
 * workloads the layout path is measured on.  No GPU is touched here.
 *
 * Binary CSR layout (reader: spectralDrawing.cpp:136-164, writer:
 * mtx2csr.cpp:281-296):
 *   int64 n, int64 m (directed nnz), int64 rest[4] = {1, 0, 1, 0},
 *   uint32 rowptr[n+1], uint32 adj[m], then int32 w[m] when weighted.
 * Neighbour lists are sorted ascending, symmetric, with no self loops and
 * no duplicates.  Largest-component extraction relabels vertices in old-id
 * order, as mtx2csr.cpp:215-269 does.
 *
 * All arrays returned inside a pg_csr are malloc'd; release with pg_free.
 */
#ifndef PHDE_GRAPH_H
#define PHDE_GRAPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pg_csr {
  int64_t n;        /* vertices */
  int64_t m;        /* directed edges (nnz) */
  uint32_t *rowptr; /* n+1 */
  uint32_t *adj;    /* m */
  int32_t *w;       /* m, or NULL when unweighted */
} pg_csr;

enum {
  PG_OK = 0,
  PG_ERR_ARG = -1,
  PG_ERR_IO = -2,
  PG_ERR_NOMEM = -3,
  PG_ERR_FORMAT = -4,
  PG_ERR_RANGE = -5 /* more than 2^32-1 edges: offsets are uint32 */
};

void pg_free(pg_csr *g);

/* 2D triangulated mesh rows x cols, vertex id = i*cols + j, edges to the
 * right (i,j+1), down (i+1,j) and down-right (i+1,j+1). */
int pg_mesh(int64_t rows, int64_t cols, pg_csr *out);

/* 3D grid nx*ny*nz with 6-neighbour edges, id = (x*ny + y)*nz + z.  When
 * wmax > 0 every undirected edge gets one int32 weight uniform in
 * [1, wmax], a pure function of (seed, min endpoint, max endpoint). */
int pg_grid3d(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, int32_t wmax,
              pg_csr *out);

/* Random geometric graph: n points uniform in [0,1)^2 drawn in vertex order
 * from xoshiro256** seeded by splitmix64(seed); an edge joins two points at
 * Euclidean distance < r, r = sqrt(avg_degree / (pi * n)).  If lcc != 0 the
 * largest connected component is kept.  xy_out (2*n doubles, optional)
 * receives the point coordinates of the kept vertices. */
int pg_rgg(int64_t n, double avg_degree, uint64_t seed, int lcc, pg_csr *out,
           double **xy_out);

/* R-MAT / Kronecker: 2^scale vertices, edgefactor * 2^scale generated edges
 * with quadrant probabilities (a, b, c, 1-a-b-c); self loops dropped,
 * symmetrised, deduplicated; largest component kept when lcc != 0. */
int pg_rmat(int scale, int edgefactor, double a, double b, double c,
            uint64_t seed, int lcc, pg_csr *out);

/* 2D grid rows x cols (4-neighbour) with `holes` seeded axis-aligned
 * rectangular holes removed, largest component kept. */
int pg_holed_grid(int64_t rows, int64_t cols, int holes, uint64_t seed,
                  pg_csr *out);

/* Build a symmetric CSR from an undirected edge list (u[i], v[i]) over n
 * vertices: self loops dropped, duplicates removed, lists sorted.  Optional
 * weights wt[i] apply to both directions (the first occurrence wins). */
int pg_from_edges(int64_t n, int64_t ne, const uint32_t *u, const uint32_t *v,
                  const int32_t *wt, pg_csr *out);

/* Largest connected component (ties: the one found first from the lowest
 * vertex), relabelled in old-id order; *n_components (may be NULL) gets the
 * number of components. */
int pg_lcc(const pg_csr *in, pg_csr *out, int64_t *n_components);

/* Matrix Market coordinate file -> CSR, read as the reference converter
 * mtx2csr.cpp:10-157 reads it: "general" square matrices as the graph of
 * A + A^T, rectangular ones as bipartite graphs (columns first), self loops
 * dropped, duplicates merged keeping the smallest weight; weights kept for
 * "integer" files only.  No component extraction.  *kind (may be NULL)
 * gets one of PG_MTX_*. */
enum { PG_MTX_PATTERN = 0, PG_MTX_REAL = 1, PG_MTX_INTEGER = 2 };
int pg_read_mtx(const char *path, pg_csr *out, int *kind);
/* The whole converter (mtx2csr.cpp main): read, largest component, write the
 * binary CSR (weights for integer files).  *n, *m (may be NULL) get the
 * written graph's size. */
int pg_mtx2csr(const char *mtx_path, const char *csr_path, int64_t *n, int64_t *m);

/* Binary CSR file I/O in the format above. */
int pg_write_csr(const char *path, const pg_csr *g);
int pg_read_csr(const char *path, int weighted, pg_csr *out);

/* PNG picture of a layout's edges, drawn as the reference's draw_graph.c
 * draws them (image size from the coordinate ranges, Bresenham lines
 * coloured by length bin; edges in order, as its single-thread run does).
 * coords: column-major n x num_coord; first / second: 0-based coordinates
 * for x and y.  *width / *height (may be NULL) get the image size. */
int pg_draw_png(const pg_csr *g, const double *coords, int num_coord, int first, int second,
                const char *path, unsigned *width, unsigned *height);

/* Checks symmetry, sortedness, no self loops / duplicates, in-range ids and
 * matching weights.  Returns PG_OK or PG_ERR_FORMAT. */
int pg_validate(const pg_csr *g);

#ifdef __cplusplus
}
#endif
#endif
