// mds.h — internal interface of the PivotMDS / PHDE arithmetic (mds.hip).
// D: n x s int32 distance columns (column-major) in the caller's
// numbering; every call is enqueued on the context stream.
#pragma once
#include "phde_common.h"

namespace phde {

// Exact 128-bit column sums (of d^2 when squares) into acc[2 i], acc[2 i + 1]
// (acc zeroed by the caller); with squares also the row sums of d^2 over
// the columns in order into rowterm (PivotMDS.cpp:39-50).
int mds_center_sums(phde_ctx *c, const int32_t *D, int64_t n, int s, bool squares, double *rowterm,
                    unsigned long long *acc);
// The sequential j-ascending double column sums (of d^2 when squares).
int mds_column_sums_seq(phde_ctx *c, const int32_t *D, int64_t n, int s, bool squares, double *out);
// C = the doubly centred -0.5 d^2 matrix (dbl, PivotMDS.cpp:51-63) or the
// column-centred distances (PHDE.cpp:30-47).
int mds_center(phde_ctx *c, const int32_t *D, int64_t n, int s, const double *colterm, const double *rowterm,
               double rN, double rK, double total, bool dbl, double *C);
// K = C^T C, one sequential j-ascending chain per entry (DrawingBase.cpp:221-244).
int mds_ctc(phde_ctx *c, const double *C, int64_t n, int s, double *K);

}  // namespace phde
