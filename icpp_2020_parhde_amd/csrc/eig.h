// eig.h — internal interface of the s x s eigensolvers (eig.hip).
#pragma once
#include "phde_common.h"

namespace phde {

// Reference power method on K_dev (dim*dim, dim <= 64, overwritten).
// init_dev holds the dim*dim start values; status_dev[0] = 0 ok, 1
// NaN/Inf, 2 iteration cap; status_dev[1] = iterations.
int power_eig_dev(phde_ctx *c, double *K_dev, const double *init_dev, double *evals_dev, int dim,
                  int max_iter, int *status_dev);
// Exact (generalised when B_dev != nullptr) symmetric eigensolve; work_dev
// 3*dim*dim doubles; status_dev[0] = 1 when B is not positive definite.
int sym_eig_dev(phde_ctx *c, const double *A_dev, const double *B_dev, int dim, double *evals_dev,
                double *evecs_dev, double *work_dev, int *status_dev);
// First `count` outputs of the C library's rand() after srand(seed).
void glibc_rand_sequence(unsigned seed, int count, int32_t *out);

}  // namespace phde
