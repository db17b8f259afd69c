// ref_harness.cpp — C entry points onto the reference's own
// ParallelComponents class (variants/ParallelComponents.cpp, compiled from
// the reference tree by this directory's Makefile into _ref/).  Test
// infrastructure only.  One OpenMP thread: the reference's reductions are
// then summed in index order, the semantics the restatement follows.
#include <omp.h>

#include "ParallelComponents.h"

extern "C" int ref_power_eig(double *K, double *eValues, long n) {
  omp_set_num_threads(1);
  ParallelComponents pc;
  return pc.eigenValuePowerMethod(K, eValues, n, (int)n);
}

extern "C" double ref_norm_l2(long n, double *v) {
  omp_set_num_threads(1);
  ParallelComponents pc;
  return pc.normL2(n, v);
}
