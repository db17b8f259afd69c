// ParallelComponents.h — drop-in for variants/ParallelComponents.h:1-15.
// eigenValuePowerMethod runs the reference's power method on the GPU
// (phde_power_eig); normL2 / dotProd are the reference's small host helpers.
#ifndef _PARALLEL_COMPONENTS_H_
#define _PARALLEL_COMPONENTS_H_
#include <cmath>

class ParallelComponents {
 public:
  ParallelComponents();
  double normL2(long n, double *vec);
  double dotProd(double *x, double *y, int n);
  int eigenValuePowerMethod(double *K, double *eValues, long n, int s);
};

#endif
