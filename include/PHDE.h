// PHDE.h — drop-in for variants/PHDE.h; run() executes on the GPU.
#ifndef _PHDE_H_
#define _PHDE_H_
#include "DrawingBase.h"

class PHDE : public DrawingBase {
 private:
  // the reference's members (variants/PHDE.h:13), kept so that the object has its
  // size and layout; the products they held live on the GPU here
  double *matrixC;

 public:
  PHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s, int r,
       int numCoord);
  ~PHDE();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
