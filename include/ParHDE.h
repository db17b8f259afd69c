// ParHDE.h — drop-in for variants/ParHDE.h; run() executes on the GPU.
#ifndef _PARHDE_H_
#define _PARHDE_H_
#include "DrawingBase.h"

class ParHDE : public DrawingBase {
 private:
  // the reference's members (variants/ParHDE.h:13-14), kept so that the object has its
  // size and layout; the products they held live on the GPU here
  double *SLSmatrix;
  double *LSmatrix;

 public:
  ParHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s, int r,
       int numCoord);
  ~ParHDE();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
