// PivotMDS.h — drop-in for variants/PivotMDS.h; run() executes on the GPU.
#ifndef _PIVOTMDS_H_
#define _PIVOTMDS_H_
#include "DrawingBase.h"

class PivotMDS : public DrawingBase {
 private:
  // the reference's members (variants/PivotMDS.h:13), kept so that the object has its
  // size and layout; the products they held live on the GPU here
  double *matrixC;

 public:
  PivotMDS(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s, int r,
       int numCoord);
  ~PivotMDS();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
