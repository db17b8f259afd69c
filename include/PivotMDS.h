// PivotMDS.h — drop-in for variants/PivotMDS.h; run() executes on the GPU.
#ifndef _PIVOTMDS_H_
#define _PIVOTMDS_H_
#include "DrawingBase.h"

class PivotMDS : public DrawingBase {
 public:
  PivotMDS(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s, int r,
       int numCoord);
  ~PivotMDS();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
