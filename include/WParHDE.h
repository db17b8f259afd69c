// WParHDE.h — drop-in for variants/WParHDE.h; run() executes on the GPU.
// numCoord is honoured (the reference forwards it into doRand and always
// computes two coordinates, WParHDE.cpp:12-13).
#ifndef _WPAR_HDE_H_
#define _WPAR_HDE_H_
#include "DrawingBase.h"

class WParHDE : public DrawingBase {
 private:
  // the reference's members (variants/WParHDE.h:13-14), kept so that the object has its
  // size and layout; the products they held live on the GPU here
  double *SLSmatrix;
  double *LSmatrix;

 public:
  WParHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s,
          int numCoord);
  ~WParHDE();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
