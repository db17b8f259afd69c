// WParHDE.h — drop-in for variants/WParHDE.h; run() executes on the GPU.
// numCoord is honoured (the reference forwards it into doRand and always
// computes two coordinates, WParHDE.cpp:12-13).
#ifndef _WPAR_HDE_H_
#define _WPAR_HDE_H_
#include "DrawingBase.h"

class WParHDE : public DrawingBase {
 public:
  WParHDE(unsigned int *rowPtrs, unsigned int *adj, int32_t *weights, long n, long m, int s,
          int numCoord);
  ~WParHDE();
  virtual int run(int tries, int alpha, int beta, int32_t delta);
};

#endif
