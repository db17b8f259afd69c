// DrawingBase.h — drop-in C++ interface of the reference's layout classes
// (variants/DrawingBase.h:1-68), implemented on the MI355X library
// (parhde.h).  Same class names, constructor and run() signatures, enum and
// free functions, so code written against the reference compiles unchanged;
// the objects also have the reference's layout (data members, their order,
// a non-virtual destructor, run() in vtable slot 0), so code compiled
// against the reference's headers links against this library.  The GPU
// state of an object (context, options, statistics) lives in a table keyed
// by the object's address, not in the object.
//
// Differences, all deliberate:
//  - the constructor uploads the graph to the GPU once; the caller's arrays
//    are still borrowed (not freed) as in the reference;
//  - run() returns the phde_status of the last try instead of always 0; the
//    reference's fatal paths keep their exit codes (exit(0) for weight and
//    pivot errors, exit(1) for eigensolver failure);
//  - options() exposes the GPU choices (precision, eigen mode, eager/lazy
//    traversals, learned pivots); stats() the per-phase times of the last
//    try; setPivotNet() attaches the learned-pivot network.
#ifndef _DRAWING_BASE_H_
#define _DRAWING_BASE_H_
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>

#include "ParallelComponents.h"
#include "parhde.h"

enum EigenChoice { First, Largest };

class DrawingBase {
 protected:
  unsigned int *rowPtrs;
  unsigned int *adj;
  int32_t *weights;
  long n;
  long m;
  int numSources;
  int numCoord;
  ParallelComponents parComp;
  double *eigenVectors;
  double *coordinates;
  const int ALIGN = 64;
  int doRand;

  // DrawingBase.cpp:221-244: C[k + i*m] = sum_j A[j + i*k] B[j + k*k]
  void matrixTransMultiplication(double *A, double *B, double *C, long m, long k, long n);
  // DrawingBase.cpp:119-194: numSources k-centers (or random) traversals
  // into *distMatrix (n x numSources, column-major, levels as doubles).
  void distancePhaseComputation(double **, double *, int, int, int, int32_t);
  // DrawingBase.cpp:110-117: BFS when weights == nullptr, else SSSP.
  void DistanceAlgorithm(unsigned int *row, unsigned int *col, long source, long numNodes,
                         long numEdges, double *dists, int32_t *weights = nullptr,
                         int32_t delta = 1, int alpha = 15, int beta = 18);
  // DrawingBase.cpp:207-219: coordinates = mat * eigenVectors.
  void computeCoordinates(double *);
  // DrawingBase.cpp:246-271: levels from `start` into distCol.
  void seqBfs(unsigned int start, double *distCol);
  // Runs `tries` layouts of algorithm `algo` on the GPU and prints the
  // reference's mean-time lines.
  int runLayout(int algo, int tries, int alpha, int beta, int32_t delta);

 public:
  DrawingBase(unsigned int *, unsigned int *, int32_t *, long, long, int, int, int numCoord = 2);
  ~DrawingBase();  // non-virtual, as variants/DrawingBase.h:42

  double *getCoordinates();

  void compute_eigens(double *, int, double *, EigenChoice);
  virtual int run(int tries = 11, int alpha = 15, int beta = 18, int32_t delta = 1) = 0;

  phde_options &options();
  // Network used by the learned pivot strategy (doRand 3 or
  // options().strategy = PHDE_PIVOTS_LEARNED); borrowed.
  int setPivotNet(phde_mpnn *net);
  const phde_stats &stats() const;
};

struct Compare {
  double val;
  long index;
};

// gapbs/bfs.cc:144 and gapbs/sssp.cc:53, same C++ signatures and output
// conventions (levels written as doubles, only reached vertices written;
// SSSP writes every vertex, 2^30-1 when unreached, and returns dist).
extern void DOBFS(unsigned int *row, unsigned int *col, long source, long numNodes, long numEdges,
                  double *dists, int alpha = 15, int beta = 18);
extern double *DeltaStep(const unsigned int *row, unsigned int *col, int32_t *weights, long source,
                         long numNodes, long numEdges, double *dist, int32_t delta);

#endif
