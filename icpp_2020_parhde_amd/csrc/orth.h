// orth.h — internal interface of the D-orthogonalisation stage (orth.hip).
#pragma once
#include <utility>
#include <vector>

#include "phde_common.h"

namespace phde {

// Slot of the ones vector in the coefficient / denominator arrays
// (accepted columns use slots 0..s-1, s <= 64).
constexpr int kOnesSlot = 64;
// Candidates per block of the blocked classical Gram-Schmidt.
constexpr int kCgsBlock = 8;
// Candidates per Gram-space classical Gram-Schmidt (with the ones vector,
// at most 64 coefficients: one per lane of a wave).
constexpr int kGramMax = 63;

// Column pointers, scales and exact sums of squares of the Gram-space
// CGS's candidates (kernel arguments, passed by value).
struct GramCols {
  const int32_t *col[kGramMax];
  double scale[kGramMax];
  double sumsq[kGramMax];
};
// Per accepted column of the Gram-space CGS: the index of its own
// candidate in C = [1, cols] (its coefficient vector is zero past it).
struct GramForm {
  int last[kGramMax];
};

// Device state of one D-orthogonalisation: S (n x s, column-major, FP64 or
// FP32 storage), the candidate a (FP64), cached denominators
// den_k = sum_i S_ik^2 deg_i, and reduction scratch.
struct Orth {
  phde_ctx *ctx = nullptr;
  int64_t n = 0;
  int s = 0;
  bool fp32 = false;
  int nblk = 0;
  void *S = nullptr;
  double *a = nullptr, *part = nullptr, *den = nullptr, *coef = nullptr, *scal = nullptr;
  double *h_scal = nullptr, *ones = nullptr;
  double *blkA = nullptr, *blkAcc = nullptr, *blkPart = nullptr, *blkCoef = nullptr;  // blocked CGS
  double *gPart = nullptr, *gG = nullptr, *gM = nullptr, *gOut = nullptr, *h_gOut = nullptr;  // Gram CGS
  int *gMax = nullptr;
  int gBlocks = 0;
  // Gram-space state kept for the tail: the candidates, the accepted
  // columns' coefficient vectors (gM, [k][64]), 2-bit residues of the
  // distances (gPack, 16 B per vertex); gram_pending while S = C M is not
  // formed (every column came from the Gram CGS)
  GramCols gCols{};
  GramForm gForm{};
  int gN = 0, gAcc = 0;
  bool gram_pending = false;
  void *gPack = nullptr;
  int *gLapPart = nullptr, *gBad = nullptr;
  double *gKC = nullptr, *gW = nullptr;
  int gLapBlocks = 0;
  const double *deg_ = nullptr;
  double u0 = 0;

  // device / pinned buffers of this orthogonalisation (context temporaries)
  std::vector<std::pair<void *, size_t>> dev_bufs, host_bufs;
  int dalloc(void **p, size_t bytes);
  int halloc(void **p, size_t bytes);

  int setup(phde_ctx *c, int64_t n, int s, bool fp32);
  void release();
  ~Orth() { release(); }
  int init_ones(const double *deg, double u0);
  int load_candidate(const int32_t *dist, double scale);  // a = dist * scale
  int cgs_pass(int ncols);  // classical GS of a against ones + S[:, :ncols]; norm in scal
  int mgs_pass(int ncols);  // modified GS, same operands
  int read_norm(double *norm);  // synchronises
  int accept(int j);            // S[:, j] = a / ||a||, den_j
  // Classical GS of nc <= kCgsBlock candidates (cols[b] * scales[b]) in
  // order, each as cgs_pass + norm + discard rule + accept would treat it
  // (FP64 storage: bit for bit; FP32 storage: CGS2, the projection applied
  // twice, with the passes over the earlier columns shared by the block),
  // until s_target columns are accepted; taken[b]: candidate b accepted,
  // norms[b]: its norm after the projection.
  int cgs_block(const int32_t *const *cols, const double *scales, int nc, int s_target, int &accepted,
                std::vector<bool> &taken, std::vector<double> &norms);
  // Classical GS of nc <= kGramMax distance candidates (cols[b] scaled by
  // 1 / sqrt(sumsq[b]), sumsq[b] the exact sum of its squares) with no
  // column accepted yet, in Gram space: the Gram matrices of [1, cols] are
  // summed exactly in one pass, the reference's projections and discard
  // rule run on coefficient vectors, S is formed in one pass.  consumed:
  // candidates looked at (0 when the Gram sums could not be exact: the
  // caller then takes every candidate through cgs_block).
  int gram_cgs(const int32_t *const *cols, const double *sumsq, int nc, int s_target, int &accepted,
               std::vector<bool> &taken, std::vector<double> &norms, int &consumed);
  // S = C M for the Gram CGS's columns when not formed yet.
  int form_S();
  // With gram_pending: K_dev = S^T L S (s x s) from C^T L C, which the
  // distance columns give exactly (BFS levels differ by at most one across
  // an edge); B_dev (optional) = S^T D S; Bfull_host (optional, (s+1)^2) =
  // [u0 1, S]^T D [u0 1, S].  ok = false when some edge's levels differ by
  // more than one (the columns are then not BFS levels of this graph, and
  // the caller forms S).
  int gram_lap(double *K_dev, double *B_dev, double *Bfull_host, bool &ok);
  // With gram_pending: X = S V (n x nc, column-major) normalised as
  // coords_compute does, evaluated as C (M V).
  int gram_coords(const double *V_dev, int nc, double *X_dev, double *part, double *inv);
  // host out[k] = sum_i deg_i u0 S_ik for k < s, out[s] = sum_i deg_i u0^2
  // (the D-inner products of the ones vector); synchronises
  int ones_dots(double *out);
};

}  // namespace phde
