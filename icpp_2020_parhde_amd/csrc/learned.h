// learned.h — internal interface of the learned pivot proposal (learned.hip):
// node features, the message-passing network (forward, backward, Adam) and
// the greedy graph non-maximum suppression that turns scores into pivots.
#pragma once
#include "phde_common.h"

namespace phde {

constexpr int kNetHidden = 64;  // width of the three message-passing layers
constexpr int kNetIn = 4;       // input features per vertex
// Parameters (FP32, flattened in this order):
//   W1 [8][64]   rows 0-3 multiply the vertex's features, 4-7 their neighbour mean
//   b1 [64]
//   W2 [128][64] rows 0-63 multiply h1, 64-127 the neighbour mean of h1
//   b2 [64]
//   W3 [128][64], b3 [64]   (same for h2)
//   wo [64], bo [1]         read-out y = h3 . wo + bo
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + 8 * 64;
constexpr int kOffW2 = kOffB1 + 64;
constexpr int kOffB2 = kOffW2 + 128 * 64;
constexpr int kOffW3 = kOffB2 + 64;
constexpr int kOffB3 = kOffW3 + 128 * 64;
constexpr int kOffWo = kOffB3 + 64;
constexpr int kOffBo = kOffWo + 64;
constexpr int kNetParams = kOffBo + 1;

// Features of the context's graph from two traversals: the BFS from
// `first` and from the vertex farthest from it (lowest index on ties).
// Leaves x (n float4) in c->lp_feat, the two level columns in c->lp_d12
// (n each) and min(d1, d2) in c->lp_sd; returns the second source and the
// two eccentricities.  PHDE_ERR_UNREACHED when the graph is disconnected.
int pivot_features(phde_ctx *c, int64_t first, int64_t *second, int *ecc1, int *ecc2);

// Network scores of every vertex into c->lp_score (device).
int net_score(phde_ctx *c, phde_mpnn *net);

// Greedy suppression: picks[first..K-1] = repeatedly the best-scoring
// vertex farther than r hops from every pick so far (if none, the farthest
// within r), with c->lp_sd holding the distances to picks[0..first-1] on
// entry.  picks (host) must hold K entries; picks[0..first-1] are inputs.
int greedy_select(phde_ctx *c, const float *score_dev, int first, int K, int r, int64_t *picks);
// greedy_select's status when the device cannot run its persistent kernel
// (the learned layout then falls back to k-centers)
constexpr int kSelectUnavailable = -100;

// Exact per-column sums of squares of n*k int32 columns, and the count of
// unreached entries (-1, or >= 2^30-1).  out: host, k + 1 entries.
int column_sumsq(phde_ctx *c, const int32_t *D, int k, double *sumsq, unsigned long long *unreached);

void learned_release(phde_ctx *c);

}  // namespace phde
