// traverse.h — internal interface of the weighted and multi-source
// traversals (sssp.hip, msbfs.hip) and the weighted Gram product.
#pragma once
#include "phde_common.h"

namespace phde {

// Delta-stepping SSSP from src into dist_out (device, 2^30-1 unreached);
// optional fused k-centers update as in bfs_run.
int sssp_run(phde_ctx *c, int64_t src, int32_t delta, int32_t *dist_out, int32_t *min_dist,
             bool sync_result);
// Up to 64 BFS in one sweep; D is n*k column-major device, -1 unreached.
int msbfs_run(phde_ctx *c, const int64_t *sources, int k, int32_t *D);
void traverse_release(phde_ctx *c);
// out = S^T D S (s*s, device) for the generalised eigenproblem.
int weighted_gram(phde_ctx *c, const void *S_dev, bool fp32, int s, double *out_dev);

}  // namespace phde
