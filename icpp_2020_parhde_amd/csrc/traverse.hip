// traverse.hip — weighted and multi-source traversals (placeholders until
// their kernels land) and the weighted Gram product S^T D S.
#include "phde_common.h"
#include "traverse.h"

namespace phde {

int sssp_run(phde_ctx *, int64_t, int32_t, int32_t *, int32_t *, bool) {
  set_error("delta-stepping SSSP not built yet");
  return PHDE_ERR_UNSUPPORTED;
}

int msbfs_run(phde_ctx *, const int64_t *, int, int32_t *) {
  set_error("multi-source BFS not built yet");
  return PHDE_ERR_UNSUPPORTED;
}

void traverse_release(phde_ctx *) {}

int weighted_gram(phde_ctx *, const void *, bool, int, double *) {
  set_error("generalised eigen mode not built yet");
  return PHDE_ERR_UNSUPPORTED;
}

}  // namespace phde
