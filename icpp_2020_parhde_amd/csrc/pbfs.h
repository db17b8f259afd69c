// pbfs.h — internal interface of the partitioned BFS (pbfs.hip).
#pragma once
#include "phde_common.h"

namespace phde {

// Cuts the context's graph into parts that fit one CU's LDS (once, after
// the internal numbering exists).  Leaves c->pg null when the graph cannot
// be cut (a vertex too wide for a part) or PHDE_PBFS=0.
int pbfs_setup(phde_ctx *c);
void pbfs_release(phde_ctx *c);
// Parts of the context's graph, 0 when the partitioned BFS is not in use.
int64_t pbfs_num_parts(const phde_ctx *c);
// Boundary vertices when the traversals run by the boundary-distance
// method, -1 otherwise.
int64_t pbfs_num_boundary(const phde_ctx *c);
// One traversal from src (internal id) with the fused k-centers step (as
// bfs_run; c->h_kc holds the result).  *ran is false when the context does
// not use the partitioned BFS, or a wait in it timed out (it then stays
// off for this context and the caller takes another traversal).
int pbfs_run(phde_ctx *c, int64_t src, int32_t *dist_out, int32_t *min_dist, bool *ran);
// Levels from k <= 64 sources (internal ids) searched together by the
// boundary-distance method: D gets n levels per source (internal ids,
// kUnvisited when unreached).  *ran is false when the context does not use
// that method (the caller takes another traversal).
int pbfs_multi_run(phde_ctx *c, const int64_t *src, int k, int32_t *D, bool *ran);

}  // namespace phde
