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
// One search from k seeds (internal ids) together, each at level 0: col
// gets every vertex's distance to the nearest seed and min_dist the least
// of that and its own value -- exact up to bound; a distance above bound
// leaves a value above bound, or none (kUnvisited: no bound).  *ran is
// false when the context does not run the boundary-distance method with
// its distance tables (the caller takes another search); the searches of
// the groups before a decline have then already lowered min_dist, which
// another search from all k seeds leaves correct.
int pbfs_seeds_run(phde_ctx *c, const int64_t *seeds, int k, int32_t *col, int32_t *min_dist, int32_t bound,
                   bool *ran);
// Whether pbfs_seeds_run can run on this context.
bool pbfs_seeds_ok(const phde_ctx *c);

}  // namespace phde
