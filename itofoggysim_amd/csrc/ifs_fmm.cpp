// MLFMM (placeholder until the octree and passes land).
#include "ifs_fmm.h"
#include "../../include/ifs.h"

namespace ifs {
namespace fmm {
struct Fmm {};
int build(ifs_ctx *, Fmm **out) {
    *out = nullptr;
    return IFS_E_ARG;
}
void destroy(Fmm *f) { delete f; }
int run(ifs_ctx *, Fmm *, const double2 *, double2 *) { return IFS_E_ARG; }
int tree_export(Fmm *, int64_t *, int32_t *) { return IFS_E_ARG; }
int stats(Fmm *, int64_t *, int32_t *) { return IFS_E_ARG; }
}  // namespace fmm
}  // namespace ifs
