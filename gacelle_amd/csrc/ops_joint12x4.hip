// ops_joint12x4.hip -- joint R1-R2*, 12 echoes x 4 flip angles.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_joint12x4() { return ops_table<JointModel<12, 4, true>>(); }
}  // namespace gc
