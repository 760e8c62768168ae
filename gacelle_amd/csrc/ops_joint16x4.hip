// ops_joint16x4.hip -- joint R1-R2*, up to 16 echoes x 4 flip angles.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_joint16x4() { return ops_table<JointModel<16, 4>>(); }
}  // namespace gc
