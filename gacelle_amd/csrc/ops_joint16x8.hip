// ops_joint16x8.hip -- joint R1-R2*, up to 16 echoes x 8 flip angles.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_joint16x8() { return ops_table<JointModel<16, 8>>(); }
}  // namespace gc
