// ops_r2s32.hip -- mono-exponential R2*, up to 32 echoes.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_r2s32() { return ops_table<R2sModel<32>>(); }
}  // namespace gc
