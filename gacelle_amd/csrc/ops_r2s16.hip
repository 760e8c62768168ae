// ops_r2s16.hip -- mono-exponential R2*, up to 16 echoes.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_r2s16() { return ops_table<R2sModel<16>>(); }
}  // namespace gc
