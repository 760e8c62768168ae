// ops_r2s12.hip -- mono-exponential R2*, exactly 12 echoes.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_r2s12() { return ops_table<R2sModel<12, true>>(); }
}  // namespace gc
