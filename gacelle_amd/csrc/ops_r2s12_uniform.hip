// ops_r2s12_uniform.hip -- mono-exponential R2*, 12 equally spaced echoes (exponential recurrence).
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_r2s12_uniform() { return ops_table<R2sModel<12, true, true>>(); }
}  // namespace gc
