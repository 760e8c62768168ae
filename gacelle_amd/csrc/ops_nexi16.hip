// ops_nexi16.hip -- NEXI spherical mean, up to 16 shells.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_nexi16() { return ops_table<NexiModel<16, false>>(); }
}  // namespace gc
