// ops_nexi16_l2.hip -- NEXI with the l = 2 rotational invariant, up to 16 shells.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_nexi16_l2() { return ops_table<NexiModel<16, true>>(); }
}  // namespace gc
