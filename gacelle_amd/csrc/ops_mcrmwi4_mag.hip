// ops_mcrmwi4_mag.hip -- MCR-MWI, 4 flip angles, no exchange (Bloch), magnitude data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_mag() { return ops_table<McrmwiModel<4, false, false>>(); }
}  // namespace gc
