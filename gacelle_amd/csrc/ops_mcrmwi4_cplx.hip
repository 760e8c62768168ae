// ops_mcrmwi4_cplx.hip -- MCR-MWI, 4 flip angles, no exchange (Bloch), complex data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_cplx() { return ops_table<McrmwiModel<4, false, true>>(); }
}  // namespace gc
