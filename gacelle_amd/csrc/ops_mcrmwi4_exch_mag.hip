// ops_mcrmwi4_exch_mag.hip -- MCR-MWI, 4 flip angles, exchange (BM), magnitude data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_exch_mag() { return ops_table<McrmwiModel<4, true, false>>(); }
}  // namespace gc
