// ops_mcrmwi4_exch_cplx.hip -- MCR-MWI, 4 flip angles, exchange (BM), complex data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_exch_cplx() { return ops_table<McrmwiModel<4, true, true>>(); }
}  // namespace gc
