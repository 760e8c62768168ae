// ops_mcrmwi4_epg_cplx.hip -- MCR-MWI, 4 flip angles, EPG-X MLP T1 path, complex data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_epg_cplx() { return ops_table<McrmwiModel<4, true, true, true>>(); }
}  // namespace gc
