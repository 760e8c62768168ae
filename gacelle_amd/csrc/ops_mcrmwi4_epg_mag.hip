// ops_mcrmwi4_epg_mag.hip -- MCR-MWI, 4 flip angles, EPG-X MLP T1 path, magnitude data.
#include "gc_ops_impl.hpp"

namespace gc {
const ModelOps* ops_mcrmwi4_epg_mag() { return ops_table<McrmwiModel<4, true, false, true>>(); }
}  // namespace gc
