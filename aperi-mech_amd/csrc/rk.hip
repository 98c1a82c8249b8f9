// RKPM set-up and smoothed-cell force (placeholder until the RK kernels land).
#include "internal.hpp"

namespace aphip {
void rk_setup(Ctx &, double) { throw Error(APHIP_ERR_UNSUPPORTED, "RKPM not built yet"); }
void rk_cell_force(Ctx &) { throw Error(APHIP_ERR_UNSUPPORTED, "RKPM not built yet"); }
void rk_node_update(Ctx &, double, int) { throw Error(APHIP_ERR_UNSUPPORTED, "RKPM not built yet"); }
void rk_mass(Ctx &, double *) { throw Error(APHIP_ERR_UNSUPPORTED, "RKPM not built yet"); }
}  // namespace aphip
