// gmc_fused.hip -- fused stage kernels used by gmc_step.
#include "gmc_internal.hpp"
#include "gmc_point.hpp"

namespace gmc {

int fused_integrate(gmc_ctx *c, int o, int n, double dt) { return ops_integrate(c, o, n, dt); }

}  // namespace gmc
