// crm_hip_mixed.hip -- the mixed-precision engine: bulk fields (velocities, scalars,
// diagnostics, eddy coefficients, Adams-Bashforth history) and the advection, SGS and
// microphysics arithmetic in FP32; the pressure solve (right-hand side, transforms, Thomas
// sweep, gradient), the reference profiles, surface precipitation and every accumulated
// statistic and output in FP64.  In this unit the reference's REAL(4) literals (R4) are
// FP32 values and the bulk kernels write every literal as R(...), so the bulk arithmetic
// stays in FP32.
#define CRM_NS crm_mixed
#define CRM_MIXED_TU 1
#include "crm_driver.hpp"

namespace crm_mixed {
CrmEngine* make_engine_mixed(int device, int* rc) { return make_engine<float>(device, rc); }
}  // namespace crm_mixed
