// Kernels of the SU(2), 18-real, float variant (see sun_kernels.inc).
#define SUN_R float
#define SUN_N 2
#define SUN_STORE 18
#define SUN_OPS_FN ops_n2_f18_r4
#define SUN_VARIANT_NS v_n2_f18_r4
#include "sun_kernels.inc"
