// Kernels of the SU(4), 18-real, float variant (see sun_kernels.inc).
#define SUN_R float
#define SUN_N 4
#define SUN_STORE 18
#define SUN_OPS_FN ops_n4_f18_r4
#define SUN_VARIANT_NS v_n4_f18_r4
#define SUN_UPDATE_TWO_WAVES 1
#include "sun_kernels.inc"
