// Kernels of the SU(3), 8-real, float variant (see sun_kernels.inc).
#define SUN_R float
#define SUN_N 3
#define SUN_STORE 8
#define SUN_OPS_FN ops_n3_f8_r4
#define SUN_VARIANT_NS v_n3_f8_r4
#include "sun_kernels.inc"
