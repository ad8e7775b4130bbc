// Kernels of the SU(3), 18-real, double variant (see sun_kernels.inc).
#define SUN_R double
#define SUN_N 3
#define SUN_STORE 18
#define SUN_OPS_FN ops_n3_f18_r8
#define SUN_VARIANT_NS v_n3_f18_r8
#include "sun_kernels.inc"
