// Kernels of the SU(2), 18-real, double variant (see sun_kernels.inc).
#define SUN_R double
#define SUN_N 2
#define SUN_STORE 18
#define SUN_OPS_FN ops_n2_f18_r8
#define SUN_VARIANT_NS v_n2_f18_r8
#include "sun_kernels.inc"
