/* FP64 instantiation of the oracle integrator (TEST INFRASTRUCTURE). */
#include "oracle.h"
#define REAL double
#define REAL_IS_FLOAT 0
#define SFX(name) name##_f64
#define R_SIN sin
#define R_COS cos
#define R_ACOS acos
#define R_ATAN2 atan2
#define R_EXP exp
#define R_LOG log
#define R_POW pow
#define R_SQRT sqrt
#define R_FLOOR floor
#define R_FABS fabs
#define R_FMAX fmax
#define R_FMIN fmin
#include "oracle_core.inc"
