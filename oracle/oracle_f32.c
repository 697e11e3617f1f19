/* FP32 instantiation of the oracle integrator (TEST INFRASTRUCTURE). */
#include "oracle.h"
#include "omath.h"
#define REAL float
#define REAL_IS_FLOAT 1
#define SFX(name) name##_f32
#define R_SIN om_sinf
#define R_COS om_cosf
#define R_ACOS om_acosf
#define R_ATAN2 om_atan2f
#define R_EXP om_expf
#define R_LOG om_logf
#define R_POW om_powf
#define R_SQRT sqrtf
#define R_FLOOR floorf
#define R_FABS fabsf
#include "oracle_core.inc"

/* exported portable math for the accuracy tests */
float oracle_sinf(float x) { return om_sinf(x); }
float oracle_cosf(float x) { return om_cosf(x); }
float oracle_atan2f(float y, float x) { return om_atan2f(y, x); }
float oracle_acosf(float x) { return om_acosf(x); }
float oracle_expf(float x) { return om_expf(x); }
float oracle_logf(float x) { return om_logf(x); }
float oracle_powf(float x, float y) { return om_powf(x, y); }
