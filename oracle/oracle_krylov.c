/*
 * oracle_krylov.c -- instantiates oracle_krylov.inc for double and float.
 * TEST INFRASTRUCTURE ONLY (see sparkit_oracle.h).
 */
#include "sparkit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define T double
#define SFX d
#define OP orc_op_d
#define CFG orc_config_d
#define SUM orc_summary_d
#define SQRT sqrt
#include "oracle_krylov.inc"
#undef T
#undef SFX
#undef OP
#undef CFG
#undef SUM
#undef SQRT

#define T float
#define SFX f
#define OP orc_op_f
#define CFG orc_config_f
#define SUM orc_summary_f
#define SQRT sqrtf
#include "oracle_krylov.inc"
#undef T
#undef SFX
#undef OP
#undef CFG
#undef SUM
#undef SQRT
