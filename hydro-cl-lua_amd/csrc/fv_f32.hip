// FP32 kernels: stored state and stencil arithmetic in float; the stage
// accumulation, constrainU and dt work in double (see fv_stage.h).
#define REAL float
#define STAGE_EXACT
#define STAGE_FAST
#define AUX_KERNELS
#include "fv_kernels.inc"
