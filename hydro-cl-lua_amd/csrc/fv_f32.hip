// FP32 kernels: stored state and stencil arithmetic in float; the stage
// accumulation, constrainU and dt work in double (see fv_stage.h).
#define REAL float
#include "fv_kernels.inc"
