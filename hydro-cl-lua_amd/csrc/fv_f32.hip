// FP32 dispatch of the stage kernels by policy and dimension, and the FP32
// auxiliary kernels: stored state and stencil arithmetic in float; the stage
// accumulation, constrainU and dt work in double (see fv_stage.h).
#define REAL float
#define AUX_KERNELS
#include "fv_kernels.inc"
