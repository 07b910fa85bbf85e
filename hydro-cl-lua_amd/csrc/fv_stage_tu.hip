// The stage kernels of one stored precision, math policy and dimension. The
// build compiles this file once per combination, with REAL (double / float),
// STAGE_EXACT or STAGE_FAST and STAGE_DIM (1, 2, 3) on the command line: the
// EXACT FP64 kernels without FP contraction (the reference's expression order,
// operation for operation), the FAST FP64 and all FP32 kernels with it (a
// product feeding a sum becomes one fused multiply-add).
#include "fv_kernels.inc"
