// FP64 stage kernels of the FAST policy. Built with -ffp-contract=fast: a
// product feeding a sum becomes one fused multiply-add, about a fifth fewer
// VALU instructions per cell than the unfused expression order of the EXACT
// kernels, at results within the FAST policy's tolerance.
#define REAL double
#define STAGE_FAST
#include "fv_kernels.inc"
