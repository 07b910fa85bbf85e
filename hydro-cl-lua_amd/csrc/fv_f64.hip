// FP64 dispatch of the stage kernels by policy and dimension, and the FP64
// auxiliary kernels (dt, constrainU, conserved totals, statistics, dt slots).
// Built with -ffp-contract=off so the arithmetic is the reference's expression
// order, operation for operation.
#define REAL double
#define AUX_KERNELS
#include "fv_kernels.inc"

namespace hydro {
hipError_t launch_reset_slots(DevScalars *dev, int mask, hipStream_t s) {
	hipLaunchKernelGGL(reset_slots_kernel, dim3(1), dim3(64), 0, s, dev, mask);
	return hipGetLastError();
}
hipError_t launch_fold_slot(DevScalars *dev, int slot, hipStream_t s) {
	hipLaunchKernelGGL(fold_slot_kernel, dim3(1), dim3(64), 0, s, dev, slot);
	return hipGetLastError();
}
} // namespace hydro
