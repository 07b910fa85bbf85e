// hydro_noalloc.cpp — hydro_conserved_totals_device: the conserved totals of
// a state the caller holds on the device, with no device allocation in the
// call (include/hydro_hip.h). The per-block partial sums that
// hydro_conserved_totals keeps in memory allocated with the solver live here
// in a file-scope device array instead, one copy per GPU, held until the
// process ends. Host code never passes the array's name to a kernel (in host
// code it is a stand-in, not a device address): the address comes from
// hipGetSymbolAddress, once per device, with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

// the solver helpers of fv_host.h that this file does not call
#pragma clang diagnostic ignored "-Wunused-function"
#include "fv_host.h"

using namespace hydro;

namespace {

// the same partition as the solver's totals (hydro_create: 1024 partial blocks)
constexpr int NOALLOC_PARTIALS = 1024;
constexpr size_t NOALLOC_SCRATCH_BYTES = sizeof(KSum) * NOALLOC_PARTIALS * HYDRO_MAX_STATES;   // 192 KiB
constexpr int NOALLOC_MAX_DEVICES = 64;

} // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex scratchMutex;
void *scratchPtr[NOALLOC_MAX_DEVICES];

// the device address of this device's scratch array (looked up on the first
// call on the device, under a lock)
int scratch_for_current_device(void **ptr) {
	int dev = 0;
	HIPCHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail("device %d beyond the %d the library tracks", dev,
		NOALLOC_MAX_DEVICES);
	std::lock_guard<std::mutex> lock(scratchMutex);
	if (!scratchPtr[dev]) {
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p) return fail("hipGetSymbolAddress of the scratch array: %s", hipGetErrorString(e));
		scratchPtr[dev] = p;
	}
	*ptr = scratchPtr[dev];
	return 0;
}

} // namespace

extern "C" int hydro_conserved_totals_device(const void *U, int precision, int num_fields, long long plane_stride,
	long long num_cells, double *out, void *stream)
{
	if (!U || !out) return fail("NULL argument");
	if (precision != HYDRO_REAL_F64 && precision != HYDRO_REAL_F32) return fail("precision %d unknown", precision);
	if (num_fields < 1 || num_fields > HYDRO_MAX_STATES)
		return fail("num_fields %d outside 1..%d: the scratch array holds the partial sums of at most %d fields "
			"(%zu bytes)", num_fields, HYDRO_MAX_STATES, HYDRO_MAX_STATES, NOALLOC_SCRATCH_BYTES);
	if (num_cells < 1) return fail("num_cells %lld must be positive", num_cells);
	if (plane_stride < num_cells) return fail("plane_stride %lld is below num_cells %lld", plane_stride, num_cells);
	void *p = nullptr;
	if (scratch_for_current_device(&p)) return -1;
	KSum *partial = static_cast<KSum *>(p);
	hipStream_t s = static_cast<hipStream_t>(stream);
	hipError_t e = precision == HYDRO_REAL_F64
		? launch_totals<double>(static_cast<const double *>(U), (long)plane_stride, (long)num_cells, num_fields,
			partial, NOALLOC_PARTIALS, out, s)
		: launch_totals<float>(static_cast<const float *>(U), (long)plane_stride, (long)num_cells, num_fields,
			partial, NOALLOC_PARTIALS, out, s);
	if (e != hipSuccess) return fail("totals kernel: %s", hipGetErrorString(e));
	return 0;
}
