// fv_launch.h — host-side launchers of the device kernels, one set per stored
// precision (defined in fv_f64.hip, fv_f32.hip and the fv_stage_tu.hip builds).
#pragma once

#include <hip/hip_runtime.h>

#include "fv_common.h"

namespace hydro {

struct KSum;

// Selects one instantiation of the stage kernel.
struct StageKey {
	int eqn;        // HYDRO_EQN_*
	int dim;        // 1, 2
	int plm;        // plm cons on
	int lim;        // Roe flux limiter on
	int limiter;    // limiter index (flux limiter if lim, slope limiter if plm)
	int fast;       // FAST math policy
	int nvel;       // Euler: active velocity components (dim or 3)
	int flux;       // HYDRO_FLUX_*
	int roeFix;     // Roe: Harten's entropy fix
	int plmKind;    // HYDRO_PLM_* of the plm kernels
};

// Limiters with a compile-time form; the rest use the runtime switch.
inline int limiter_template_id(int index) {
	switch (index) {
	case HYDRO_LIMITER_MINMOD:
	case HYDRO_LIMITER_SUPERBEE:
	case HYDRO_LIMITER_VAN_LEER:
		return index;
	}
	return -1;
}

template <typename R> struct GridArgs;

// grid.x > 0: plain launch of that many blocks; grid.x == 0: persistent
// batched launch over a.grids; grid.x == 0 and block.x == 0: no launch, the
// plain kernel's resident blocks per CU go to stage_occupancy (this thread's).
template <typename R>
hipError_t launch_stage(const StageKey &k, const StageArgs<R> &a, dim3 grid, dim3 block, hipStream_t s);
extern thread_local int stage_occupancy;
// the same for one math policy and dimension (launch_stage picks by k.fast and
// k.dim); each is defined in a translation unit of its own
template <typename R, bool FAST, int DIM>
hipError_t launch_stage_dim(const StageKey &k, const StageArgs<R> &a, dim3 grid, dim3 block, hipStream_t s);

template <typename R>
hipError_t launch_dt(int eqn, int dim, int fast, const GridArgs<R> &g, hipStream_t s);

template <typename R>
hipError_t launch_constrain(int eqn, int fast, const GridArgs<R> &g, hipStream_t s);

// *flag = 1 when any value of the q + nstat planes (interior cells) is not finite
template <typename R>
hipError_t launch_nonfinite(const R *U, const R *stat, long ps, long ncell, int q, int nstat, unsigned *flag,
	hipStream_t s);

template <typename R>
hipError_t launch_totals(const R *U, long ps, long ncell, int q, KSum *partial, int nPartial, double *out, hipStream_t s);

// min, max, average, standard deviation of one stored field (out: 4 doubles
// on the device; partial: 2 nPartial KSums)
template <typename R>
hipError_t launch_stats(int eqn, int dim, const GridDesc<R> &g, int field, KSum *partial, int nPartial, double nInterior,
	double *out, const double *eqnPar /* gamma, rhoMin, PMin */, hipStream_t s);

hipError_t launch_reset_slots(DevScalars *dev, int mask, hipStream_t s);
hipError_t launch_fold_slot(DevScalars *dev, int slot, hipStream_t s);

} // namespace hydro
