// fv_aux.h — the small kernels around the fused stage: standalone calcDT with
// its min-reduction, constrainU, dt-slot reset, and deterministic compensated
// conserved-total sums.
#pragma once

#include "fv_common.h"
#include "fv_euler.h"
#include "fv_glm.h"

namespace hydro {

// Shape of the interior state for the elementwise kernels.
template <typename R>
struct GridArgs {
	R *U;            // q integrated planes
	const R *stat;   // static planes
	long ps;         // plane stride
	long ncell;      // interior cells (planes are dense)
	int dim;
	double dx[3];
	double gamma, rhoMin, PMin, divPhi, divPsi;
	DevScalars *dev;
	int slot;
};

[[maybe_unused]] static __global__ void reset_slots_kernel(DevScalars *dev, int mask) {
	if (blockIdx.x == 0 && threadIdx.x < DT_LANES) {
		for (int i = 0; i < 3; ++i)
			if (mask & (1 << i)) dev->dtmin[i][threadIdx.x] = INFINITY;
	}
}

// min of the 64 sub-slots into sub-slot 0 (for host readback)
[[maybe_unused]] static __global__ void fold_slot_kernel(DevScalars *dev, int slot) {
	double m = read_dt_slot(dev, slot);
	if (threadIdx.x == 0) dev->dtmin[slot][0] = m;
}

// calcDT (eqn.lua:1187-1224 + calcDT.cl:38-73) and reduceMin: per-cell dt in
// double, wave min, one atomic min per wave into dev->dtmin[slot].
template <typename R, template <typename> class EQ, int DIM>
__global__ void __launch_bounds__(256) dt_kernel(const GridArgs<R> g) {
	using E = EQ<R>;
	constexpr int Q = E::Q;
	double m = INFINITY;
	for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long)gridDim.x * blockDim.x) {
		double u[Q + E::NSTAT];
#pragma unroll
		for (int j = 0; j < Q; ++j) u[j] = (double)g.U[j * g.ps + c];
#pragma unroll
		for (int j = 0; j < E::NSTAT; ++j) u[Q + j] = (double)g.stat[j * g.ps + c];
		double d = E::template dt_cell<DIM>(g.gamma, g.rhoMin, g.PMin, g.dx, u, g.divPhi, g.divPsi);
		m = fmin(m, d);
	}
	m = wave_min(m);
	int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	if ((threadIdx.x & 63) == 0 && m < INFINITY) atomic_min_pos(&g.dev->dtmin[g.slot][w & (DT_LANES - 1)], m);
}

// constrainU (euler.cl:697-717) over the interior; ghosts follow by mapping.
template <typename R, template <typename> class EQ>
__global__ void __launch_bounds__(256) constrain_kernel(const GridArgs<R> g) {
	using E = EQ<R>;
	constexpr int Q = E::Q;
	for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.ncell; c += (long)gridDim.x * blockDim.x) {
		double u[Q];
#pragma unroll
		for (int j = 0; j < Q; ++j) u[j] = (double)g.U[j * g.ps + c];
		E::constrain(g.gamma, g.rhoMin, g.PMin, u);
#pragma unroll
		for (int j = 0; j < Q; ++j) g.U[j * g.ps + c] = R(u[j]);
	}
}

// ------------------------------------------------------ compensated totals
// Neumaier summation per thread over a fixed cell subset, then a fixed-order
// tree over the block and a second single-block pass over block partials.
struct KSum {
	double s, c;
};
__device__ __forceinline__ KSum ksum_add(KSum a, double x) {
	double t = a.s + x;
	if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
	else a.c += (x - t) + a.s;
	a.s = t;
	return a;
}
__device__ __forceinline__ KSum ksum_merge(KSum a, KSum b) {
	a = ksum_add(a, b.s);
	a.c += b.c;
	return a;
}

constexpr int TOTALS_BLOCK = 256;

template <typename R>
__global__ void __launch_bounds__(TOTALS_BLOCK) totals_partial_kernel(const R *U, long ps, long ncell, int q, KSum *partial) {
	__shared__ KSum sh[TOTALS_BLOCK];
	for (int j = 0; j < q; ++j) {
		KSum a = {0., 0.};
		for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x)
			a = ksum_add(a, (double)U[j * ps + c]);
		sh[threadIdx.x] = a;
		__syncthreads();
		for (int w = TOTALS_BLOCK / 2; w > 0; w >>= 1) {
			if ((int)threadIdx.x < w) sh[threadIdx.x] = ksum_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
			__syncthreads();
		}
		if (threadIdx.x == 0) partial[j * gridDim.x + blockIdx.x] = sh[0];
		__syncthreads();
	}
}

[[maybe_unused]] static __global__ void __launch_bounds__(TOTALS_BLOCK) totals_final_kernel(const KSum *partial, int nPartial, int q, double *out) {
	__shared__ KSum sh[TOTALS_BLOCK];
	for (int j = 0; j < q; ++j) {
		KSum a = {0., 0.};
		for (int i = threadIdx.x; i < nPartial; i += blockDim.x) a = ksum_merge(a, partial[j * nPartial + i]);
		sh[threadIdx.x] = a;
		__syncthreads();
		for (int w = TOTALS_BLOCK / 2; w > 0; w >>= 1) {
			if ((int)threadIdx.x < w) sh[threadIdx.x] = ksum_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
			__syncthreads();
		}
		if (threadIdx.x == 0) out[j] = sh[0].s + sh[0].c;
		__syncthreads();
	}
}

} // namespace hydro
