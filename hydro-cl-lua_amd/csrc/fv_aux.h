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
	EqnParD pd;
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

// Whether any stored value of a state is not finite: the q integrated and the
// nstat static planes over the interior cells. Sets *flag to 1 and never
// clears it (the caller starts it at 0).
template <typename R>
__global__ void __launch_bounds__(256) nonfinite_kernel(const R *U, const R *stat, long ps, long ncell, int q,
	int nstat, unsigned *flag)
{
	bool bad = false;
	for (int f = 0; f < q + nstat; ++f) {
		const R *p = f < q ? U + (long)f * ps : stat + (long)(f - q) * ps;
		for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x)
			bad = bad || !__builtin_isfinite(p[c]);
	}
	if (wave_any(bad) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
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
		double d = E::template dt_cell<DIM>(g.pd, g.dx, u);
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
		E::constrain(g.pd.gamma, g.pd.rhoMin, g.pd.PMin, u);
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

// ------------------------------------------------------------ var stats
// calcDisplayVarRangeAndAvgAndStdDev (solverbase.lua:2871-2994) of one
// stored field: min and max over every cell of the ghosted grid (ghost values
// as the boundary fills them: periodic/freeflow copies, mirror flips, fixed
// values), the average and the population standard deviation over the
// interior (the reference zeroes the ghosts for its sums and divides by the
// interior count). Two passes, FP64 compensated sums.
// Derived Euler display variables (euler.lua:240-260) of one cell's cons
// (rho, m, ETotal) in double: W from primFromCons (euler.cl:140-159, P by
// calc_P, euler.cl:63-83), then the variable's formula from euler.cl:1-94.
struct StatsPar {
	double gamma, rhoMin, PMin;
};
__device__ inline double euler_display(const StatsPar &p, int id, const double *U) {
	const double rho = U[0];
	const double inv = 1. / rho;
	const double v[3] = {U[1] * inv, U[2] * inv, U[3] * inv};
	const double mSq = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
	const double EKinC = rho < p.rhoMin ? 0. : (.5 * mSq / rho);
	const double P = rho < p.rhoMin ? 0. : ((p.gamma - 1.) * (U[4] - EKinC));
	const double vSq = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
	auto Cs = [&]() { return P <= p.PMin ? 0. : rho < p.rhoMin ? (double)INFINITY : sqrt(p.gamma * P / rho); };
	switch (id) {
	case HYDRO_VAR_P: return P;
	case HYDRO_VAR_E_INT_SPEC: return P / (p.gamma - 1.) / rho;
	case HYDRO_VAR_E_KIN_SPEC: return .5 * vSq;
	case HYDRO_VAR_E_TOTAL_SPEC: return U[4] / rho;
	case HYDRO_VAR_E_INT: return P / (p.gamma - 1.);
	case HYDRO_VAR_E_KIN: return rho * (.5 * vSq);
	case HYDRO_VAR_S: return P / pow(rho, p.gamma);
	case HYDRO_VAR_H: return P * (p.gamma / (p.gamma - 1.));
	case HYDRO_VAR_H_SPEC: return P * (p.gamma / (p.gamma - 1.)) / rho;
	case HYDRO_VAR_H_TOTAL: return P + U[4];
	case HYDRO_VAR_H_TOTAL_SPEC: return (P + U[4]) / rho;
	case HYDRO_VAR_CS: return Cs();
	case HYDRO_VAR_MACH: return sqrt(vSq) / Cs();
	case HYDRO_VAR_V_X: return v[0];
	case HYDRO_VAR_V_Y: return v[1];
	default: return v[2];
	}
}

template <typename R, template <typename> class EQ>
__global__ void __launch_bounds__(TOTALS_BLOCK) stats_partial_kernel(const GridDesc<R> g, int dim, int field, int pass,
	const double *out, KSum *partial, StatsPar sp)
{
	using E = EQ<R>;
	__shared__ KSum sh[TOTALS_BLOCK];
	__shared__ KSum shm[TOTALS_BLOCK];
	int gs[3];
	for (int a = 0; a < 3; ++a) gs[a] = a < dim ? g.n[a] + 2 * NG : 1;
	const long total = (long)gs[0] * gs[1] * gs[2];
	const double avg = pass ? out[2] : 0.;
	double mn = INFINITY, mx = -INFINITY;
	KSum s = {0., 0.};
	for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < total; c += (long)gridDim.x * blockDim.x) {
		int cc[3];
		cc[0] = (int)(c % gs[0]) - NG;
		cc[1] = dim > 1 ? (int)((c / gs[0]) % gs[1]) - NG : 0;
		cc[2] = dim > 2 ? (int)(c / ((long)gs[0] * gs[1])) - NG : 0;
		bool interior = true;
		Mapped m[3];
		for (int a = 0; a < 3; ++a) {
			if (a < dim) {
				interior = interior && cc[a] >= 0 && cc[a] < g.n[a];
				m[a] = map_coord(cc[a], g.n[a], g.bc[2 * a], g.bc[2 * a + 1], 2 * a);
			} else {
				m[a].idx = 0; m[a].mirror = 0; m[a].fixedFace = -1;
			}
		}
		if (pass && !interior) continue;
		const long src = ((long)m[2].idx * g.n[1] + m[1].idx) * g.step1 + m[0].idx;
		// axis by axis as the boundary passes run: a fixed face writes its
		// value, a mirror face flips the reflected fields
		auto fieldAt = [&](int f) {
			R v = f < E::Q ? g.U[0][f * g.ps + src] : g.stat[(f - E::Q) * g.ps + src];
			for (int a = 0; a < dim; ++a) {
				if (m[a].fixedFace >= 0) v = g.fixedv[m[a].fixedFace][f];
				else if (m[a].mirror && E::reflects(f, a)) v = -v;
			}
			return (double)v;
		};
		double x;
		if constexpr (E::Q == 5) {
			if (field >= HYDRO_VAR_P) {
				double U[5];
				for (int f = 0; f < 5; ++f) U[f] = fieldAt(f);
				x = euler_display(sp, field, U);
			} else {
				x = fieldAt(field);
			}
		} else {
			x = fieldAt(field);
		}
		if (pass == 0) {
			mn = fmin(mn, x);
			mx = fmax(mx, x);
			if (interior) s = ksum_add(s, x);
		} else {
			const double d = x - avg;
			s = ksum_add(s, d * d);
		}
	}
	sh[threadIdx.x] = s;
	shm[threadIdx.x] = KSum{mn, mx};
	__syncthreads();
	for (int w = TOTALS_BLOCK / 2; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) {
			sh[threadIdx.x] = ksum_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
			shm[threadIdx.x].s = fmin(shm[threadIdx.x].s, shm[threadIdx.x + w].s);
			shm[threadIdx.x].c = fmax(shm[threadIdx.x].c, shm[threadIdx.x + w].c);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		partial[blockIdx.x] = sh[0];
		partial[gridDim.x + blockIdx.x] = shm[0];
	}
}

// out = {min, max, avg, stddev}: pass 0 writes the first three, pass 1 the last
[[maybe_unused]] static __global__ void __launch_bounds__(TOTALS_BLOCK) stats_final_kernel(const KSum *partial,
	int nPartial, int pass, double nInterior, double *out)
{
	__shared__ KSum sh[TOTALS_BLOCK];
	__shared__ KSum shm[TOTALS_BLOCK];
	KSum a = {0., 0.};
	KSum m = {INFINITY, -INFINITY};
	for (int i = threadIdx.x; i < nPartial; i += blockDim.x) {
		a = ksum_merge(a, partial[i]);
		m.s = fmin(m.s, partial[nPartial + i].s);
		m.c = fmax(m.c, partial[nPartial + i].c);
	}
	sh[threadIdx.x] = a;
	shm[threadIdx.x] = m;
	__syncthreads();
	for (int w = TOTALS_BLOCK / 2; w > 0; w >>= 1) {
		if ((int)threadIdx.x < w) {
			sh[threadIdx.x] = ksum_merge(sh[threadIdx.x], sh[threadIdx.x + w]);
			shm[threadIdx.x].s = fmin(shm[threadIdx.x].s, shm[threadIdx.x + w].s);
			shm[threadIdx.x].c = fmax(shm[threadIdx.x].c, shm[threadIdx.x + w].c);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const double sum = sh[0].s + sh[0].c;
		if (pass == 0) {
			out[0] = shm[0].s;
			out[1] = shm[0].c;
			out[2] = sum / nInterior;
		} else {
			out[3] = sqrt(sum / nInterior);
		}
	}
}

} // namespace hydro
