// fv_sw.h — shallow water equations (hydro/eqn/shallow-water.lua / .cl) for
// the fused stage kernels, in the reference's depthInCell = false form:
// cons_t = h (0), m.xyz (1-3), integrated, and the depth (4), a static field
// kept in its own plane (the reference's other form keeps it in cell_t; the
// arithmetic is the same). extraTermInFlux = false, as the reference sets it.
// The transforms keep the reference's expressions term for term, including
// the products with the normal components (constants 0 and 1 here), so the
// EXACT policy's arithmetic is the reference's; the FAST policy's transforms,
// flux and source drop the products with a zero component.
#pragma once

#include "fv_common.h"
#include "fv_math.h"

namespace hydro {

template <bool FAST>
struct ShallowWaterT {
template <typename R>
struct Eqn {
	using M = Math<R, FAST>;
	static constexpr bool FASTM = FAST;
	static constexpr int FLUXID = HYDRO_FLUX_ROE;
	static constexpr bool PLM_PRIM = false;
	static constexpr bool PLM_EIG = false;
	static constexpr bool PLM_CONS_FLUX = false;
	static constexpr bool PLM_ATHENA = false;
	static constexpr bool PLM_PPM = false;
	static constexpr bool PLM_EIG_PRIM = false;
	static constexpr bool PLM_EIG_PRIM_REF = false;
	static constexpr int Q = 4;        // numIntStates (shallow-water.lua:56)
	static constexpr int NSTAT = 1;    // depth
	static constexpr int NW = 4;       // numWaves (shallow-water.lua:55)
	static constexpr int NSTATES = 5;
	static constexpr bool SOURCE = true;
	static constexpr int SRC_FIELD = 4;   // the source reads the neighbours' depth
	__device__ static constexpr bool active(int f) { return f < 5; }

	struct Par {
		R g;
	};
	struct Prim {
		R h, v[3];
	};
	struct Eig {
		R h, v[3], C;
	};

	__device__ static __forceinline__ Par par(const StageArgs<R> &a) {
		Par p;
		p.g = a.gravity;
		return p;
	}

	// primFromCons (shallow-water.cl:3-15): v = m * (1 / h)
	__device__ static __forceinline__ Prim prim(const Par &, const R *U) {
		Prim W;
		W.h = U[0];
		const R inv = M::rcp(U[0]);
#pragma unroll
		for (int i = 0; i < 3; ++i) W.v[i] = U[1 + i] * inv;
		return W;
	}

	// eigen_forInterface (shallow-water.cl:228-265): the arithmetic mean of the
	// heights, the velocities weighted by sqrt(h)
	__device__ static __forceinline__ Eig eigen(const Par &p, const R *UL, const R *UR, const Prim &WL,
		const Prim &WR)
	{
		Eig e;
		const R sqrtHL = M::sqrt_(WL.h), sqrtHR = M::sqrt_(WR.h);
		const R invDenom = M::rcp(sqrtHL + sqrtHR);
		e.h = R(.5) * (UL[0] + UR[0]);
#pragma unroll
		for (int i = 0; i < 3; ++i) e.v[i] = WL.v[i] * (sqrtHL * invDenom) + WR.v[i] * (sqrtHR * invDenom);
		const R CSq = p.g * e.h;
		e.C = M::sqrt_(CSq);
		return e;
	}

	// normal_l<j>.<i> / normal_u<j>.<i> of axis S (coord.lua:2376-2385)
	template <int S, int J, int I>
	__device__ static constexpr R nl() { return (((I - J) % 3 + 3) % 3) == S ? R(1) : R(0); }

	// The tangential momentum fields of axis S: T1 the one with
	// normal_l<2> = 1, T2 the one with normal_l<3> = 1 (1-based cons indices)
	template <int S> static constexpr int T1 = (S + 1) % 3 + 1;
	template <int S> static constexpr int T2 = (S + 2) % 3 + 1;

	// eigen_leftTransform (shallow-water.cl:286-318), nLen = 1. The fast
	// policy drops the terms whose normal component is 0 (products with an
	// exact zero), keeping the order of the others
	template <int S>
	__device__ static __forceinline__ void left(const Par &, const Eig &e, const R *X, R *Y) {
		if constexpr (FAST) {
			const R v_n = e.v[S];
			const R invDenom = M::rcp(R(2) * e.h * e.C);
			const R tmp2 = e.C * X[0];
			const R tmp5 = -(v_n * X[0]);
			const R tmp31 = invDenom * e.C;
			Y[0] = invDenom * (X[1 + S] + tmp5 + -tmp2);
			Y[1] = R(2) * tmp31 * (X[T1<S>] + -(e.v[T1<S> - 1] * X[0]));
			Y[2] = R(2) * tmp31 * (X[T2<S>] + -(e.v[T2<S> - 1] * X[0]));
			Y[3] = invDenom * (X[1 + S] + tmp2 + tmp5);
			return;
		}
		const R nLen = R(1);
		const R v_n = e.v[S];
		const R invDenom = M::rcp(R(2) * nLen * e.h * e.C);
		const R tmp1 = e.C * X[0];
		const R tmp2 = nLen * tmp1;
		const R tmp3 = v_n * X[0];
		const R tmp5 = -tmp3;
		const R tmp6 = nl<S, 1, 3>() * X[3];
		const R tmp8 = nl<S, 1, 2>() * X[2];
		const R tmp10 = nl<S, 1, 1>() * X[1];
		const R tmp13 = e.v[0] * X[0];
		const R tmp15 = e.v[1] * X[0];
		const R tmp17 = e.v[2] * X[0];
		const R tmp29 = nLen * e.C;
		const R tmp31 = invDenom * tmp29;
		Y[0] = invDenom * (tmp10 + tmp8 + tmp6 + tmp5 + -tmp2);
		Y[1] = R(2) * tmp31 * (nl<S, 2, 1>() * X[1] + nl<S, 2, 2>() * X[2] + nl<S, 2, 3>() * X[3]
			+ -nl<S, 2, 1>() * tmp13 + -nl<S, 2, 3>() * tmp17 + -nl<S, 2, 2>() * tmp15);
		Y[2] = R(2) * tmp31 * (nl<S, 3, 1>() * X[1] + nl<S, 3, 2>() * X[2] + nl<S, 3, 3>() * X[3]
			+ -nl<S, 3, 1>() * tmp13 + -nl<S, 3, 3>() * tmp17 + -nl<S, 3, 2>() * tmp15);
		Y[3] = invDenom * (tmp10 + tmp8 + tmp6 + tmp2 + tmp5);
	}

	// eigen_rightTransform (shallow-water.cl:322-362), nLen = n2Len = n3Len = 1
	template <int S>
	__device__ static __forceinline__ void right(const Par &, const Eig &e, const R *X, R *Y) {
		if constexpr (FAST) {
			// the same sums without their zero normal terms; the lengths are 1
			const R h = e.h, C = e.C;
			const R invDenom = M::rcp(C);
			const R tmp10 = C * C * X[0];
			const R tmp16 = X[3] * (C * C);
			const R tmp22 = C * X[0];
			const R tmp37 = C * X[3];
			const R vn = e.v[S], v1 = e.v[T1<S> - 1], v2 = e.v[T2<S> - 1];
			Y[0] = invDenom * (X[3] + -X[0]) * h * C;
			Y[1 + S] = invDenom * (h * (tmp10 + tmp16 + -vn * tmp22 + vn * tmp37));
			Y[T1<S>] = invDenom * (h * (-v1 * tmp22 + C * X[1] + v1 * tmp37));
			Y[T2<S>] = invDenom * (h * (-v2 * tmp22 + v2 * tmp37 + C * X[2]));
			return;
		}
		const R nLen = R(1), n2Len = R(1), n3Len = R(1);
		const R h = e.h, C = e.C;
		const R invDenom = M::rcp(C);
		const R tmp5 = n3Len * n3Len;
		const R tmp6 = n2Len * n2Len;
		const R tmp7 = tmp5 * X[0];
		const R tmp8 = C * C;
		const R tmp9 = tmp6 * tmp7;
		const R tmp10 = tmp8 * tmp9;
		const R tmp13 = X[3] * tmp5;
		const R tmp15 = tmp13 * tmp6;
		const R tmp16 = tmp15 * tmp8;
		const R tmp17 = C * X[0];
		const R tmp19 = nLen * tmp17;
		const R tmp21 = tmp19 * tmp5;
		const R tmp22 = tmp21 * tmp6;
		const R tmp24 = C * X[1];
		const R tmp26 = nLen * tmp24;
		const R tmp27 = tmp26 * tmp5;
		const R tmp28 = C * X[2];
		const R tmp30 = nLen * tmp28;
		const R tmp31 = tmp30 * tmp6;
		const R tmp32 = C * X[3];
		const R tmp34 = nLen * tmp32;
		const R tmp36 = tmp34 * tmp5;
		const R tmp37 = tmp36 * tmp6;
		const R tmp51 = nLen * tmp5;
		const R tmp53 = tmp51 * tmp6;
		const R inv53 = M::rcp(tmp53);   // tmp53 = 1: the division is exact either way
		Y[0] = invDenom * (X[3] + -X[0]) * h * C;
		Y[1] = invDenom * (h * (nl<S, 1, 1>() * tmp10 + nl<S, 1, 1>() * tmp16 + -e.v[0] * tmp22
			+ nl<S, 2, 1>() * tmp27 + e.v[0] * tmp37 + nl<S, 3, 1>() * tmp31)) * inv53;
		Y[2] = invDenom * (h * (nl<S, 1, 2>() * tmp10 + nl<S, 1, 2>() * tmp16 + -e.v[1] * tmp22
			+ nl<S, 2, 2>() * tmp27 + e.v[1] * tmp37 + nl<S, 3, 2>() * tmp31)) * inv53;
		Y[3] = invDenom * (h * (nl<S, 1, 3>() * tmp10 + nl<S, 1, 3>() * tmp16 + -e.v[2] * tmp22
			+ nl<S, 2, 3>() * tmp27 + e.v[2] * tmp37 + nl<S, 3, 3>() * tmp31)) * inv53;
	}

	// fluxFromCons (shallow-water.cl:186-206), extraTermInFlux = false. The
	// fast policy takes v_n from the cell's prims (m_n * (1 / h), one rounding
	// from m_n / h) instead of a second division (2 fewer divisions per
	// interface) and leaves out the zero normal terms
	template <int S>
	__device__ static __forceinline__ void flux(const Par &p, const R *U, const Prim &W, R *F) {
		const R m_n = U[1 + S];
		const R v_n = FAST ? W.v[S] : M::div(m_n, U[0]);
		const R hP = R(.5) * p.g * U[0] * (U[0]);
		if constexpr (FAST) {
			F[0] = m_n;
#pragma unroll
			for (int i = 1; i <= 3; ++i) F[i] = i == 1 + S ? U[i] * v_n + hP : U[i] * v_n;
			return;
		}
		F[0] = m_n;
		F[1] = U[1] * v_n + nl<S, 1, 1>() * hP;
		F[2] = U[2] * v_n + nl<S, 1, 2>() * hP;
		F[3] = U[3] * v_n + nl<S, 1, 3>() * hP;
	}

	// eigenWaveCodePrefix / eigenWaveCode (shallow-water.lua:247-275)
	template <int S>
	__device__ static __forceinline__ void waves(const Par &p, const Eig &e, R *l) {
		const R CSq = p.g * e.h;
		const R C = M::sqrt_(CSq);
		R C_nLen = C * R(1);
		R v_n = e.v[S];
		if (C <= R(0)) {
			C_nLen = R(0);
			v_n = R(0);
		}
		l[0] = v_n - C_nLen;
		l[1] = v_n;
		l[2] = v_n;
		l[3] = v_n + C_nLen;
	}

	__device__ static __forceinline__ void constrain(double, double, double, double *) {}

	// addSource (shallow-water.cl:366-428, its enabled block) for one interior
	// cell: em / ep the depth of the -1 / +1 neighbours per axis; the
	// second-order central difference of the depth (makepartial.lua:13-20,
	// 85-110: coefficient 1, denominator 2 dx), divided by the basis lengths
	// (1) and raised (identity), times g h, added to the momentum derivative
	template <int DIM>
	__device__ static __forceinline__ void source(const Par &p, const GridDesc<R> &g, const R *U, const R *em,
		const R *ep, R *d)
	{
		const R gh = p.g * U[0];
#pragma unroll
		for (int j = 0; j < 3; ++j) {
			if (j < DIM) {
				const R partial = (R(0) + (ep[j] - em[j]) * R(1)) * g.inv2dx[j];
				d[1 + j] = d[1 + j] + partial * gh;
			} else if constexpr (!FAST) {
				d[1 + j] = d[1 + j] + R(0) * gh;
			}
		}
	}

	// calcDTCell (calcDT.cl:38-73) with consWaveCodeMinMaxAllSides
	// (shallow-water.lua:318-354), in double
	template <int DIM>
	__device__ static __forceinline__ double dt_cell(const EqnParD &pd, const double *dx, const double *U) {
		using MD = Math<double, FAST>;
		const double CSq = pd.gravity * U[0];
		const double C = MD::sqrt_(CSq);
		const double inv = MD::rcp(U[0]);
		double dt = INFINITY;
#pragma unroll
		for (int s = 0; s < DIM; ++s) {
			if (dx[s] > 0.) {
				double C_nLen = C * 1.;
				double v_n = U[1 + s] * inv;
				if (C <= 0.) {
					C_nLen = 0.;
					v_n = 0.;
				}
				const double lmin = v_n - C_nLen, lmax = v_n + C_nLen;
				double absLambdaMax = fmax(fabs(lmin), fabs(lmax));
				absLambdaMax = fmax(1e-9, absLambdaMax);
				dt = fmin(dt, MD::div(dx[s], absLambdaMax));
			}
		}
		return dt;
	}

	// mirror reflects the momentum along the axis (reflectVars.mirror, eqn.lua:366-370)
	__device__ static __forceinline__ bool reflects(int field, int axis) { return field == 1 + axis; }
};
};

} // namespace hydro
