// fv_euler.h — Euler equations (ideal gas) for the fused stage kernels.
//
// Arithmetic follows hydro/eqn/euler.cl with Cartesian normals resolved at
// compile time: the reference multiplies by normal components that are the
// constants 0 and 1 (coord/coord.lua:2376-2385); here those terms are dropped,
// which changes nothing but the sign of an exact zero. Evaluation order, every
// division and every square root are kept as the reference writes them.
#pragma once

#include "fv_common.h"

namespace hydro {

template <typename R>
struct Euler {
	static constexpr int Q = 5;       // numIntStates (euler.lua:14)
	static constexpr int NSTAT = 0;   // no static fields on the device
	static constexpr int NW = 5;      // numWaves (euler.lua:13)
	static constexpr int NSTATES = 6; // cons_t reals incl. ePot

	struct Par {
		R gamma, g1, rhoMin, PMin;
	};
	struct Prim {
		R rho, v[3], P;
	};
	struct Eig {
		R rho, v[3], hTotal, Cs, vSq;   // vL == v for Cartesian
	};

	__device__ static __forceinline__ Par par(const StageArgs<R> &a) {
		Par p;
		p.gamma = a.gamma;
		p.g1 = a.gamma - R(1);
		p.rhoMin = a.rhoMin;
		p.PMin = a.PMin;
		return p;
	}

	// primFromCons (euler.cl:140-159; the rhoMin branch is dead code there)
	__device__ static __forceinline__ Prim prim(const Par &p, const R *U) {
		Prim W;
		R inv = R(1) / U[0];
		W.rho = U[0];
		W.v[0] = U[1] * inv;
		W.v[1] = U[2] * inv;
		W.v[2] = U[3] * inv;
		R lenSq = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
		R EKin = U[0] < p.rhoMin ? R(0) : (R(.5) * lenSq / U[0]);
		W.P = U[0] < p.rhoMin ? R(0) : (p.g1 * (U[4] - EKin));
		return W;
	}

	__device__ static __forceinline__ R calc_Cs(const Par &p, const Prim &W) {
		if (W.P <= p.PMin) return R(0);
		if (W.rho < p.rhoMin) return R(INFINITY);
		return sqrt(p.gamma * W.P / W.rho);
	}

	// eigen_forInterface (euler.cl:347-430) from the two states and their prims
	__device__ static __forceinline__ Eig eigen(const Par &p, const R *UL, const R *UR,
		const Prim &WL, const Prim &WR)
	{
		const R rhoEpsilon = R(1e-5);
		Eig e;
		if (UL[0] < rhoEpsilon && UR[0] < rhoEpsilon) {
			e.rho = 0; e.v[0] = e.v[1] = e.v[2] = 0; e.hTotal = 0; e.Cs = 0; e.vSq = 0;
			return e;
		}
		if (UL[0] < rhoEpsilon || UR[0] < rhoEpsilon) {
			bool useR = UL[0] < rhoEpsilon;
			const R *Us = useR ? UR : UL;
			const Prim &W = useR ? WR : WL;
			e.rho = Us[0];
			e.v[0] = W.v[0]; e.v[1] = W.v[1]; e.v[2] = W.v[2];
			e.vSq = W.v[0] * W.v[0] + W.v[1] * W.v[1] + W.v[2] * W.v[2];
			e.hTotal = (W.P + Us[4]) / W.rho;
			e.Cs = calc_Cs(p, W);
			return e;
		}
		R sqrtRhoL = sqrt(WL.rho);
		R hTotalL = (WL.P + UL[4]) / WL.rho;
		R sqrtRhoR = sqrt(WR.rho);
		R hTotalR = (WR.P + UR[4]) / WR.rho;
		R invDenom = R(1) / (sqrtRhoL + sqrtRhoR);
		e.rho = sqrtRhoL * sqrtRhoR;
		R wl = sqrtRhoL * invDenom, wr = sqrtRhoR * invDenom;
		for (int i = 0; i < 3; ++i) e.v[i] = WL.v[i] * wl + WR.v[i] * wr;
		R hTotal = invDenom * (sqrtRhoL * hTotalL + sqrtRhoR * hTotalR);
		R vSq = e.v[0] * e.v[0] + e.v[1] * e.v[1] + e.v[2] * e.v[2];
		R eKin = R(.5) * vSq;
		R h = hTotal - eKin;
		if (h < rhoEpsilon) {
			e.hTotal = eKin;
			e.Cs = 0;
		} else {
			e.hTotal = hTotal;
			e.Cs = sqrt(p.g1 * h);
		}
		e.vSq = vSq;
		return e;
	}

	// eigen_leftTransform (euler.cl:434-489), normal along axis S
	template <int S>
	__device__ static __forceinline__ void left(const Par &p, const Eig &e, const R *X, R *Y) {
		if (e.rho < p.rhoMin) {
			for (int i = 0; i < 5; ++i) Y[i] = X[i];
			return;
		}
		constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
		const R Cs = e.Cs, vSq = e.vSq, g1 = p.g1;
		R denom = R(2) * Cs * Cs;
		R invDenom = R(1) / denom;
		R hg1vSq = R(.5) * g1 * vSq;
		R Csvn = Cs * e.v[A];
		R t[3];
		for (int i = 0; i < 3; ++i) t[i] = -g1 * e.v[i];
		R m[3] = {t[0], t[1], t[2]};
		m[A] = t[A] - Cs;   // -g1 vL_a - Cs n1_a
		R pl[3] = {t[0], t[1], t[2]};
		pl[A] = t[A] + Cs;  // -g1 vL_a + Cs n1_a
		Y[0] = (X[0] * (hg1vSq + Csvn) + X[1] * m[0] + X[2] * m[1] + X[3] * m[2] + X[4] * g1) * invDenom;
		Y[1] = (X[0] * (denom - g1 * vSq)
			+ X[1] * R(2) * g1 * e.v[0]
			+ X[2] * R(2) * g1 * e.v[1]
			+ X[3] * R(2) * g1 * e.v[2]
			+ X[4] * R(-2) * g1) * invDenom;
		Y[2] = X[0] * -e.v[B] + X[1 + B];
		Y[3] = X[0] * -e.v[C] + X[1 + C];
		Y[4] = (X[0] * (hg1vSq - Csvn) + X[1] * pl[0] + X[2] * pl[1] + X[3] * pl[2] + X[4] * g1) * invDenom;
	}

	// eigen_rightTransform (euler.cl:493-542)
	template <int S>
	__device__ static __forceinline__ void right(const Par &p, const Eig &e, const R *X, R *Y) {
		if (e.rho < p.rhoMin) {
			for (int i = 0; i < 5; ++i) Y[i] = X[i];
			return;
		}
		constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
		const R Cs = e.Cs;
		Y[0] = X[0] + X[1] + X[4];
		Y[1 + A] = X[0] * (e.v[A] - Cs) + X[1] * e.v[A] + X[4] * (e.v[A] + Cs);
		Y[1 + B] = X[0] * e.v[B] + X[1] * e.v[B] + X[2] + X[4] * e.v[B];
		Y[1 + C] = X[0] * e.v[C] + X[1] * e.v[C] + X[3] + X[4] * e.v[C];
		R Csvn = Cs * e.v[A];
		Y[4] = X[0] * (e.hTotal - Csvn) + X[1] * R(.5) * e.vSq + X[2] * e.v[B] + X[3] * e.v[C]
			+ X[4] * (e.hTotal + Csvn);
	}

	// fluxFromCons (euler.cl:273-291) given the state's prim
	template <int S>
	__device__ static __forceinline__ void flux(const R *U, const Prim &W, R *F) {
		R vn = W.v[S];
		F[0] = U[0] * vn;
		F[1] = U[1] * vn;
		F[2] = U[2] * vn;
		F[3] = U[3] * vn;
		F[1 + S] = F[1 + S] + W.P;
		F[4] = (U[4] + W.P) * vn;
	}

	// eigenWaveCode (euler.lua:309-325)
	template <int S>
	__device__ static __forceinline__ void waves(const Eig &e, R *l) {
		R vn = e.v[S];
		l[0] = vn - e.Cs;
		l[1] = vn;
		l[2] = vn;
		l[3] = vn;
		l[4] = vn + e.Cs;
	}

	// constrainU (euler.cl:697-717) on a state in type T
	template <typename T>
	__device__ static __forceinline__ void constrain(T gamma, T rhoMin, T PMin, T *U) {
		if (U[0] < rhoMin) U[0] = rhoMin;
		T g1 = gamma - T(1);
		T inv = T(1) / U[0];
		T v0 = U[1] * inv, v1 = U[2] * inv, v2 = U[3] * inv;
		T lenSq = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
		T EKin = U[0] < rhoMin ? T(0) : (T(.5) * lenSq / U[0]);
		T P = U[0] < rhoMin ? T(0) : (g1 * (U[4] - EKin));
		if (P < PMin) P = PMin;
		T rho = U[0];
		U[1] = v0 * rho;
		U[2] = v1 * rho;
		U[3] = v2 * rho;
		U[4] = rho * (T(.5) * (v0 * v0 + v1 * v1 + v2 * v2)) + P / g1;
	}

	// calcDTCell (calcDT.cl:38-73, euler.lua:346-373) on a state in type T
	template <typename T, int DIM>
	__device__ static __forceinline__ T dt_cell(T gamma, T rhoMin, T PMin, const T *dx, const T *U, T, T) {
		T lenSq = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
		T EKin = U[0] < rhoMin ? T(0) : (T(.5) * lenSq / U[0]);
		T P = U[0] < rhoMin ? T(0) : ((gamma - T(1)) * (U[4] - EKin));
		T Cs;
		if (P <= PMin) Cs = 0;
		else if (U[0] < rhoMin) Cs = T(INFINITY);
		else Cs = sqrt(gamma * P / U[0]);
		T dt = T(INFINITY);
		for (int s = 0; s < DIM; ++s) {
			if (dx[s] > T(0)) {
				T vn = U[0] < rhoMin ? T(0) : U[1 + s] / U[0];
				T lmin = vn - Cs, lmax = vn + Cs;
				T absLambdaMax = fmax(fabs(lmin), fabs(lmax));
				absLambdaMax = fmax(T(1e-9), absLambdaMax);
				dt = fmin(dt, dx[s] / absLambdaMax);
			}
		}
		return dt;
	}

	// mirror reflects the normal momentum component
	__device__ static __forceinline__ bool reflects(int field, int axis) { return field == 1 + axis; }
};

} // namespace hydro
