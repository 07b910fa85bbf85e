// fv_euler.h — Euler equations (ideal gas) for the fused stage kernels.
//
// Arithmetic follows hydro/eqn/euler.cl with Cartesian normals resolved at
// compile time: the reference multiplies by normal components that are the
// constants 0 and 1 (coord/coord.lua:2376-2385); here those terms are dropped,
// which changes nothing but the sign of an exact zero. Evaluation order is kept
// as the reference writes it; divisions and square roots go through the Math
// policy (IEEE, or the faithful FAST sequences).
//
// NVEL < 3 means velocity components NVEL..2 are identically zero (2D / 1D
// runs whose initial state has no out-of-plane momentum). Every term such a
// component enters is an exact zero in the reference too (its wave amplitude
// L dU and its flux are +-0), so those terms are dropped and its plane is not
// streamed.
#pragma once

#include "fv_common.h"
#include "fv_math.h"

namespace hydro {

// FLUX: the interface flux the stage kernels build from this equation
// (HYDRO_FLUX_ROE, _HLL, _RUSANOV, _HLLC, _MARQUINA, or FLUX_ROE_ENTROPY_FIX).
// PLMK: the reconstruction of the PLM kernels, HYDRO_PLM_CONS, HYDRO_PLM_PRIM
// (plm.cl:191-254), HYDRO_PLM_EIG (plm.cl:256-432), HYDRO_PLM_CONS_FLUX
// (plm.cl:95-188), HYDRO_PLM_ATHENA (plm.cl:782-879), HYDRO_PLM_PPM
// (plm.cl:882-944) or HYDRO_PLM_EIG_PRIM(_REF) (plm.cl:536-780).
template <bool FAST, int NVEL, int FLUX = HYDRO_FLUX_ROE, int PLMK = HYDRO_PLM_CONS>
struct EulerT {
	template <typename R>
	struct Eqn {
		using M = Math<R, FAST>;
		static constexpr bool FASTM = FAST;
		static constexpr int FLUXID = FLUX;
		static constexpr bool PLM_PRIM = PLMK == HYDRO_PLM_PRIM;
		static constexpr bool PLM_EIG = PLMK == HYDRO_PLM_EIG;
		static constexpr bool PLM_CONS_FLUX = PLMK == HYDRO_PLM_CONS_FLUX;
		static constexpr bool PLM_ATHENA = PLMK == HYDRO_PLM_ATHENA;
		static constexpr bool PLM_PPM = PLMK == HYDRO_PLM_PPM;
		static constexpr bool PLM_EIG_PRIM = PLMK == HYDRO_PLM_EIG_PRIM;
		static constexpr bool PLM_EIG_PRIM_REF = PLMK == HYDRO_PLM_EIG_PRIM_REF;
		static constexpr int Q = 5;       // numIntStates (euler.lua:14)
		static constexpr int NSTAT = 0;   // no static fields on the device
		static constexpr int NW = 5;      // numWaves (euler.lua:13)
		static constexpr int NSTATES = 6; // cons_t reals incl. ePot
		static constexpr bool SOURCE = false;

		// fields the kernel streams: rho, the active momenta, ETotal
		__device__ static constexpr bool active(int f) { return f == 0 || f == 4 || (f >= 1 && f <= 3 && f - 1 < NVEL); }

		struct Par {
			R gamma, g1, rhoMin, PMin;
			R fixFactor;   // Roe: Harten's entropy fix factor (FLUX_ROE_ENTROPY_FIX kernels)
		};
		struct Prim {
			R rho, v[3], P, inv;   // inv = 1 / rho
		};
		struct Eig {
			R rho, v[3], hTotal, Cs, vSq;   // vL == v for Cartesian
		};

		__device__ static __forceinline__ Par par(const StageArgs<R> &a) {
			Par p;
			p.gamma = a.gamma;
			p.g1 = a.gamma - R(1);
			p.rhoMin = a.rhoMin;
			p.fixFactor = a.roeFixFactor;
			p.PMin = a.PMin;
			return p;
		}

		template <typename T>
		__device__ static __forceinline__ T lenSq3(const T *m) {
			T s = m[0] * m[0];
			if constexpr (NVEL > 1) s = s + m[1] * m[1];
			if constexpr (NVEL > 2) s = s + m[2] * m[2];
			return s;
		}
		template <typename T>
		__device__ static __forceinline__ T dot3(const T *a, const T *b) {
			T s = a[0] * b[0];
			if constexpr (NVEL > 1) s = s + a[1] * b[1];
			if constexpr (NVEL > 2) s = s + a[2] * b[2];
			return s;
		}

		// primFromCons (euler.cl:140-159; the rhoMin branch is dead code there)
		__device__ static __forceinline__ Prim prim(const Par &p, const R *U) {
			Prim W;
			R inv = M::rcp(U[0]);
			W.inv = inv;
			W.rho = U[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) W.v[i] = i < NVEL ? U[1 + i] * inv : R(0);
			R lenSq = lenSq3(U + 1);
			R EKin = M::div_inv(R(.5) * lenSq, U[0], inv);
			W.P = p.g1 * (U[4] - EKin);
			// below rhoMin the reference's P is 0 (its EKin 0 only feeds P);
			// wave-uniform test, so states above rhoMin pay no select
			if (wave_any(U[0] < p.rhoMin)) W.P = U[0] < p.rhoMin ? R(0) : W.P;
			return W;
		}

		__device__ static __forceinline__ R calc_Cs(const Par &p, const Prim &W) {
			if (W.P <= p.PMin) return R(0);
			if (W.rho < p.rhoMin) return R(INFINITY);
			return M::sqrt_pos(M::div_inv(p.gamma * W.P, W.rho, W.inv));   // P > PMin, rho >= rhoMin
		}

		// consWaveCodeMinMax (eqn.lua:1134-1146, euler.lua:329-337): the
		// state's slowest and fastest wave along S, v_n -+ calc_Cs_fromCons
		// (euler.cl:96-112), v_n = m_n / rho (0 below rhoMin)
		template <int S>
		__device__ static __forceinline__ void cons_waves(const Par &p, const R *U, const Prim &W, R &lmin, R &lmax) {
			R Cs = calc_Cs(p, W);
			R vn = U[0] < p.rhoMin ? R(0) : M::div_inv(U[1 + S], U[0], W.inv);
			lmin = vn - Cs;
			lmax = vn + Cs;
		}

		// eigen_forInterface (euler.cl:347-430) from the two states and their prims
		__device__ static __forceinline__ Eig eigen(const Par &p, const R *UL, const R *UR,
			const Prim &WL, const Prim &WR)
		{
			const R rhoEpsilon = R(1e-5);
			// every input is read once into a value up front; the branches below
			// only select among values (selecting between the two states'
			// addresses would move them to scratch memory)
			const R rhoL = UL[0], rhoR = UR[0], EL = UL[4], ER = UR[4];
			const R PL = WL.P, PR = WR.P, invL = WL.inv, invR = WR.inv;
			R vL[3], vR[3];
#pragma unroll
			for (int i = 0; i < 3; ++i) { vL[i] = WL.v[i]; vR[i] = WR.v[i]; }
			Eig e;
			if (rhoL < rhoEpsilon && rhoR < rhoEpsilon) {
				e.rho = 0; e.v[0] = e.v[1] = e.v[2] = 0; e.hTotal = 0; e.Cs = 0; e.vSq = 0;
				return e;
			}
			if (rhoL < rhoEpsilon || rhoR < rhoEpsilon) {
				// the non-vacuum side
				const bool useR = rhoL < rhoEpsilon;
				Prim W;
				W.rho = useR ? rhoR : rhoL;
				W.inv = useR ? invR : invL;
				W.P = useR ? PR : PL;
#pragma unroll
				for (int i = 0; i < 3; ++i) W.v[i] = useR ? vR[i] : vL[i];
				const R ETotalS = useR ? ER : EL;
				e.rho = W.rho;
				e.v[0] = W.v[0]; e.v[1] = W.v[1]; e.v[2] = W.v[2];
				e.vSq = lenSq3(W.v);
				e.hTotal = M::div_inv(W.P + ETotalS, W.rho, W.inv);
				e.Cs = calc_Cs(p, W);
				return e;
			}
			R sqrtRhoL = M::sqrt_pos(rhoL);      // both >= rhoEpsilon on this path
			R hTotalL = M::div_inv(PL + EL, rhoL, invL);
			R sqrtRhoR = M::sqrt_pos(rhoR);
			R hTotalR = M::div_inv(PR + ER, rhoR, invR);
			R invDenom = M::rcp(sqrtRhoL + sqrtRhoR);
			e.rho = sqrtRhoL * sqrtRhoR;
			R wl = sqrtRhoL * invDenom, wr = sqrtRhoR * invDenom;
#pragma unroll
			for (int i = 0; i < 3; ++i) e.v[i] = i < NVEL ? vL[i] * wl + vR[i] * wr : R(0);
			R hTotal = invDenom * (sqrtRhoL * hTotalL + sqrtRhoR * hTotalR);
			R vSq = lenSq3(e.v);
			R eKin = R(.5) * vSq;
			R h = hTotal - eKin;
			if (h < rhoEpsilon) {
				e.hTotal = eKin;
				e.Cs = 0;
			} else {
				e.hTotal = hTotal;
				e.Cs = M::sqrt_pos(p.g1 * h);        // h >= rhoEpsilon
			}
			e.vSq = vSq;
			return e;
		}

		// consFromPrim (euler.cl:161-173): m = v rho, ETotal = rho (v^2 / 2) + P / (gamma - 1)
		__device__ static __forceinline__ void cons_from_prim(const Par &p, const R *W, R *U) {
			U[0] = W[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) U[1 + i] = i < NVEL ? W[1 + i] * W[0] : R(0);
			U[4] = W[0] * (R(.5) * lenSq3(W + 1)) + M::div(W[4], p.g1);
		}

		// apply_dU_dW (euler.cl:179-195): the cons difference of the prim
		// difference dW about the state WA; prims as (rho, v, P)
		__device__ static __forceinline__ void apply_dU_dW(const Par &p, const R *WA, const R *dW, R *dU) {
			dU[0] = dW[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) dU[1 + i] = i < NVEL ? WA[1 + i] * dW[0] + dW[1 + i] * WA[0] : R(0);
			dU[4] = dW[0] * R(.5) * lenSq3(WA + 1) + WA[0] * dot3(dW + 1, WA + 1) + M::div(dW[4], p.g1);
		}

		// apply_dW_dU (euler.cl:201-223): the prim difference of the cons
		// difference dU about WA; velocity and pressure 0 where dU's density
		// is below rhoMin, as the reference tests it
		__device__ static __forceinline__ void apply_dW_dU(const Par &p, const R *WA, const R *dU, R *dW) {
			const bool low = dU[0] < p.rhoMin;
			const R invRho = M::rcp(WA[0]);
			const R rhoRatio = M::div(dU[0], WA[0]);
			dW[0] = dU[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) dW[1 + i] = i < NVEL && !low ? dU[1 + i] * invRho - WA[1 + i] * rhoRatio : R(0);
			const R P = p.g1 * (R(.5) * lenSq3(WA + 1) * dU[0] - dot3(dU + 1, WA + 1) + dU[4]);
			dW[4] = low ? R(0) : P;
		}

		// eigen_forCell (euler.cl:315-341): one cell's own state, no guards
		__device__ static __forceinline__ Eig eigen_cell(const Par &p, const R *U, const Prim &W) {
			Eig e;
			e.rho = W.rho;
#pragma unroll
			for (int i = 0; i < 3; ++i) e.v[i] = W.v[i];
			R vSq = lenSq3(W.v);
			R eKin = R(.5) * vSq;
			R hTotal = M::div_inv(W.P + U[4], W.rho, W.inv);   // calc_hTotal = (P + ETotal) / rho
			R CsSq = p.g1 * (hTotal - eKin);
			e.hTotal = hTotal;
			e.Cs = M::sqrt_(CsSq);
			e.vSq = vSq;
			return e;
		}

		// eigen_leftTransform (euler.cl:434-489), normal along axis S
		template <int S>
		__device__ static __forceinline__ void left(const Par &p, const Eig &e, const R *X, R *Y) {
			// identity below rhoMin: both forms are computed and, when any lane
			// of the wave is below, selected per element; every store into Y has
			// a constant index (a branch-merged store with a computed index
			// would move Y to scratch memory)
			const bool vac = e.rho < p.rhoMin;
			const bool anyVac = wave_any(vac);
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			const R Cs = e.Cs, vSq = e.vSq, g1 = p.g1;
			R denom = R(2) * Cs * Cs;
			R invDenom = M::rcp(denom);
			R hg1vSq = R(.5) * g1 * vSq;
			R Csvn = Cs * e.v[A];
			R t[3];
#pragma unroll
			for (int i = 0; i < 3; ++i) t[i] = -g1 * e.v[i];
			// (X0 (.5 g1 vSq +- Cs vn) + sum_i X_{1+i} (-g1 vL_i -+ Cs n1_i) + X4 g1) / (2 Cs^2)
			R y0 = X[0] * (hg1vSq + Csvn);
			R y4 = X[0] * (hg1vSq - Csvn);
#pragma unroll
			for (int i = 0; i < 3; ++i) {
				if (i < NVEL || i == A) {
					R cm = i == A ? t[i] - Cs : t[i];
					R cp = i == A ? t[i] + Cs : t[i];
					y0 = y0 + X[1 + i] * cm;
					y4 = y4 + X[1 + i] * cp;
				}
			}
			R z[5];
			z[0] = (y0 + X[4] * g1) * invDenom;
			z[4] = (y4 + X[4] * g1) * invDenom;
			R y1 = X[0] * (denom - g1 * vSq);
#pragma unroll
			for (int i = 0; i < 3; ++i)
				if (i < NVEL) y1 = y1 + X[1 + i] * R(2) * g1 * e.v[i];
			z[1] = (y1 + X[4] * R(-2) * g1) * invDenom;
			z[2] = B < NVEL ? X[0] * -e.v[B] + X[1 + B] : R(0);
			z[3] = C < NVEL ? X[0] * -e.v[C] + X[1 + C] : R(0);
			if (anyVac) {
#pragma unroll
				for (int i = 0; i < 5; ++i) Y[i] = vac ? X[i] : z[i];
			} else {
#pragma unroll
				for (int i = 0; i < 5; ++i) Y[i] = z[i];
			}
		}

		// eigen_rightTransform (euler.cl:493-542)
		template <int S>
		__device__ static __forceinline__ void right(const Par &p, const Eig &e, const R *X, R *Y) {
			const bool vac = e.rho < p.rhoMin;   // identity: selected per element, as in left
			const bool anyVac = wave_any(vac);
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			const R Cs = e.Cs;
			R z[5];
			z[0] = X[0] + X[1] + X[4];
			if constexpr (FAST) {
				// the same sums factored: z_n = vn z0 + c (X4 - X0), z_t = v_t z0 + X_t,
				// z4 = H (X0 + X4) + c vn (X4 - X0) + vSq/2 X1 + v_t . X_t
				const R vn = e.v[A], d40 = X[4] - X[0];
				z[1 + A] = vn * z[0] + Cs * d40;
				z[1 + B] = B < NVEL ? e.v[B] * z[0] + X[2] : R(0);
				z[1 + C] = C < NVEL ? e.v[C] * z[0] + X[3] : R(0);
				R y4 = e.hTotal * (X[0] + X[4]) + (Cs * vn) * d40 + (R(.5) * e.vSq) * X[1];
				if constexpr (B < NVEL) y4 = y4 + e.v[B] * X[2];
				if constexpr (C < NVEL) y4 = y4 + e.v[C] * X[3];
				if (anyVac) {
#pragma unroll
					for (int i = 0; i < 5; ++i) Y[i] = vac ? X[i] : (i == 4 ? y4 : z[i]);
				} else {
#pragma unroll
					for (int i = 0; i < 4; ++i) Y[i] = z[i];
					Y[4] = y4;
				}
				return;
			}
			z[1 + A] = X[0] * (e.v[A] - Cs) + X[1] * e.v[A] + X[4] * (e.v[A] + Cs);
			if constexpr (B < NVEL) z[1 + B] = X[0] * e.v[B] + X[1] * e.v[B] + X[2] + X[4] * e.v[B];
			else z[1 + B] = R(0);
			if constexpr (C < NVEL) z[1 + C] = X[0] * e.v[C] + X[1] * e.v[C] + X[3] + X[4] * e.v[C];
			else z[1 + C] = R(0);
			R Csvn = Cs * e.v[A];
			R y4 = X[0] * (e.hTotal - Csvn) + X[1] * R(.5) * e.vSq;
			if constexpr (B < NVEL) y4 = y4 + X[2] * e.v[B];
			if constexpr (C < NVEL) y4 = y4 + X[3] * e.v[C];
			z[4] = y4 + X[4] * (e.hTotal + Csvn);
			if (anyVac) {
#pragma unroll
				for (int i = 0; i < 5; ++i) Y[i] = vac ? X[i] : z[i];
			} else {
#pragma unroll
				for (int i = 0; i < 5; ++i) Y[i] = z[i];
			}
		}

		// fluxFromCons (euler.cl:273-291) given the state's prim
		template <int S>
		__device__ static __forceinline__ void flux(const Par &, const R *U, const Prim &W, R *F) {
			R vn = W.v[S];
			F[0] = U[0] * vn;
#pragma unroll
			for (int i = 0; i < 3; ++i) F[1 + i] = i < NVEL ? U[1 + i] * vn : R(0);
			F[1 + S] = F[1 + S] + W.P;
			F[4] = (U[4] + W.P) * vn;
		}

		// The Roe quantities of an interface whose states are regular: both
		// densities at least max(1e-5, rhoMin) and the Roe average's
		// h = hTotal - v^2/2 at least 1e-5, so that none of
		// eigen_forInterface's vacuum and cold branches (euler.cl:357-430) and
		// no identity transform (euler.cl:432-542) applies. FAST policy only:
		// the same functions of the two states as the reference's, with the
		// algebra collected:
		//   Roe weight of the left state  wl = 1 / (1 + sqrt(rhoR / rhoL)),
		//   v~ = vR + wl (vL - vR), hTotal~ likewise,
		//   L dU through dp = g1 (dE - v~.dm + v~^2/2 drho) and
		//   c (vn drho - dm_n): y0,4 = (dp +- c (vn drho - dm_n)) / (2 c^2),
		//   y1 = drho - dp / c^2, tangential y = dm_t - v_t drho,
		// c and 1 / c from one reciprocal square root. Computed whatever the
		// states; ok tells whether this lane's interface is regular (the
		// values are used only then: otherwise the caller takes the general
		// form, which has every branch).
		struct Collected {
			R vL[3], vR[3], PL, PR, EPL, EPR;   // the states' velocities, pressures, E + P
			R v[3], H, hq, c;                   // Roe velocity, hTotal, v^2 / 2, sound speed
			R y[5];                             // L (UR - UL)
			bool ok;
		};
		template <int S>
		__device__ static __forceinline__ Collected roe_collect(const Par &p, const R *UL, const R *UR) {
			static_assert(FAST, "the collected Roe arithmetic belongs to the FAST policy");
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			Collected q;
			const R rhoFloor = fmax(R(1e-5), p.rhoMin);
			const bool okRho = UL[0] >= rhoFloor && UR[0] >= rhoFloor;   // NaN: not regular
			const R iL = M::rcp(UL[0]), iR = M::rcp(UR[0]);
#pragma unroll
			for (int i = 0; i < 3; ++i) {
				q.vL[i] = i < NVEL ? UL[1 + i] * iL : R(0);
				q.vR[i] = i < NVEL ? UR[1 + i] * iR : R(0);
			}
			q.PL = p.g1 * (UL[4] - R(.5) * dot3(UL + 1, q.vL));
			q.PR = p.g1 * (UR[4] - R(.5) * dot3(UR + 1, q.vR));
			q.EPL = UL[4] + q.PL;
			q.EPR = UR[4] + q.PR;
			const R HL = q.EPL * iL, HR = q.EPR * iR;
			const R wl = M::rcp(R(1) + M::sqrt_pos(UR[0] * iL));
#pragma unroll
			for (int i = 0; i < 3; ++i) q.v[i] = i < NVEL ? q.vR[i] + wl * (q.vL[i] - q.vR[i]) : R(0);
			q.H = HR + wl * (HL - HR);
			q.hq = R(.5) * lenSq3(q.v);
			const R h = q.H - q.hq;
			q.ok = okRho && h >= R(1e-5);
			// c = sqrt(g1 h) and 1 / c: two Goldschmidt steps from v_rsq
			const R c2 = p.g1 * h;
			R s, hh;
			M::sqrt_and_half_rsqrt(c2, s, hh);
			const R c = s, ic2 = R(4) * hh * hh;   // 1 / c^2
			q.c = c;
			// L (UR - UL)
			R X[5];
#pragma unroll
			for (int j = 0; j < 5; ++j) X[j] = UR[j] - UL[j];
			R t = X[4] + q.hq * X[0];
#pragma unroll
			for (int i = 0; i < 3; ++i)
				if (i < NVEL) t = t - q.v[i] * X[1 + i];
			const R dp = p.g1 * t;
			const R vn = q.v[A];
			const R cB = c * (vn * X[0] - X[1 + A]);
			const R hic2 = R(.5) * ic2;
			q.y[0] = (dp + cB) * hic2;
			q.y[4] = (dp - cB) * hic2;
			q.y[1] = X[0] - dp * ic2;
			q.y[2] = B < NVEL ? X[1 + B] - q.v[B] * X[0] : R(0);
			q.y[3] = C < NVEL ? X[1 + C] - q.v[C] * X[0] : R(0);
			return q;
		}

		// The Roe flux F = (F(UL) + F(UR)) / 2 - R |Lambda| L (UR - UL) / 2
		// (roe.cl:17-163 without a flux limiter) of a regular interface, from
		// roe_collect, with R in the matching factored form. Returns whether
		// this lane's interface is regular (F holds the result only then).
		template <int S>
		__device__ static __forceinline__ bool roe_regular_lane(const Par &p, const R *UL, const R *UR, R *F) {
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			const Collected q = roe_collect<S>(p, UL, UR);
			const R c = q.c, vn = q.v[A];
			// |Lambda| L dU
			const R an = fabs(vn);
			const R f0 = fabs(vn - c) * q.y[0], f4 = fabs(vn + c) * q.y[4];
			const R f1 = an * q.y[1], fB = an * q.y[2], fC = an * q.y[3];
			// R (|Lambda| L dU)
			const R z0 = f0 + f1 + f4;
			const R d40 = f4 - f0;
			R zm[3];
			zm[A] = vn * z0 + c * d40;
			zm[B] = B < NVEL ? q.v[B] * z0 + fB : R(0);
			zm[C] = C < NVEL ? q.v[C] * z0 + fC : R(0);
			R z4 = q.H * (f0 + f4) + (c * vn) * d40 + q.hq * f1;
			if constexpr (B < NVEL) z4 = z4 + q.v[B] * fB;
			if constexpr (C < NVEL) z4 = z4 + q.v[C] * fC;
			// (F(UL) + F(UR) - z) / 2
			const R vnL = q.vL[A], vnR = q.vR[A];
			F[0] = R(.5) * ((UL[1 + A] + UR[1 + A]) - z0);
#pragma unroll
			for (int i = 0; i < 3; ++i) {
				if (i < NVEL || i == A) {
					R fs = UL[1 + i] * vnL + UR[1 + i] * vnR;
					if (i == A) fs = fs + (q.PL + q.PR);
					F[1 + i] = R(.5) * (fs - zm[i]);
				} else {
					F[1 + i] = R(0);
				}
			}
			F[4] = R(.5) * ((q.EPL * vnL + q.EPR * vnR) - z4);
			return q.ok;
		}

		// The interface quantities the flux-limited Roe flux (roe.cl:17-163)
		// keeps of a regular interface, from roe_collect: the Roe state as
		// eigen_forInterface gives it (rho only tells the transforms that the
		// state is above rhoMin), L (UR - UL), and the mean of the two
		// physical fluxes (euler.cl:273-291). Returns whether this lane's
		// interface is regular (the outputs hold the results only then).
		template <int S>
		__device__ static __forceinline__ bool roe_prep_lane(const Par &p, const R *UL, const R *UR, Eig &e, R *dW,
			R *Favg) {
			constexpr int A = S;
			const Collected q = roe_collect<S>(p, UL, UR);
			e.rho = UL[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) e.v[i] = q.v[i];
			e.hTotal = q.H;
			e.Cs = q.c;
			e.vSq = R(2) * q.hq;
#pragma unroll
			for (int j = 0; j < 5; ++j) dW[j] = q.y[j];
			const R vnL = q.vL[A], vnR = q.vR[A];
			Favg[0] = R(.5) * (UL[1 + A] + UR[1 + A]);
#pragma unroll
			for (int i = 0; i < 3; ++i) {
				if (i < NVEL || i == A) {
					R fs = UL[1 + i] * vnL + UR[1 + i] * vnR;
					if (i == A) fs = fs + (q.PL + q.PR);
					Favg[1 + i] = R(.5) * fs;
				} else {
					Favg[1 + i] = R(0);
				}
			}
			Favg[4] = R(.5) * (q.EPL * vnL + q.EPR * vnR);
			return q.ok;
		}

		// eigenWaveCode (euler.lua:309-325)
		template <int S>
		__device__ static __forceinline__ void waves(const Par &, const Eig &e, R *l) {
			R vn = e.v[S];
			l[0] = vn - e.Cs;
			l[1] = vn;
			l[2] = vn;
			l[3] = vn;
			l[4] = vn + e.Cs;
		}

		// constrainU (euler.cl:697-717) on a state in double
		__device__ static __forceinline__ void constrain(double gamma, double rhoMin, double PMin, double *U) {
			using MD = Math<double, FAST>;
			if (U[0] < rhoMin) U[0] = rhoMin;
			double g1 = gamma - 1.;
			double inv = MD::rcp(U[0]);
			double v[3];
#pragma unroll
			for (int i = 0; i < 3; ++i) v[i] = i < NVEL ? U[1 + i] * inv : 0.;
			double lenSq = lenSq3(U + 1);
			double EKin = U[0] < rhoMin ? 0. : MD::div_inv(.5 * lenSq, U[0], inv);
			double P = U[0] < rhoMin ? 0. : (g1 * (U[4] - EKin));
			if (P < PMin) P = PMin;
			double rho = U[0];
#pragma unroll
			for (int i = 0; i < 3; ++i) U[1 + i] = i < NVEL ? v[i] * rho : 0.;
			U[4] = rho * (.5 * lenSq3(v)) + MD::div_pos(P, g1);
		}

		// calcDTCell (calcDT.cl:38-73, euler.lua:346-373) in double
		template <int DIM>
		__device__ static __forceinline__ double dt_cell(const EqnParD &pd, const double *dx, const double *U) {
			using MD = Math<double, FAST>;
			const double gamma = pd.gamma, rhoMin = pd.rhoMin, PMin = pd.PMin;
			double inv = MD::rcp(U[0]);
			double lenSq = lenSq3(U + 1);
			double EKin = U[0] < rhoMin ? 0. : MD::div_inv(.5 * lenSq, U[0], inv);
			double P = U[0] < rhoMin ? 0. : ((gamma - 1.) * (U[4] - EKin));
			double Cs;
			if (P <= PMin) Cs = 0;
			else if (U[0] < rhoMin) Cs = INFINITY;
			else Cs = MD::sqrt_pos(MD::div_inv(gamma * P, U[0], inv));
			double dt = INFINITY;
#pragma unroll
			for (int s = 0; s < DIM; ++s) {
				if (dx[s] > 0.) {
					double vn = U[0] < rhoMin ? 0. : (s < NVEL ? MD::div_inv(U[1 + s], U[0], inv) : 0.);
					double lmin = vn - Cs, lmax = vn + Cs;
					double absLambdaMax = fmax(fabs(lmin), fabs(lmax));
					absLambdaMax = fmax(1e-9, absLambdaMax);
					dt = fmin(dt, MD::div(dx[s], absLambdaMax));
				}
			}
			return dt;
		}

		template <int DIM>
		__device__ static __forceinline__ void source(const Par &, const GridDesc<R> &, const R *, const R *,
			const R *, R *) {}

		// mirror reflects the normal momentum component
		__device__ static __forceinline__ bool reflects(int field, int axis) { return field == 1 + axis; }
	};
};

} // namespace hydro
