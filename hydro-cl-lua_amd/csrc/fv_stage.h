// fv_stage.h — the fused finite-volume stage kernel (1D / 2D).
//
// One launch performs one integrator stage of the reference's
//   calcLR -> calcFlux -> calcDerivFromFlux -> multAdd... -> boundary -> constrainU
// sequence (fvsolver.lua:225-302, rk.lua:152-183) for every interior cell, and
// on the last stage also the per-cell CFL dt of the new state and its
// min-reduction (calcDT.cl:38-73, solverbase.lua:3004-3023).
//
// Work decomposition: one 64-lane wave owns a strip of 64 columns, one output
// column per lane, and marches down the strip in y. x neighbours are exchanged
// with DPP wave shifts; y neighbours stay in registers (a sliding window of
// rows), so every state value is loaded once per stage from HBM/L2. The two
// interfaces at the strip's edges need cells outside the strip: before its
// rows, the wave computes them for all its rows at once (a lane per row and
// side, at most 32 rows) and keeps them in its own LDS slice, from which lanes
// 0 and 63 take them row by row. A 1D strip has one row, too few to repay that
// pass: there lanes 0-1 and 62-63 load a 2-cell halo and 60 lanes output.
// Ghost cells are not stored: loads outside the interior are mapped through
// the boundary conditions.
#pragma once

#include "fv_common.h"
#include "fv_euler.h"
#include "fv_glm.h"
#include "fv_limiters.h"
#include "fv_sw.h"

namespace hydro {

template <typename R, class E>
struct Cell {
	static constexpr int NV = E::Q + E::NSTAT;
	R v[NV];
};

// ---------------------------------------------------------- state loading
template <typename R, class E>
__device__ __forceinline__ void load_cell(const GridDesc<R> &g, const R *base, const Mapped &mx, const Mapped &my,
	Cell<R, E> &c)
{
	long off = (long)my.idx * g.step1 + mx.idx;
	// axis by axis, a later axis wins: the last fixed axis gives the value and
	// the mirrors of the axes after it flip it. One table lookup from a face
	// index chosen by value (selecting between two lookups would copy the
	// whole argument block to scratch memory).
	int fixedFace = -1, lastFixed = -1;
	if (mx.fixedFace >= 0) { fixedFace = mx.fixedFace; lastFixed = 0; }
	if (my.fixedFace >= 0) { fixedFace = my.fixedFace; lastFixed = 1; }
#pragma unroll
	for (int f = 0; f < Cell<R, E>::NV; ++f) {
		if (!E::active(f)) {
			c.v[f] = R(0);
			continue;
		}
		R v;
		if (f < E::Q) v = base[f * g.ps + off];
		else v = g.stat[(f - E::Q) * g.ps + off];
		R fv = g.fixedv[fixedFace < 0 ? 0 : fixedFace][f];
		v = fixedFace >= 0 ? fv : v;
		if (lastFixed < 0 && mx.mirror && E::reflects(f, 0)) v = -v;
		if (lastFixed < 1 && my.mirror && E::reflects(f, 1)) v = -v;
		c.v[f] = v;
	}
}

// A cell whose coordinates map to the interior or through periodic faces
// only: no fixed values or mirror flips (the common case, taken by a
// wave-uniform branch around load_cell).
template <typename R, class E>
__device__ __forceinline__ void load_plain(const GridDesc<R> &g, const R *base, long off, Cell<R, E> &c) {
#pragma unroll
	for (int f = 0; f < Cell<R, E>::NV; ++f) {
		if (!E::active(f)) c.v[f] = R(0);
		else if (f < E::Q) c.v[f] = base[f * g.ps + off];
		else c.v[f] = g.stat[(f - E::Q) * g.ps + off];
	}
}

template <typename R, class E>
__device__ __forceinline__ void shift_left(const Cell<R, E> &in, Cell<R, E> &out) {
#pragma unroll
	for (int f = 0; f < Cell<R, E>::NV; ++f) out.v[f] = from_left(in.v[f]);
}
template <typename R, class E>
__device__ __forceinline__ void shift_right(const Cell<R, E> &in, Cell<R, E> &out) {
#pragma unroll
	for (int f = 0; f < Cell<R, E>::NV; ++f) out.v[f] = from_right(in.v[f]);
}

// The half step of 'plm cons with flux' and 'plm eig' (plm.cl:169-188,
// 410-428): both edge states hL, hR add (dt/dx)/2 (F(hR) - F(hL)).
template <typename R, class E, int S>
__device__ __forceinline__ void plm_half_step(const typename E::Par &par, const Cell<R, E> &U, const R *hL,
	const R *hR, R dt_dx, Cell<R, E> &L, Cell<R, E> &Rr)
{
	constexpr int Q = E::Q;
	R FL[Q], FR[Q];
	E::template flux<S>(par, hL, E::prim(par, hL), FL);
	E::template flux<S>(par, hR, E::prim(par, hR), FR);
	const R h = R(.5) * dt_dx;
#pragma unroll
	for (int j = 0; j < Cell<R, E>::NV; ++j) {
		if (j < Q) {
			const R dF = FR[j] - FL[j];
			L.v[j] = hL[j] + h * dF;
			Rr.v[j] = hR[j] + h * dF;
		} else {
			L.v[j] = U.v[j];
			Rr.v[j] = U.v[j];
		}
	}
}

// ------------------------------------------------------------- plm cons
// plm.cl:27-91: one division per field (numerator / denominator chosen by the
// sign of dUC, as the reference's two branches do). S: the axis of the
// reconstruction, dt_dx = dt / dx along it (used by 'plm eig' only).
template <typename R, class E, int SLIM, int S>
__device__ __forceinline__ void plm_cons(const typename E::Par &par, int lim, const Cell<R, E> &Um, const Cell<R, E> &U,
	const Cell<R, E> &Up, Cell<R, E> &L, Cell<R, E> &Rr, R dt_dx)
{
	if constexpr (E::PLM_EIG) {
		// 'plm eig' (plm.cl:256-432, its enabled form): the slopes in the
		// characteristic variables of the cell's own eigensystem, limited
		// with r = dL / dR, each wave's slope kept on the side it moves
		// towards, back to cons, then half a step of the edge fluxes
		constexpr int Q = E::Q;
		R dUL[Q], dUR[Q];
#pragma unroll
		for (int j = 0; j < Q; ++j) {
			dUL[j] = U.v[j] - Um.v[j];
			dUR[j] = Up.v[j] - U.v[j];
		}
		const auto W = E::prim(par, U.v);
		const auto e = E::eigen_cell(par, U.v, W);
		R aL[E::NW], aR[E::NW], lam[E::NW];
		E::template left<S>(par, e, dUL, aL);
		E::template left<S>(par, e, dUR, aR);
		E::template waves<S>(par, e, lam);
#pragma unroll
		for (int j = 0; j < E::NW; ++j) {
			const R sigma = limited<SLIM, E::FASTM>(lim, aL[j], aR[j]);
			aR[j] = lam[j] >= R(0) ? R(0) : sigma;
			aL[j] = lam[j] <= R(0) ? R(0) : sigma;
		}
		R sL[Q], sR[Q];
		E::template right<S>(par, e, aL, sL);
		E::template right<S>(par, e, aR, sR);
		R hL[Q], hR[Q];
#pragma unroll
		for (int j = 0; j < Q; ++j) {
			hL[j] = U.v[j] - R(.5) * sL[j];
			hR[j] = U.v[j] + R(.5) * sR[j];
		}
		plm_half_step<R, E, S>(par, U, hL, hR, dt_dx, L, Rr);
		return;
	} else if constexpr (E::PLM_CONS_FLUX) {
		// 'plm cons with flux' (plm.cl:95-188): per field r = dUL / dUR,
		// sigma = phi(r) dUR, edges U -+ sigma / 2, then the half step
		constexpr int Q = E::Q;
		R hL[Q], hR[Q];
#pragma unroll
		for (int j = 0; j < Q; ++j) {
			const R dUL = U.v[j] - Um.v[j];
			const R dUR = Up.v[j] - U.v[j];
			const R sigma = limited<SLIM, E::FASTM>(lim, dUL, dUR);
			hL[j] = U.v[j] - R(.5) * sigma;
			hR[j] = U.v[j] + R(.5) * sigma;
		}
		plm_half_step<R, E, S>(par, U, hL, hR, dt_dx, L, Rr);
		return;
	} else if constexpr (E::PLM_EIG_PRIM || E::PLM_EIG_PRIM_REF) {
		// 'plm eig prim' / 'plm eig prim ref' (plm.cl:536-780): primitive
		// differences mapped by dU/dW into the cell's characteristic
		// variables, the doubled minmod-of-three slope, per-wave slopes for
		// each edge (with the reference states of the 'ref' form), back by
		// dW/dU, and the edge primitives to cons
		const auto W = E::prim(par, U.v), Wm = E::prim(par, Um.v), Wp = E::prim(par, Up.v);
		const auto e = E::eigen_cell(par, U.v, W);
		const R w0[5] = {W.rho, W.v[0], W.v[1], W.v[2], W.P};
		const R wm[5] = {Wm.rho, Wm.v[0], Wm.v[1], Wm.v[2], Wm.P};
		const R wp[5] = {Wp.rho, Wp.v[0], Wp.v[1], Wp.v[2], Wp.P};
		R dWL[5], dWR[5], dWC[5];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			dWL[j] = w0[j] - wm[j];
			dWR[j] = wp[j] - w0[j];
			dWC[j] = R(.5) * (wp[j] - wm[j]);
		}
		R t[5], eL[5], eR[5], eC[5];
		E::apply_dU_dW(par, w0, dWL, t);
		E::template left<S>(par, e, t, eL);
		E::apply_dU_dW(par, w0, dWR, t);
		E::template left<S>(par, e, t, eR);
		E::apply_dU_dW(par, w0, dWC, t);
		E::template left<S>(par, e, t, eC);
		R dM[5], lam[5];
#pragma unroll
		for (int j = 0; j < 5; ++j)
			dM[j] = eL[j] * eR[j] < R(0) ? R(0)
				: (eC[j] >= R(0) ? R(1) : R(-1)) * R(2) * fmin(fmin(fabs(eL[j]), fabs(eR[j])), fabs(eC[j]));
		E::template waves<S>(par, e, lam);
		R aL[5], aR[5], sL[5], sR[5], w2L[5], w2R[5];
		if constexpr (E::PLM_EIG_PRIM) {
#pragma unroll
			for (int j = 0; j < 5; ++j) {
				aL[j] = lam[j] < R(0) ? R(0) : dM[j] * R(.5) * (R(1) - lam[j] * dt_dx);
				aR[j] = lam[j] > R(0) ? R(0) : dM[j] * R(.5) * (R(1) + lam[j] * dt_dx);
			}
			E::template right<S>(par, e, aL, t);
			E::apply_dW_dU(par, w0, t, sL);
			E::template right<S>(par, e, aR, t);
			E::apply_dW_dU(par, w0, t, sR);
#pragma unroll
			for (int j = 0; j < 5; ++j) {
				w2L[j] = w0[j] + sL[j];
				w2R[j] = w0[j] - sR[j];
			}
		} else {
			const R waveMin = fmin(R(0), lam[0]), waveMax = fmax(R(0), lam[4]);
			R dWM[5];
			E::template right<S>(par, e, dM, t);
			E::apply_dW_dU(par, w0, t, dWM);
#pragma unroll
			for (int j = 0; j < 5; ++j) {
				aL[j] = lam[j] < R(0) ? R(0) : dM[j] * dt_dx * (waveMax - lam[j]);
				aR[j] = lam[j] > R(0) ? R(0) : dM[j] * dt_dx * (waveMin - lam[j]);
			}
			E::template right<S>(par, e, aL, t);
			E::apply_dW_dU(par, w0, t, sL);
			E::template right<S>(par, e, aR, t);
			E::apply_dW_dU(par, w0, t, sR);
#pragma unroll
			for (int j = 0; j < 5; ++j) {
				const R wLRef = w0[j] + R(.5) * (R(1) - dt_dx * waveMax) * dWM[j];
				const R wRRef = w0[j] - R(.5) * (R(1) + dt_dx * waveMin) * dWM[j];
				w2R[j] = wRRef + R(.5) * sR[j];
				w2L[j] = wLRef + R(.5) * sL[j];
			}
		}
		E::cons_from_prim(par, w2L, L.v);
		E::cons_from_prim(par, w2R, Rr.v);
		return;
	} else if constexpr (E::PLM_ATHENA) {
		// 'plm athena' (plm.cl:782-879; the later branch of that name is never
		// reached): differences of the primitives put through the cell's cons
		// eigensystem as they are, the Athena slope (the harmonic mean dWG
		// beside left, right and center), back, each edge value clamped
		// between the cell and its neighbour and the other edge moved to keep
		// their sum; the cell's left edge takes the right value's cons and the
		// right edge the left value's, as the reference stores them
		using M = typename E::M;
		const auto W = E::prim(par, U.v), Wm = E::prim(par, Um.v), Wp = E::prim(par, Up.v);
		const auto e = E::eigen_cell(par, U.v, W);
		const R w0[5] = {W.rho, W.v[0], W.v[1], W.v[2], W.P};
		const R wm[5] = {Wm.rho, Wm.v[0], Wm.v[1], Wm.v[2], Wm.P};
		const R wp[5] = {Wp.rho, Wp.v[0], Wp.v[1], Wp.v[2], Wp.P};
		R dWL[5], dWR[5], dWC[5], dWG[5];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			dWL[j] = w0[j] - wm[j];
			dWR[j] = wp[j] - w0[j];
			dWC[j] = R(.5) * (wp[j] - wm[j]);
			dWG[j] = dWL[j] * dWR[j] <= R(0) ? R(0) : M::div(R(2) * dWL[j] * dWR[j], dWL[j] + dWR[j]);
		}
		R dal[5], dar[5], dac[5], dag[5];
		E::template left<S>(par, e, dWL, dal);
		E::template left<S>(par, e, dWR, dar);
		E::template left<S>(par, e, dWC, dac);
		E::template left<S>(par, e, dWG, dag);
		R da[5];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			da[j] = R(0);
			if (dal[j] * dar[j] > R(0)) {
				const R s1 = fmin(fabs(dal[j]), fabs(dar[j]));
				const R s2 = fmin(fabs(dac[j]), fabs(dag[j]));
				// OpenCL sign(): +-1, or the zero itself
				const R sg = dac[j] > R(0) ? R(1) : dac[j] < R(0) ? R(-1) : dac[j];
				da[j] = sg * fmin(R(2) * s1, s2);
			}
		}
		R dWm[5];
		E::template right<S>(par, e, da, dWm);
		R wl[5], wr[5];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			R lv = w0[j] - R(.5) * dWm[j];
			R rv = w0[j] + R(.5) * dWm[j];
			const R C = rv + lv;
			lv = fmax(fmin(w0[j], wm[j]), lv);
			lv = fmin(fmax(w0[j], wm[j]), lv);
			rv = C - lv;
			rv = fmax(fmin(w0[j], wp[j]), rv);
			rv = fmin(fmax(w0[j], wp[j]), rv);
			lv = C - rv;
			wl[j] = lv;
			wr[j] = rv;
		}
		E::cons_from_prim(par, wr, L.v);
		E::cons_from_prim(par, wl, Rr.v);
		return;
	} else if constexpr (E::PLM_PPM) {
		// 'ppm' (plm.cl:882-944): the kappa = 1/3 edges of the three-cell
		// stencil scaled by phi(dUL / dUR); the second neighbours' differences
		// it takes feed no output
		using M = typename E::M;
		constexpr R kappa = R(1) / R(3);
#pragma unroll
		for (int j = 0; j < Cell<R, E>::NV; ++j) {
			if (j < E::Q) {
				const R dUL = U.v[j] - Um.v[j];
				const R dUR = Up.v[j] - U.v[j];
				const R r = dUR == R(0) ? R(0) : (E::FASTM ? M::div(dUL, dUR) : dUL / dUR);
				const R phi = limiter<SLIM>(lim, r);
				Rr.v[j] = U.v[j] + R(.25) * phi * ((R(1) - kappa) * dUL + (R(1) + kappa) * dUR);
				L.v[j] = U.v[j] - R(.25) * phi * ((R(1) - kappa) * dUR + (R(1) + kappa) * dUL);
			} else {
				L.v[j] = U.v[j];
				Rr.v[j] = U.v[j];
			}
		}
		return;
	} else if constexpr (E::PLM_PRIM) {
		// 'plm prim' (plm.cl:191-254): the same slopes of the primitives
		// (rho, v, P), r = dWL / dWR always, back to cons at the two edges
		const auto Wm = E::prim(par, Um.v), W0 = E::prim(par, U.v), Wp = E::prim(par, Up.v);
		const R wm[5] = {Wm.rho, Wm.v[0], Wm.v[1], Wm.v[2], Wm.P};
		const R w0[5] = {W0.rho, W0.v[0], W0.v[1], W0.v[2], W0.P};
		const R wp[5] = {Wp.rho, Wp.v[0], Wp.v[1], Wp.v[2], Wp.P};
		R nL[5], nR[5];
#pragma unroll
		for (int j = 0; j < 5; ++j) {
			const R dWR = wp[j] - w0[j];
			const R dWL = w0[j] - wm[j];
			const R sigma = limited<SLIM, E::FASTM>(lim, dWL, dWR);
			nL[j] = w0[j] - R(.5) * sigma;
			nR[j] = w0[j] + R(.5) * sigma;
		}
		E::cons_from_prim(par, nL, L.v);
		E::cons_from_prim(par, nR, Rr.v);
		return;
	} else {
#pragma unroll
	for (int j = 0; j < Cell<R, E>::NV; ++j) {
		if (j < E::Q) {
			R dUR = Up.v[j] - U.v[j];
			R dUL = U.v[j] - Um.v[j];
			R dUC = R(.5) * (dUR - dUL);
			R sigma;
			if constexpr (E::FASTM && limited_symmetric(SLIM)) {
				// phi(num/den) den is symmetric in its operands for these
				// limiters: the choice by the sign of dUC drops out
				sigma = limited<SLIM, true>(lim, dUL, dUR);
			} else {
				bool pos = dUC >= R(0);
				R num = pos ? dUL : dUR;
				R den = pos ? dUR : dUL;
				sigma = limited<SLIM, E::FASTM>(lim, num, den);
			}
			Rr.v[j] = U.v[j] + R(.5) * sigma;
			L.v[j] = U.v[j] - R(.5) * sigma;
		} else {
			L.v[j] = U.v[j];
			Rr.v[j] = U.v[j];
		}
	}
	}
}

// ------------------------------------------------------------------ Roe
// Interface quantities shared by neighbouring interfaces when the flux
// limiter is on: the Roe state, L * (UR - UL), and the mean physical flux.
template <typename R, class E>
struct Iface {
	typename E::Eig e;
	R dW[E::NW];
	R Favg[E::Q];
};

// The FAST kernels of the Euler equations with the Roe flux prepare a regular
// interface in the collected form (EulerT::roe_prep_lane) and the others in
// the general one, lane by lane.
template <class E>
__host__ __device__ constexpr bool roe_collected_prep() {
	return E::FASTM && E::FLUXID == HYDRO_FLUX_ROE && E::Q == 5 && E::NW == 5;
}

// eigen_forInterface, eigen_leftTransform of UR - UL and the mean of the
// physical fluxes, with every branch of the reference (euler.cl:347-542)
template <typename R, class E, int S>
__device__ __forceinline__ void iface_prep_general(const typename E::Par &p, const Cell<R, E> &UL,
	const Cell<R, E> &UR, Iface<R, E> &I)
{
	auto WL = E::prim(p, UL.v);
	auto WR = E::prim(p, UR.v);
	I.e = E::eigen(p, UL.v, UR.v, WL, WR);
	R dU[E::Q];
#pragma unroll
	for (int j = 0; j < E::Q; ++j) dU[j] = UR.v[j] - UL.v[j];
	E::template left<S>(p, I.e, dU, I.dW);
	R FL[E::Q], FR[E::Q];
	E::template flux<S>(p, UL.v, WL, FL);
	E::template flux<S>(p, UR.v, WR, FR);
#pragma unroll
	for (int j = 0; j < E::Q; ++j) I.Favg[j] = R(.5) * (FL[j] + FR[j]);
}

template <typename R, class E, int S>
__device__ __forceinline__ void iface_prep(const typename E::Par &p, const Cell<R, E> &UL,
	const Cell<R, E> &UR, Iface<R, E> &I)
{
	if constexpr (roe_collected_prep<E>()) {
		const bool ok = E::template roe_prep_lane<S>(p, UL.v, UR.v, I.e, I.dW, I.Favg);
		// the irregular lanes overwrite their own results (the wave-wide
		// tests inside the general form only skip work none of their lanes needs)
		if (!ok) iface_prep_general<R, E, S>(p, UL, UR, I);
	} else {
		iface_prep_general<R, E, S>(p, UL, UR, I);
	}
}

// roe.cl:91-160
template <typename R, class E, int S, bool LIM, int FLIM>
__device__ __forceinline__ void iface_flux(const typename E::Par &p, const Iface<R, E> &I,
	const R *dWL, const R *dWR, R dt_dx, int lim, R *F)
{
	R fluxEig[E::NW], lam[E::NW];
	E::template waves<S>(p, I.e, lam);
#pragma unroll
	for (int j = 0; j < E::NW; ++j) {
		R lambda = lam[j];
		if constexpr (E::FLUXID == FLUX_ROE_ENTROPY_FIX) {
			// Harten's entropy fix (roe.cl:83-106): |lambda| below dws is
			// replaced by (lambda^2/dws + dws)/2 with OpenCL's sign (0 at 0)
			const R dws = p.fixFactor * I.e.Cs;
			if (fabs(lambda) < dws) {
				const R sgn = lambda > R(0) ? R(1) : (lambda < R(0) ? R(-1) : lambda);
				lambda = sgn * R(.5) * (E::M::div(lambda * lambda, dws) + dws);
			}
		}
		R sgnLambda = lambda >= R(0) ? R(1) : R(-1);
		if constexpr (LIM && E::FASTM) {
			// -lambda/2 (dW sgn + phi(r) dW (lambda dt/dx - sgn)), phi(r) dW without r
			R up = lambda >= R(0) ? dWL[j] : dWR[j];
			R phidW = limited<FLIM, true>(lim, up, I.dW[j]);
			fluxEig[j] = R(0) - R(.5) * lambda * (I.dW[j] * sgnLambda + phidW * (lambda * dt_dx - sgnLambda));
		} else if constexpr (LIM) {
			R rEig;
			if (I.dW[j] == R(0)) rEig = R(0);
			else rEig = lambda >= R(0) ? dWL[j] / I.dW[j] : dWR[j] / I.dW[j];
			R phi = limiter<FLIM>(lim, rEig);
			fluxEig[j] = R(0) - R(.5) * lambda * I.dW[j] * (sgnLambda + phi * (lambda * dt_dx - sgnLambda));
		} else {
			// = 0 - .5 lambda dW sgn(lambda), bit for bit: scaling by +-1 commutes
			// with round-to-nearest, and 0 - (+-0) is +0 either way
			fluxEig[j] = R(0) - R(.5) * fabs(lambda) * I.dW[j];
		}
	}
	R Y[E::Q];
	E::template right<S>(p, I.e, fluxEig, Y);
#pragma unroll
	for (int j = 0; j < E::Q; ++j) F[j] = Y[j] + I.Favg[j];
}

// ------------------------------------------------ the Roe flux's two forms
// How a FAST kernel of the Euler equations with the Roe flux and no flux
// limiter computes an interface: ROE_INLINE takes, lane by lane, the collected
// form where the interface is regular and the general form (every branch of
// eigen_forInterface and the transforms) where it is not; ROE_FLAGGED takes
// the collected form only and reports per lane whether the interface is
// regular, so that the stage kernel leaves a strip from its first row with an
// irregular interface on to the redo kernel (stage_kernel_redo), which runs
// ROE_INLINE (1D and 3D kernels take ROE_INLINE throughout). Either way each
// interface's form depends on its own states only, so a grid advances to the
// same bits whatever its strips' heights and whether it runs alone or
// batched. Keeping the general form out of the 2D stage kernel keeps its
// registers at the collected form's (127 VGPRs FP64: 4 waves per SIMD,
// against 186 with both forms).
enum { ROE_GENERAL = 0, ROE_INLINE = 1, ROE_FLAGGED = 2 };

template <class E>
__host__ __device__ constexpr bool roe_collected(bool lim) {
	return E::FASTM && E::FLUXID == HYDRO_FLUX_ROE && E::Q == 5 && E::NW == 5 && !lim;
}

// The stage kernels that take ROE_FLAGGED and a redo launch: 2D. A 1D strip
// is one row, so the general form costs it few registers (94 VGPRs FP64
// against 74) and saves a launch per stage; 1D kernels take ROE_INLINE.
template <class E, int DIM>
__host__ __device__ constexpr bool redo_form(bool lim) {
	return roe_collected<E>(lim) && DIM == 2;
}

// ------------------------------------------------- one interface, no limiter
// The interface flux of the equation's flux plugin without a flux limiter:
// Roe (roe.cl:17-163), HLL with the 'Davis direct bounded' estimates
// (hll.cl:13-73, hll.lua:10), Rusanov (rusanov.cl:13-39, whose dissipation
// speed is the right state's: the reference assigns it per side in turn), or
// HLLC (euler-hllc.cl:13-240, hllcMethod 2, HLL's estimates), or Marquina
// (marquina.cl:5-93). The cases of HLL and HLLC are selects over all
// candidate fluxes, in the reference's expression order each.
template <typename R, class E, int S, int RM = ROE_INLINE>
__device__ __forceinline__ void iface_solve(const typename E::Par &p, const Cell<R, E> &UL, const Cell<R, E> &UR,
	R dt_dx, int lim, R *F, bool *ok = nullptr)
{
	if (ok) *ok = true;
	if constexpr (E::FLUXID == HYDRO_FLUX_ROE || E::FLUXID == FLUX_ROE_ENTROPY_FIX) {
		if constexpr (roe_collected<E>(false) && RM == ROE_FLAGGED) {
			*ok = E::template roe_regular_lane<S>(p, UL.v, UR.v, F);
			return;
		} else if constexpr (roe_collected<E>(false) && RM == ROE_INLINE) {
			// a regular interface: the collected form
			if (E::template roe_regular_lane<S>(p, UL.v, UR.v, F)) return;
		}
		Iface<R, E> I;
		iface_prep_general<R, E, S>(p, UL, UR, I);
		iface_flux<R, E, S, false, 0>(p, I, nullptr, nullptr, dt_dx, lim, F);
	} else {
		using M = typename E::M;
		auto WL = E::prim(p, UL.v);
		auto WR = E::prim(p, UR.v);
		R FL[E::Q], FR[E::Q];
		E::template flux<S>(p, UL.v, WL, FL);
		E::template flux<S>(p, UR.v, WR, FR);
		R lminL, lmaxL, lminR, lmaxR;
		E::template cons_waves<S>(p, UL.v, WL, lminL, lmaxL);
		E::template cons_waves<S>(p, UR.v, WR, lminR, lmaxR);
		if constexpr (E::FLUXID == HYDRO_FLUX_HLL) {
			auto e = E::eigen(p, UL.v, UR.v, WL, WR);
			R lam[E::NW];
			E::template waves<S>(p, e, lam);
			const R sL = fmin(lminL, lam[0]);
			const R sR = fmax(lmaxR, lam[E::NW - 1]);
			const R inv = M::rcp(sR - sL);
#pragma unroll
			for (int j = 0; j < E::Q; ++j) {
				R mid = M::div_inv(sR * FL[j] - sL * FR[j] + sL * sR * (UR.v[j] - UL.v[j]), sR - sL, inv);
				F[j] = R(0) <= sL ? FL[j] : (sR <= R(0) ? FR[j] : mid);
			}
		} else if constexpr (E::FLUXID == HYDRO_FLUX_MARQUINA) {
			// each side's state and flux in its own eigenbasis; per wave,
			// upwind when both sides' speeds share a sign, else the local
			// Lax-Friedrichs split with the larger |lambda|
			auto eL = E::eigen_cell(p, UL.v, WL);
			auto eR = E::eigen_cell(p, UR.v, WR);
			R omegaL[E::NW], omegaR[E::NW], phiL[E::NW], phiR[E::NW];
			E::template left<S>(p, eL, UL.v, omegaL);
			E::template left<S>(p, eR, UR.v, omegaR);
			E::template left<S>(p, eL, FL, phiL);
			E::template left<S>(p, eR, FR, phiR);
			R lL[E::NW], lR[E::NW];
			E::template waves<S>(p, eL, lL);
			E::template waves<S>(p, eR, lR);
			R pp[E::NW], pm[E::NW];
#pragma unroll
			for (int k = 0; k < E::NW; ++k) {
				const bool same = lL[k] * lR[k] > R(0);
				const bool right = lL[k] > R(0);
				const R alpha = fmax(fabs(lL[k]), fabs(lR[k]));
				const R splitP = (phiL[k] + alpha * omegaL[k]) * R(.5);
				const R splitM = (phiR[k] - alpha * omegaR[k]) * R(.5);
				pp[k] = same ? (right ? phiL[k] : R(0)) : splitP;
				pm[k] = same ? (right ? R(0) : phiR[k]) : splitM;
			}
			R YL[E::Q], YR[E::Q];
			E::template right<S>(p, eL, pp, YL);
			E::template right<S>(p, eR, pm, YR);
#pragma unroll
			for (int j = 0; j < E::Q; ++j) F[j] = YL[j] + YR[j];
		} else if constexpr (E::FLUXID == HYDRO_FLUX_HLLC) {
			auto e = E::eigen(p, UL.v, UR.v, WL, WR);
			R lam[E::NW];
			E::template waves<S>(p, e, lam);
			const R sL = fmin(lminL, lam[0]);
			const R sR = fmax(lmaxR, lam[E::NW - 1]);
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;   // normal basis (coord.lua:2393-2408)
			const R vnL = WL.v[A], vnR = WR.v[A];
			const R sStar = M::div(WR.rho * vnR * (sR - vnR) - WL.rho * vnL * (sL - vnL) + WL.P - WR.P,
				WR.rho * (sR - vnR) - WL.rho * (sL - vnL));
			const R PLR = R(.5) * (WL.P + WR.P + WL.rho * (sL - vnL) * (sStar - vnL)
				+ WR.rho * (sR - vnR) * (sStar - vnR));
			const R dL = sL - sStar, dR = sR - sStar;
			const R iL = M::rcp(dL), iR = M::rcp(dR);
			R FsL[E::Q], FsR[E::Q];
			FsL[0] = M::div_inv((sL * UL.v[0] - FL[0]) * sStar, dL, iL);
			FsL[1 + A] = M::div_inv((sL * UL.v[1 + A] - FL[1 + A]) * sStar + sL * PLR, dL, iL);
			FsL[1 + B] = M::div_inv(sStar * (sL * UL.v[1 + B] - FL[1 + B]), dL, iL);
			FsL[1 + C] = M::div_inv(sStar * (sL * UL.v[1 + C] - FL[1 + C]), dL, iL);
			FsL[4] = M::div_inv(sStar * (sL * UL.v[4] - FL[4]) + sL * PLR * sStar, dL, iL);
			FsR[0] = M::div_inv(sStar * (sR * UR.v[0] - FR[0]), dR, iR);
			FsR[1 + A] = M::div_inv(sStar * (sR * UR.v[1 + A] - FR[1 + A]) + sR * PLR, dR, iR);
			FsR[1 + B] = M::div_inv(sStar * (sR * UR.v[1 + B] - FR[1 + B]), dR, iR);
			FsR[1 + C] = M::div_inv(sStar * (sR * UR.v[1 + C] - FR[1 + C]), dR, iR);
			FsR[4] = M::div_inv(sStar * (sR * UR.v[4] - FR[4]) + sR * PLR * sStar, dR, iR);
			const bool useFL = R(0) <= sL;
			const bool useL = sL <= R(0) && R(0) <= sStar;
			const bool useR = sStar <= R(0) && R(0) <= sR;
			const bool useFR = sR <= R(0);
#pragma unroll
			for (int j = 0; j < E::Q; ++j) F[j] = useFL ? FL[j] : useL ? FsL[j] : useR ? FsR[j] : FR[j];
			// HLL where sStar is not a number (and the NaN case): a wave-uniform
			// branch, as no lane of a finite state takes it
			const bool hllCase = !(useFL || useL || useR || useFR);
			if (wave_any(hllCase)) {
				const R dLR = sR - sL, iLR = M::rcp(dLR);
#pragma unroll
				for (int j = 0; j < E::Q; ++j) {
					R hll = M::div_inv(sR * FL[j] - sL * FR[j] + sL * sR * (UR.v[j] - UL.v[j]), dLR, iLR);
					F[j] = hllCase ? hll : F[j];
				}
			}
		} else {
			const R a = fmax(fabs(lminR), fabs(lmaxR));
#pragma unroll
			for (int j = 0; j < E::Q; ++j) F[j] = R(.5) * (FL[j] + FR[j] - a * (UR.v[j] - UL.v[j]));
		}
	}
}

// ------------------------------------------------------------ flux guards
// The reference zeroes a flux where the cell area is <= 1e-7 and the whole
// derivative where the volume is (fvsolver.lua:105, fvsolver.cl:97). EXACT
// kernels apply both as selects; FAST kernels fold them into one factor per
// side (GridDesc::aovFast), one multiply and no selects.
template <class E, typename R>
__device__ __forceinline__ bool flux_on(const GridDesc<R> &g, int side) {
	return E::FASTM ? true : g.fluxOn[side] != 0;
}
template <class E, typename R>
__device__ __forceinline__ R flux_aov(const GridDesc<R> &g, int side) {
	return E::FASTM ? g.aovFast[side] : g.areaOverVol[side];
}

// ------------------------------------------------------------ strip edges
// One wave's LDS slice for up to NR rows: for row j of the wave, slot j
// describes the strip's left edge interface and slot NR + j its right edge
// interface: the area-weighted x flux, the neighbour cell outside the strip
// (needed by PLM slopes and by a source term reading x neighbours), and the
// interface's characteristic jump dW (needed by the flux limiter).
template <typename R, class E, bool PLM, bool LIM, int NR>
struct XEdge {
	static constexpr bool U = PLM || E::SOURCE;
	static constexpr int S = 2 * NR;
	R fa[E::Q][S];
	R u[U ? Cell<R, E>::NV : 1][S];
	R dw[LIM ? E::NW : 1][S];
	int ok[S];                // the interface is regular (ROE_FLAGGED kernels)
};

// x flux (area-weighted) and dW of the interface xb - 1/2 from the cells
// xb-2 .. xb+1 of one row: the same arithmetic as the lanes' own interfaces
template <typename R, class E, bool PLM, bool LIM, int FLIM, int SLIM, int RM = ROE_INLINE>
__device__ __forceinline__ void edge_flux(const typename E::Par &par, int lim, int slim, const Cell<R, E> *c, R dt_dx0,
	bool fluxOn, R aov, R *Fa, R *dW, bool &ok)
{
	constexpr int Q = E::Q;
	R F[Q];
	ok = true;
	if constexpr (PLM) {
		Cell<R, E> L1, R1, L2, R2;
		plm_cons<R, E, SLIM, 0>(par, slim, c[0], c[1], c[2], L1, R1, dt_dx0);
		plm_cons<R, E, SLIM, 0>(par, slim, c[1], c[2], c[3], L2, R2, dt_dx0);
		iface_solve<R, E, 0, RM>(par, R1, L2, dt_dx0, lim, F, &ok);
#pragma unroll
		for (int j = 0; j < E::NW; ++j) dW[j] = R(0);   // kept by the flux limiter only
	} else if constexpr (LIM) {
		Iface<R, E> Im, I, Ip;
		iface_prep<R, E, 0>(par, c[0], c[1], Im);
		iface_prep<R, E, 0>(par, c[1], c[2], I);
		iface_prep<R, E, 0>(par, c[2], c[3], Ip);
		iface_flux<R, E, 0, true, FLIM>(par, I, Im.dW, Ip.dW, dt_dx0, lim, F);
#pragma unroll
		for (int j = 0; j < E::NW; ++j) dW[j] = I.dW[j];
	} else {
		iface_solve<R, E, 0, RM>(par, c[1], c[2], dt_dx0, lim, F, &ok);
#pragma unroll
		for (int j = 0; j < E::NW; ++j) dW[j] = R(0);
	}
#pragma unroll
	for (int j = 0; j < Q; ++j) Fa[j] = fluxOn ? F[j] * aov : R(0);
}

// LDS stores of a wave, visible to its own lanes' later loads
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Edge pass of a strip: lane j < NR computes the left edge interface of the
// wave's row j, lane NR + j the right one (lanes from 2 NR on repeat them and
// store nothing); cb holds the lane's cells xb-2 .. xb+1 around its
// interface xb - 1/2.
template <typename R, class E, bool PLM, bool LIM, int FLIM, int SLIM, int NR, int RM = ROE_INLINE>
__device__ __forceinline__ void edge_pass(const typename E::Par &par, int lim, int slim, const Cell<R, E> *cb, R dt_dx0,
	bool fluxOn, R aov, int lane, XEdge<R, E, PLM, LIM, NR> &edge)
{
	constexpr int Q = E::Q;
	R Fa[Q], dW[E::NW];
	bool ok;
	edge_flux<R, E, PLM, LIM, FLIM, SLIM, RM>(par, lim, slim, cb, dt_dx0, fluxOn, aov, Fa, dW, ok);
	const bool right = (lane & NR) != 0;
	wave_lds_sync();   // the previous item's reads (persistent waves) are done
	if (lane < 2 * NR) {
		edge.ok[lane] = ok ? 1 : 0;
#pragma unroll
		for (int j = 0; j < Q; ++j) edge.fa[j][lane] = Fa[j];
		if constexpr (XEdge<R, E, PLM, LIM, NR>::U) {
#pragma unroll
			for (int f = 0; f < Cell<R, E>::NV; ++f) edge.u[f][lane] = right ? cb[2].v[f] : cb[1].v[f];
		}
		if constexpr (LIM) {
#pragma unroll
			for (int j = 0; j < E::NW; ++j) edge.dw[j][lane] = dW[j];
		}
	}
	wave_lds_sync();
}

// x part of a row: the x flux divergence (area-weighted, -(F(x+1/2) - F(x-1/2)))
// of row U, lanes exchanging neighbours by DPP and, with EDGES, the strip
// edges (lanes 0 and 63) taken from the edge slots slotL and slotR of edge
// (the cell outside the right edge from edgeR's slot; a wave's own slice
// holds both) (without, lanes 0-1 and 62-63 are a halo whose results are not
// used). Um / Up return the x neighbours.
template <typename R, class E, bool PLM, bool LIM, int FLIM, int SLIM, int NR, bool EDGES = true, int RM = ROE_INLINE>
__device__ __forceinline__ void strip_x(const typename E::Par &par, int lim, int slim, const Cell<R, E> &U, R dt_dx0,
	bool fluxOn, R aov, int lane, int slotL, int slotR, const XEdge<R, E, PLM, LIM, NR> &edge,
	const XEdge<R, E, PLM, LIM, NR> &edgeR, Cell<R, E> &Um, Cell<R, E> &Up, R *divx, bool &okx)
{
	using C = Cell<R, E>;
	constexpr int Q = E::Q;
	okx = true;
	const int slot = lane == WAVE - 1 ? slotR : slotL;
	if constexpr (EDGES && XEdge<R, E, PLM, LIM, NR>::U) {
		// lane 0 reads the left edge slot, lane 63 the right one: the shifts'
		// empty lanes take them
		const XEdge<R, E, PLM, LIM, NR> &eu = lane == WAVE - 1 ? edgeR : edge;
#pragma unroll
		for (int f = 0; f < C::NV; ++f) {
			R e = eu.u[f][slot];
			Um.v[f] = from_left_or(U.v[f], e);
			Up.v[f] = from_right_or(U.v[f], e);
		}
	} else {
		shift_left<R, E>(U, Um);
		shift_right<R, E>(U, Up);
	}
	if constexpr (EDGES && !XEdge<R, E, PLM, LIM, NR>::U) {
		// lane 0's own interface is replaced by the edge's; a valid state
		// (its own) in place of the zero fill keeps it out of the vacuum paths
#pragma unroll
		for (int f = 0; f < C::NV; ++f) Um.v[f] = lane == 0 ? U.v[f] : Um.v[f];
	}
	R Fx[Q];
	if constexpr (PLM) {
		C Lx, Rx, RxL;
		plm_cons<R, E, SLIM, 0>(par, slim, Um, U, Up, Lx, Rx, dt_dx0);
		shift_left<R, E>(Rx, RxL);
		if constexpr (EDGES) {
			// as above: lane 0's interface is replaced, its left state only
			// needs to be valid
#pragma unroll
			for (int f = 0; f < C::NV; ++f) RxL.v[f] = lane == 0 ? Rx.v[f] : RxL.v[f];
		}
		iface_solve<R, E, 0, RM>(par, RxL, Lx, dt_dx0, lim, Fx, &okx);
	} else if constexpr (LIM) {
		Iface<R, E> I;
		iface_prep<R, E, 0>(par, Um, U, I);
		R dWL[E::NW], dWR[E::NW];
#pragma unroll
		for (int k = 0; k < E::NW; ++k) {
			if constexpr (EDGES) {
				// lane 0's own interface is the left edge one: its dW (from
				// the cell outside the strip) is lane 1's left neighbour jump
				R e = edge.dw[k][slot];
				dWL[k] = from_left(lane == 0 ? e : I.dW[k]);
				dWR[k] = from_right_or(I.dW[k], e);
			} else {
				dWL[k] = from_left(I.dW[k]);
				dWR[k] = from_right(I.dW[k]);
			}
		}
		iface_flux<R, E, 0, true, FLIM>(par, I, dWL, dWR, dt_dx0, lim, Fx);
	} else {
		iface_solve<R, E, 0, RM>(par, Um, U, dt_dx0, lim, Fx, &okx);
	}
	if constexpr (RM == ROE_FLAGGED && EDGES) {
		// lane 0's interface is the left edge one; lane 63 also needs the right edge
		const bool eok = edge.ok[slot] != 0;
		okx = lane == 0 ? eok : (lane == WAVE - 1 ? okx && eok : okx);
	} else if constexpr (RM == ROE_FLAGGED) {
		// halo strip: only the interfaces of the output lanes NG..63-NG count
		// (the outer halo lanes' left states come from the zero fill)
		okx = okx || lane < NG || lane > WAVE - NG;
	}
#pragma unroll
	for (int k = 0; k < Q; ++k) {
		R Fa = fluxOn ? Fx[k] * aov : R(0);
		R FaR;
		if constexpr (EDGES) {
			R ef = edge.fa[k][slot];             // lane 0: left edge, lane 63: right edge
			Fa = lane == 0 ? ef : Fa;
			FaR = from_right_or(Fa, ef);
		} else {
			FaR = from_right(Fa);
		}
		divx[k] = R(0) - (FaR - Fa);
	}
}

// ------------------------------------------------------- one wave strip
// Advances the strip (sx, sy) of grid g by one stage: 64 output columns (one
// per lane; 1D: 60) by rowsPerWave (<= 32) rows. Returns the strip's dt minimum of the new
// state when the stage reduces dt.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM,
	int RM = ROE_INLINE>
__device__ __forceinline__ double stage_strip(const StageArgs<R> &a, const GridDesc<R> &g, int sx, int sy, double dt,
	int lane, XEdge<R, EQ<R>, PLM, LIM, EDGE_ROWS> *edges, int w, bool share, int yStart = -1)
{
	using E = EQ<R>;
	using C = Cell<R, E>;
	constexpr int Q = E::Q;
	const typename E::Par par = E::par(a);
	const int lim = PLM ? 0 : a.limiter;   // runtime limiter index when FLIM/SLIM == -1
	const int slim = a.limiter;
	double dtLocal = INFINITY;

	const int nx = g.n[0];
	const int ny = DIM >= 2 ? g.n[1] : 1;
	const R *src = g.U[a.src];
	R *out = a.out >= 0 ? g.U[a.out] : g.ping;
	R *storeD = a.storeD >= 0 ? g.D[a.storeD] : nullptr;
	constexpr bool EDGES = strip_edges(DIM, E::SOURCE);   // edge pass, else a halo
	const int x = EDGES ? sx * STRIP_OUT + lane : sx * strip_out(false) - NG + lane;
	const bool outLane = EDGES ? x < nx : lane >= NG && lane < WAVE - NG && x < nx;
	const Mapped mx = map_coord(x, nx, g.bc[0], g.bc[1], 0);
	Mapped myIn;
	myIn.idx = 0; myIn.mirror = 0; myIn.fixedFace = -1;
	const R rdt = R(dt);
	const R dt_dx0 = rdt / g.dx[0];
	const R dt_dx1 = DIM >= 2 ? rdt / g.dx[1] : R(0);
	// the strip's rows; the redo kernel starts at the first row it takes over
	const int y0 = yStart >= 0 ? yStart : sy * g.rowsPerWave;
	const int y1 = min(sy * g.rowsPerWave + g.rowsPerWave, ny);

	// ---- the strip's edge interfaces for rows y0 .. y1-1: lane j (< 32) the
	// left edge of row y0 + j, lane 32 + j the right edge. Shared (plain 2D
	// launch, a block's four waves x-adjacent strips of one row band): wave 0
	// computes the block's five edge columns, lane c * rows + j the interface
	// left of strip sx + c in row y0 + j, into slice 0 (the right cells of the
	// interfaces into slice 1); one barrier publishes them.
	using XE = XEdge<R, E, PLM, LIM, EDGE_ROWS>;
	XE &edge = edges[share ? 0 : w];
	const XE &edgeR = edges[share ? 1 : w];
	const int rpw = g.rowsPerWave;
	if constexpr (EDGES) {
		if (share) {
			if (w == 0) {
				int c = lane / rpw, j = lane - c * rpw;
				if (c > 4) { c = 4; j = 0; }       // lanes past the five columns repeat one
				const int yb = y0 + (j < y1 - y0 ? j : 0);
				const int xb = (sx + c) * STRIP_OUT;   // interface xb - 1/2
				const Mapped myb = map_coord(yb, ny, g.bc[2], g.bc[3], 2);
				Mapped mxb[4];
				bool plain = myb.fixedFace < 0 && !myb.mirror;
#pragma unroll
				for (int k = 0; k < 4; ++k) {
					mxb[k] = map_coord(xb - 2 + k, nx, g.bc[0], g.bc[1], 0);
					plain = plain && mxb[k].fixedFace < 0 && !mxb[k].mirror;
				}
				C cb[4];
				if (wave_all(plain)) {
#pragma unroll
					for (int k = 0; k < 4; ++k) load_plain<R, E>(g, src, (long)myb.idx * g.step1 + mxb[k].idx, cb[k]);
				} else {
#pragma unroll
					for (int k = 0; k < 4; ++k) load_cell<R, E>(g, src, mxb[k], myb, cb[k]);
				}
				R Fa[Q], dW[E::NW];
				bool ok;
				edge_flux<R, E, PLM, LIM, FLIM, SLIM, RM>(par, lim, slim, cb, dt_dx0, flux_on<E>(g, 0),
					flux_aov<E>(g, 0), Fa, dW, ok);
				edges[0].ok[lane] = ok ? 1 : 0;
#pragma unroll
				for (int q = 0; q < Q; ++q) edges[0].fa[q][lane] = Fa[q];
				if constexpr (XE::U) {
#pragma unroll
					for (int f = 0; f < C::NV; ++f) {
						edges[0].u[f][lane] = cb[1].v[f];
						edges[1].u[f][lane] = cb[2].v[f];
					}
				}
				if constexpr (LIM) {
#pragma unroll
					for (int q = 0; q < E::NW; ++q) edges[0].dw[q][lane] = dW[q];
				}
			}
			__syncthreads();
		} else {
			const int bj = lane & (EDGE_ROWS - 1);
			const bool right = lane >= EDGE_ROWS;
			const int yb = y0 + (bj < y1 - y0 ? bj : 0);
			const int xb = sx * STRIP_OUT + (right ? STRIP_OUT : 0);   // interface xb - 1/2
			const Mapped myb = DIM >= 2 ? map_coord(yb, ny, g.bc[2], g.bc[3], 2) : myIn;
			Mapped mxb[4];
			bool plain = myb.fixedFace < 0 && !myb.mirror;
#pragma unroll
			for (int k = 0; k < 4; ++k) {
				mxb[k] = map_coord(xb - 2 + k, nx, g.bc[0], g.bc[1], 0);
				plain = plain && mxb[k].fixedFace < 0 && !mxb[k].mirror;
			}
			C cb[4];
			if (wave_all(plain)) {
#pragma unroll
				for (int k = 0; k < 4; ++k) load_plain<R, E>(g, src, (long)myb.idx * g.step1 + mxb[k].idx, cb[k]);
			} else {
#pragma unroll
				for (int k = 0; k < 4; ++k) load_cell<R, E>(g, src, mxb[k], myb, cb[k]);
			}
			edge_pass<R, E, PLM, LIM, FLIM, SLIM, EDGE_ROWS, RM>(par, lim, slim, cb, dt_dx0, flux_on<E>(g, 0),
				flux_aov<E>(g, 0), lane, edge);
		}
	}

	// y pipeline state
	C Wm, W0, W1, W2;       // rows r-1 .. r+2
	C Ry0;                  // PLM: R edge of row r (from the previous iteration)
	Iface<R, E> I1;         // LIM: interface r+1/2
	R dW0[E::NW];           // LIM: dW of interface r-1/2
	R FaYprev[Q];
	bool okYprev = true;    // ROE_FLAGGED: interface r-1/2 is regular
#pragma unroll
	for (int f = 0; f < C::NV; ++f) { Wm.v[f] = W0.v[f] = W1.v[f] = W2.v[f] = Ry0.v[f] = R(0); }
#pragma unroll
	for (int j = 0; j < Q; ++j) FaYprev[j] = R(0);
#pragma unroll
	for (int j = 0; j < E::NW; ++j) dW0[j] = R(0);
	I1 = Iface<R, E>{};

	// the wave's x coordinates need no fixed values or mirror flips
	const bool xPlain = wave_all(mx.fixedFace < 0 && !mx.mirror);
	auto loadRow = [&](int r, C &c) {
		Mapped my = DIM >= 2 ? map_coord(r, ny, g.bc[2], g.bc[3], 2) : myIn;
		if (xPlain && my.fixedFace < 0 && !my.mirror) load_plain<R, E>(g, src, (long)my.idx * g.step1 + mx.idx, c);
		else load_cell<R, E>(g, src, mx, my, c);
	};

	// iterations before y0 only fill the pipeline
	constexpr int WARM = DIM < 2 ? 0 : (PLM ? 2 : (LIM ? 3 : 1));
	if constexpr (DIM >= 2) {
		if constexpr (PLM) { loadRow(y0 - 2, W0); loadRow(y0 - 1, W1); }
		else if constexpr (LIM) { loadRow(y0 - 2, W1); }
		else { loadRow(y0 - 1, W0); }
	}

	// two passes: the warm-up rows (y pipeline only), then the output rows, so
	// that the output rows' body is one block with the y and x work together
#pragma unroll
	for (int pass = 0; pass < 2; ++pass) {
	const int rBegin = pass == 0 ? y0 - WARM : y0;
	const int rEnd = pass == 0 ? y0 : y1;
	for (int r = rBegin; r < rEnd; ++r) {
		// ---- y direction: interface r+1/2, dyterm of row r
		R dyterm[Q];
		bool okY = true;
		if constexpr (DIM >= 2) {
			R F[Q];
			// the first PLM warm-up iteration's interface is never used
			const bool wantFlux = !(PLM && r == y0 - WARM);
			if constexpr (PLM) {
				// W0 = U(r), W1 = U(r+1): cell r+1's edges, interface r+1/2
				loadRow(r + 2, W2);
				C L1, R1;
				plm_cons<R, E, SLIM, 1>(par, slim, W0, W1, W2, L1, R1, dt_dx1);
				if (wantFlux) {
					iface_solve<R, E, 1, RM>(par, Ry0, L1, dt_dx1, lim, F, &okY);
				} else {
#pragma unroll
					for (int j = 0; j < Q; ++j) F[j] = R(0);
				}
				Ry0 = R1;
			} else if constexpr (LIM) {
				// W1 = U(r+1): interface r+3/2, flux at r+1/2 from its neighbours
				loadRow(r + 2, W2);
				Iface<R, E> I2;
				iface_prep<R, E, 1>(par, W1, W2, I2);
				iface_flux<R, E, 1, true, FLIM>(par, I1, dW0, I2.dW, dt_dx1, lim, F);
#pragma unroll
				for (int j = 0; j < E::NW; ++j) dW0[j] = I1.dW[j];
				I1 = I2;
			} else {
				// W0 = U(r): interface r+1/2
				loadRow(r + 1, W1);
				iface_solve<R, E, 1, RM>(par, W0, W1, dt_dx1, lim, F, &okY);
			}
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				R FaY = flux_on<E>(g, 1) ? F[j] * flux_aov<E>(g, 1) : R(0);
				dyterm[j] = FaY - FaYprev[j];
				FaYprev[j] = FaY;
			}
		}

		if (pass == 1) {
			// ---- x direction on row r (lanes)
			C U;
			if constexpr (DIM >= 2) U = W0;
			else loadRow(0, U);
			C Um, Up;
			R deriv[Q];
			bool okX;
			const int jr = r - y0;
			strip_x<R, E, PLM, LIM, FLIM, SLIM, EDGE_ROWS, EDGES, RM>(par, lim, slim, U, dt_dx0, flux_on<E>(g, 0),
				flux_aov<E>(g, 0), lane, share ? w * rpw + jr : jr, share ? (w + 1) * rpw + jr : EDGE_ROWS + jr, edge,
				edgeR, Um, Up, deriv, okX);
			if constexpr (RM == ROE_FLAGGED) {
				// an irregular interface of this row: the redo kernel takes the
				// strip from this row on; nothing of it is stored here
				if (!wave_all(okYprev && okY && okX)) {
					if (lane == 0) {
						const unsigned k = atomicAdd(&g.dev->redoCount, 1u);
						g.redo[k] = make_int2(sy * g.stripsX + sx, r);
					}
					break;
				}
			}
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				if constexpr (DIM >= 2) deriv[j] = deriv[j] - dyterm[j];
				if (!E::FASTM && !g.derivOn) deriv[j] = R(0);
			}
			if constexpr (E::SOURCE) {
				// eqn addSource after calcDerivFromFlux (solverbase.lua:3208-3214)
				R em[3], ep[3];
				constexpr int SF = E::SRC_FIELD;
				em[0] = Um.v[SF];
				ep[0] = Up.v[SF];
				if constexpr (DIM >= 2) { em[1] = Wm.v[SF]; ep[1] = W1.v[SF]; }
				E::template source<DIM>(par, g, U.v, em, ep, deriv);
			}

			// ---- integrator combine (rk.lua:155-174, fe.lua:50), constrainU
			if (outLane) {
				long off = (long)r * g.step1 + x;
				double acc[Q];
#pragma unroll
				for (int j = 0; j < Q; ++j) acc[j] = 0.;
				for (int k = 0; k < a.nAlpha; ++k) {
					const bool fromSrc = a.accU[k] == a.src;
					const R *Uk = g.U[a.accU[k]];
					double al = a.alpha[k];
#pragma unroll
					for (int j = 0; j < Q; ++j) {
						if (!E::active(j)) continue;
						R u = fromSrc ? U.v[j] : Uk[j * g.ps + off];
						acc[j] = acc[j] + (double)u * al;
					}
				}
				for (int k = 0; k < a.nStoredD; ++k) {
					const R *Dk = g.D[a.accD[k]];
					double c = (double)R(a.beta[k] * dt);
#pragma unroll
					for (int j = 0; j < Q; ++j)
						if (E::active(j)) acc[j] = acc[j] + (double)Dk[j * g.ps + off] * c;
				}
				if (storeD) {
#pragma unroll
					for (int j = 0; j < Q; ++j)
						if (E::active(j)) storeD[j * g.ps + off] = deriv[j];
				}
				if (a.betaNew != 0.) {
					double c = (double)R(a.betaNew * dt);
#pragma unroll
					for (int j = 0; j < Q; ++j)
						if (E::active(j)) acc[j] = acc[j] + (double)deriv[j] * c;
				}
				E::constrain(a.pd.gamma, a.pd.rhoMin, a.pd.PMin, acc);
				R o[Q];
#pragma unroll
				for (int j = 0; j < Q; ++j) {
					o[j] = R(acc[j]);
					if (E::active(j)) out[j * g.ps + off] = o[j];
				}
				if (a.reduceSlot >= 0) {
					double ud[C::NV];
#pragma unroll
					for (int j = 0; j < Q; ++j) ud[j] = (double)o[j];
#pragma unroll
					for (int j = Q; j < C::NV; ++j) ud[j] = (double)U.v[j];
					double d = E::template dt_cell<DIM>(a.pd, g.dxD, ud);
					dtLocal = fmin(dtLocal, d);
				}
			}
		}

		// ---- advance the row window
		okYprev = okY;
		if constexpr (DIM >= 2) {
			if constexpr (E::SOURCE) Wm = W0;
			if constexpr (PLM || LIM) { W0 = W1; W1 = W2; }
			else { W0 = W1; }
		}
	}
	}
	return dtLocal;
}

// dt of this step for grid g (solverbase.lua:3016): cfl * min, or a host value
template <typename R>
__device__ __forceinline__ double stage_dt(const StageArgs<R> &a, const GridDesc<R> &g) {
	return a.dtSlot >= 0 ? g.cfl * read_dt_slot(g.dev, a.dtSlot) : a.hostDt;
}

// Per-grid bookkeeping done by the wave that owns the grid's first strip:
// clear the next-but-one dt slot (first stage) and advance t (last stage).
template <typename R>
__device__ __forceinline__ void stage_first_strip(const StageArgs<R> &a, const GridDesc<R> &g, double dt, int lane) {
	if (a.resetSlot >= 0) g.dev->dtmin[a.resetSlot][lane] = INFINITY;
	if (a.advanceTime && lane == 0) {
		g.dev->t = g.dev->t + dt;
		g.dev->dt = dt;
	}
}

template <typename R>
__device__ __forceinline__ void stage_reduce(const StageArgs<R> &a, const GridDesc<R> &g, double dtLocal, long item,
	int lane)
{
	if (a.reduceSlot >= 0) {
		double m = wave_min(dtLocal);
		if (lane == 0 && m < INFINITY) atomic_min_pos(&g.dev->dtmin[a.reduceSlot][item & (DT_LANES - 1)], m);
	}
}

// ---------------------------------------------------------- the kernels
// Plain form: one grid, one wave per strip. The 2D FAST kernels of the Euler
// equations with the Roe flux and no flux limiter take the collected Roe form
// only (ROE_FLAGGED) and leave any strip with an irregular interface, from
// that row on, to stage_kernel_redo, launched after them on the same stream.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(256) stage_kernel(const StageArgs<R> a)
{
	constexpr int RM = redo_form<EQ<R>, DIM>(LIM) ? ROE_FLAGGED : ROE_INLINE;
	const int lane = threadIdx.x & 63;
	int blk = a.xcdRemap ? xcd_block(blockIdx.x, a.nBlocks) : (int)blockIdx.x;
	const int waveInBlock = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const long item = (long)blk * (blockDim.x >> 6) + waveInBlock;
	__shared__ XEdge<R, EQ<R>, PLM, LIM, EDGE_ROWS> edges[strip_edges(DIM, EQ<R>::SOURCE) ? 4 : 1];
	const GridDesc<R> &g = a.g;
	double dt = stage_dt(a, g);
	if (item >= g.nItems) return;     // whole waves leave; no barriers follow
	const int sx = (int)(item % g.stripsX);
	const int sy = (int)(item / g.stripsX);
	if (item == 0) stage_first_strip(a, g, dt, lane);
	// shared edges: the block's four waves are strips sx0 .. sx0+3 of one row
	// band (the host sets the flag only where the strips per row are a
	// multiple of 4, so every block is whole)
	const bool share = DIM == 2 && strip_edges(DIM, EQ<R>::SOURCE) && a.shareEdges != 0 && blockDim.x == 256;
	double dtLocal = stage_strip<R, EQ, DIM, PLM, LIM, FLIM, SLIM, RM>(a, g, sx, sy, dt, lane, edges, waveInBlock,
		share);
	stage_reduce(a, g, dtLocal, item, lane);
}

// The strips a stage kernel left (from their first irregular row on), with
// the Roe form chosen per interface (ROE_INLINE): the entries of every grid's
// redo list (one grid for the plain form, a.grids for the batched one) are
// numbered one after the other, and wave k takes entries k, k + W, k + 2W, ...
// (W = waves launched). The rows it takes were not written by the stage
// kernel, so their inputs are as the stage began; their dt joins the same
// slot. The lists' lengths are read 64 grids at a time, one per lane; when
// every list is empty (the usual case) the kernel ends there, with nothing to
// clear. Otherwise the last block to finish empties the lists for the next
// stage (every wave has read the lengths by then); no wave depends on another.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(256) stage_kernel_redo(const StageArgs<R> a)
{
	const int lane = threadIdx.x & 63;
	const int waveInBlock = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const long wave = (long)blockIdx.x * (blockDim.x >> 6) + waveInBlock;
	const long nWaves = (long)gridDim.x * (blockDim.x >> 6);
	__shared__ XEdge<R, EQ<R>, PLM, LIM, EDGE_ROWS> edges[4];
	__shared__ unsigned lastBlock;
	const bool batched = a.grids != nullptr;
	const int nGrids = batched ? a.nGrids : 1;
	long base = 0;     // entries of the grids before the current one
	for (int g0 = 0; g0 < nGrids; g0 += 64) {
		const int gi = g0 + lane;
		unsigned c = 0u;
		if (gi < nGrids) c = (batched ? a.grids[gi].dev : a.g.dev)->redoCount;
		unsigned long long pending = __ballot(c != 0u);
		while (pending) {
			const int j = __ffsll((long long)pending) - 1;
			pending &= pending - 1;
			const long n = (long)__shfl(c, j);
			const GridDesc<R> &g = batched ? a.grids[g0 + j] : a.g;
			const double dt = stage_dt(a, g);
			// this wave's first entry at or after base
			for (long k = base + ((wave - base) % nWaves + nWaves) % nWaves; k < base + n; k += nWaves) {
				const int2 e = g.redo[k - base];
				const int sx = e.x % g.stripsX, sy = e.x / g.stripsX;
				double dtLocal = stage_strip<R, EQ, DIM, PLM, LIM, FLIM, SLIM, ROE_INLINE>(a, g, sx, sy, dt, lane,
					edges, waveInBlock, false, e.y);
				stage_reduce(a, g, dtLocal, e.x, lane);
			}
			base += n;
		}
	}
	if (base == 0) return;     // the same for every wave of the launch
	DevScalars *d = batched ? a.grids[0].dev : a.g.dev;
	__syncthreads();
	if (threadIdx.x == 0) lastBlock = atomicAdd(&d->redoDone, 1u) == gridDim.x - 1 ? 1u : 0u;
	__syncthreads();
	if (lastBlock && threadIdx.x < 64) {
		for (int gi = lane; gi < nGrids; gi += 64) (batched ? a.grids[gi].dev : a.g.dev)->redoCount = 0u;
		if (lane == 0) d->redoDone = 0u;
	}
}

// Dynamic LDS of the batched kernel: the strip-edge slices of its waves
// (waves per block x sizeof(XEdge), set at launch; none for halo strips).
extern __shared__ __align__(16) unsigned char batch_lds[];

// Batched form: a persistent launch over every strip of every active grid of
// the batch; wave k takes items k, k + W, k + 2W, ... (W = waves launched).
// Items are numbered grid by grid (GridDesc::itemBase), so the owning grid of
// an item is found by a wave-uniform binary search. The Roe form is the plain
// kernel's (ROE_FLAGGED where it applies, with the same redo kernel after it),
// so a grid advances to the same bits batched or alone.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(256) stage_kernel_batch(const StageArgs<R> a)
{
	constexpr int RM = redo_form<EQ<R>, DIM>(LIM) ? ROE_FLAGGED : ROE_INLINE;
	const int lane = threadIdx.x & 63;
	int blk = a.xcdRemap ? xcd_block(blockIdx.x, a.nBlocks) : (int)blockIdx.x;
	const int waveInBlock = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const long wave = (long)blk * (blockDim.x >> 6) + waveInBlock;
	const long nWaves = (long)gridDim.x * (blockDim.x >> 6);
	auto *edges = reinterpret_cast<XEdge<R, EQ<R>, PLM, LIM, EDGE_ROWS> *>(batch_lds);
	int gi = 0;
	for (long item = wave; item < a.totalItems; item += nWaves) {
		// grids are ordered by itemBase; items only increase, so search upward
		int lo = gi, hi = a.nGrids - 1;
		while (lo < hi) {
			int mid = (lo + hi + 1) >> 1;
			if (a.grids[mid].itemBase <= item) lo = mid;
			else hi = mid - 1;
		}
		gi = lo;
		const GridDesc<R> &g = a.grids[gi];
		if (!g.active || g.stepsLeft <= a.stepInTick) continue;
		const long local = item - g.itemBase;
		double dt = stage_dt(a, g);
		const int sx = (int)(local % g.stripsX);
		const int sy = (int)(local / g.stripsX);
		if (local == 0) stage_first_strip(a, g, dt, lane);
		double dtLocal = stage_strip<R, EQ, DIM, PLM, LIM, FLIM, SLIM, RM>(a, g, sx, sy, dt, lane, edges, waveInBlock,
			false);
		stage_reduce(a, g, dtLocal, local, lane);
	}
}

} // namespace hydro
