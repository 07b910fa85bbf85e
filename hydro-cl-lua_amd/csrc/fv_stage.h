// fv_stage.h — the fused finite-volume stage kernel (1D / 2D).
//
// One launch performs one integrator stage of the reference's
//   calcLR -> calcFlux -> calcDerivFromFlux -> multAdd... -> boundary -> constrainU
// sequence (fvsolver.lua:225-302, rk.lua:152-183) for every interior cell, and
// on the last stage also the per-cell CFL dt of the new state and its
// min-reduction (calcDT.cl:38-73, solverbase.lua:3004-3023).
//
// Work decomposition: one 64-lane wave owns a strip of 64 columns (60 output
// columns plus a 2-cell halo each side) and marches down the strip in y.
// x neighbours are exchanged with DPP wave shifts; y neighbours stay in
// registers (a sliding window of rows), so every state value is loaded once
// per stage from HBM/L2 and no LDS or barriers are needed. Ghost cells are not
// stored: loads outside the interior are mapped through the boundary
// conditions.
#pragma once

#include "fv_common.h"
#include "fv_euler.h"
#include "fv_glm.h"
#include "fv_limiters.h"

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

// ------------------------------------------------------------- plm cons
// plm.cl:27-91: one division per field (numerator / denominator chosen by the
// sign of dUC, as the reference's two branches do)
template <typename R, class E, int SLIM>
__device__ __forceinline__ void plm_cons(int lim, const Cell<R, E> &Um, const Cell<R, E> &U,
	const Cell<R, E> &Up, Cell<R, E> &L, Cell<R, E> &Rr)
{
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

// ------------------------------------------------------------------ Roe
// Interface quantities shared by neighbouring interfaces when the flux
// limiter is on: the Roe state, L * (UR - UL), and the mean physical flux.
template <typename R, class E>
struct Iface {
	typename E::Eig e;
	R dW[E::NW];
	R lambda[E::NW];
	R Favg[E::Q];
};

template <typename R, class E, int S>
__device__ __forceinline__ void iface_prep(const typename E::Par &p, const Cell<R, E> &UL,
	const Cell<R, E> &UR, Iface<R, E> &I)
{
	auto WL = E::prim(p, UL.v);
	auto WR = E::prim(p, UR.v);
	I.e = E::eigen(p, UL.v, UR.v, WL, WR);
	E::template waves<S>(p, I.e, I.lambda);
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

// roe.cl:91-160
template <typename R, class E, int S, bool LIM, int FLIM>
__device__ __forceinline__ void iface_flux(const typename E::Par &p, const Iface<R, E> &I,
	const R *dWL, const R *dWR, R dt_dx, int lim, R *F)
{
	R fluxEig[E::NW];
#pragma unroll
	for (int j = 0; j < E::NW; ++j) {
		R lambda = I.lambda[j];
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

// ------------------------------------------------------- one wave strip
// Advances the strip (sx, sy) of grid g by one stage: 60 output columns
// (lanes 2..61) by rowsPerWave rows. Returns the strip's dt minimum of the new
// state when the stage reduces dt.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__device__ __forceinline__ double stage_strip(const StageArgs<R> &a, const GridDesc<R> &g, int sx, int sy, double dt,
	int lane)
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
	const int x = sx * STRIP_OUT - NG + lane;
	const bool outLane = lane >= NG && lane < WAVE - NG && x < nx;
	const Mapped mx = map_coord(x, nx, g.bc[0], g.bc[1], 0);
	Mapped myIn;
	myIn.idx = 0; myIn.mirror = 0; myIn.fixedFace = -1;
	const R rdt = R(dt);
	const R dt_dx0 = rdt / g.dx[0];
	const R dt_dx1 = DIM >= 2 ? rdt / g.dx[1] : R(0);
	const int y0 = sy * g.rowsPerWave;
	const int y1 = min(y0 + g.rowsPerWave, ny);

	// y pipeline state
	C Wm, W0, W1, W2;       // rows r-1 .. r+2
	C Ry0;                  // PLM: R edge of row r (from the previous iteration)
	Iface<R, E> I1;         // LIM: interface r+1/2
	R dW0[E::NW];           // LIM: dW of interface r-1/2
	R FaYprev[Q];
#pragma unroll
	for (int f = 0; f < C::NV; ++f) { Wm.v[f] = W0.v[f] = W1.v[f] = W2.v[f] = Ry0.v[f] = R(0); }
#pragma unroll
	for (int j = 0; j < Q; ++j) FaYprev[j] = R(0);
#pragma unroll
	for (int j = 0; j < E::NW; ++j) dW0[j] = R(0);
	I1 = Iface<R, E>{};

	auto loadRow = [&](int r, C &c) {
		Mapped my = DIM >= 2 ? map_coord(r, ny, g.bc[2], g.bc[3], 2) : myIn;
		load_cell<R, E>(g, src, mx, my, c);
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
		if constexpr (DIM >= 2) {
			R F[Q];
			// the first PLM warm-up iteration's interface is never used
			const bool wantFlux = !(PLM && r == y0 - WARM);
			if constexpr (PLM) {
				// W0 = U(r), W1 = U(r+1): cell r+1's edges, interface r+1/2
				loadRow(r + 2, W2);
				C L1, R1;
				plm_cons<R, E, SLIM>(slim, W0, W1, W2, L1, R1);
				if (wantFlux) {
					Iface<R, E> I;
					iface_prep<R, E, 1>(par, Ry0, L1, I);
					iface_flux<R, E, 1, false, 0>(par, I, nullptr, nullptr, dt_dx1, lim, F);
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
				Iface<R, E> I;
				iface_prep<R, E, 1>(par, W0, W1, I);
				iface_flux<R, E, 1, false, 0>(par, I, nullptr, nullptr, dt_dx1, lim, F);
			}
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				R FaY = g.fluxOn[1] ? F[j] * g.areaOverVol[1] : R(0);
				dyterm[j] = FaY - FaYprev[j];
				FaYprev[j] = FaY;
			}
		}

		if (pass == 1) {
			// ---- x direction on row r (lanes)
			C U;
			if constexpr (DIM >= 2) U = W0;
			else loadRow(0, U);
			R Fx[Q];
			if constexpr (PLM) {
				C Um, Up, Lx, Rx, RxL;
				shift_left<R, E>(U, Um);
				shift_right<R, E>(U, Up);
				plm_cons<R, E, SLIM>(slim, Um, U, Up, Lx, Rx);
				shift_left<R, E>(Rx, RxL);
				Iface<R, E> I;
				iface_prep<R, E, 0>(par, RxL, Lx, I);
				iface_flux<R, E, 0, false, 0>(par, I, nullptr, nullptr, dt_dx0, lim, Fx);
			} else if constexpr (LIM) {
				C Um;
				shift_left<R, E>(U, Um);
				Iface<R, E> I;
				iface_prep<R, E, 0>(par, Um, U, I);
				R dWL[E::NW], dWR[E::NW];
#pragma unroll
				for (int j = 0; j < E::NW; ++j) { dWL[j] = from_left(I.dW[j]); dWR[j] = from_right(I.dW[j]); }
				iface_flux<R, E, 0, true, FLIM>(par, I, dWL, dWR, dt_dx0, lim, Fx);
			} else {
				C Um;
				shift_left<R, E>(U, Um);
				Iface<R, E> I;
				iface_prep<R, E, 0>(par, Um, U, I);
				iface_flux<R, E, 0, false, 0>(par, I, nullptr, nullptr, dt_dx0, lim, Fx);
			}
			R deriv[Q];
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				R Fa = g.fluxOn[0] ? Fx[j] * g.areaOverVol[0] : R(0);
				R FaR = from_right(Fa);
				deriv[j] = R(0) - (FaR - Fa);
				if constexpr (DIM >= 2) deriv[j] = deriv[j] - dyterm[j];
				if (!g.derivOn) deriv[j] = R(0);
			}
			if constexpr (E::SOURCE) {
				// eqn addSource after calcDerivFromFlux (solverbase.lua:3208-3214)
				R em[3], ep[3];
				em[0] = from_left(U.v[10]);
				ep[0] = from_right(U.v[10]);
				if constexpr (DIM >= 2) { em[1] = Wm.v[10]; ep[1] = W1.v[10]; }
				E::template source<DIM>(par, g.invdx, U.v, em, ep, deriv);
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
				E::constrain(a.gammaD, a.rhoMinD, a.PMinD, acc);
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
					double d = E::template dt_cell<DIM>(a.gammaD, a.rhoMinD, a.PMinD, g.dxD, ud, a.divPhiD, a.divPsiD);
					dtLocal = fmin(dtLocal, d);
				}
			}
		}

		// ---- advance the row window
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
// Plain form: one grid, one wave per strip.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(256) stage_kernel(const StageArgs<R> a)
{
	const int lane = threadIdx.x & 63;
	int blk = a.xcdRemap ? xcd_block(blockIdx.x, a.nBlocks) : (int)blockIdx.x;
	const int waveInBlock = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const long item = (long)blk * (blockDim.x >> 6) + waveInBlock;
	const GridDesc<R> &g = a.g;
	double dt = stage_dt(a, g);
	if (item >= g.nItems) return;     // whole waves leave; no barriers follow
	const int sx = (int)(item % g.stripsX);
	const int sy = (int)(item / g.stripsX);
	if (item == 0) stage_first_strip(a, g, dt, lane);
	double dtLocal = stage_strip<R, EQ, DIM, PLM, LIM, FLIM, SLIM>(a, g, sx, sy, dt, lane);
	stage_reduce(a, g, dtLocal, item, lane);
}

// Batched form: a persistent launch over every strip of every active grid of
// the batch; wave k takes items k, k + W, k + 2W, ... (W = waves launched).
// Items are numbered grid by grid (GridDesc::itemBase), so the owning grid of
// an item is found by a wave-uniform binary search.
template <typename R, template <typename> class EQ, int DIM, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(256) stage_kernel_batch(const StageArgs<R> a)
{
	const int lane = threadIdx.x & 63;
	int blk = a.xcdRemap ? xcd_block(blockIdx.x, a.nBlocks) : (int)blockIdx.x;
	const int waveInBlock = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const long wave = (long)blk * (blockDim.x >> 6) + waveInBlock;
	const long nWaves = (long)gridDim.x * (blockDim.x >> 6);
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
		double dtLocal = stage_strip<R, EQ, DIM, PLM, LIM, FLIM, SLIM>(a, g, sx, sy, dt, lane);
		stage_reduce(a, g, dtLocal, local, lane);
	}
}

} // namespace hydro
