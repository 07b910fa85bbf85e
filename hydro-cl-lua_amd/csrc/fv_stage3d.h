// fv_stage3d.h — the fused finite-volume stage kernel in 3D (Euler).
//
// Same stage as the 1D/2D kernel (calcLR -> calcFlux -> calcDerivFromFlux ->
// multAdd... -> constrainU, and the new state's dt on the last stage;
// fvsolver.lua:225-302, rk.lua:152-183, calcDT.cl:38-73), with a third axis.
//
// Work decomposition: a block owns a 64-column strip (x) of W consecutive
// rows (y) and marches a chunk of at most 16 planes in z. It has W row waves,
// one per row, and one halo wave. In every row wave:
//   x neighbours: DPP lane shifts, and the strip's edge interfaces of all the
//     chunk's planes computed in one pass before the planes (as in 2D);
//   z neighbours: a sliding register window of planes (as y in 2D);
//   y neighbours: the block's rows of the current plane, exchanged through
//     LDS; wave w computes the interface below its row.
// The halo wave loads the rows outside the block (one plane ahead) and
// computes the block's y interfaces that no row owns: the top one, and the
// reconstruction edge (PLM) or characteristic jump (flux limiter) of the
// interface below the bottom row. So every row wave does the same work in
// each of the three barrier phases of a plane (with the z step in the phase
// where the halo wave computes its interfaces), and a plane takes as long as
// one row's work. Every wave of a block runs the same planes, so all of them
// reach every barrier.
#pragma once

#include "fv_stage.h"

namespace hydro {

template <typename R, class E>
__device__ __forceinline__ void load_cell3(const GridDesc<R> &g, const R *base, const Mapped &mx, const Mapped &my,
	const Mapped &mz, Cell<R, E> &c)
{
	long off = ((long)mz.idx * g.n[1] + my.idx) * g.step1 + mx.idx;
	int fixedFace = -1, lastFixed = -1;
	if (mx.fixedFace >= 0) { fixedFace = mx.fixedFace; lastFixed = 0; }
	if (my.fixedFace >= 0) { fixedFace = my.fixedFace; lastFixed = 1; }
	if (mz.fixedFace >= 0) { fixedFace = mz.fixedFace; lastFixed = 2; }
#pragma unroll
	for (int f = 0; f < Cell<R, E>::NV; ++f) {
		if (!E::active(f)) {
			c.v[f] = R(0);
			continue;
		}
		// axis by axis, a later axis wins (corners never enter the axis-split
		// stencil): the last fixed axis gives the value, and mirrors of the
		// axes after it flip it. One table lookup, from a face index chosen by
		// value (selecting between lookups would copy the argument block to
		// scratch memory).
		const int face = fixedFace;
		R v = base[f * g.ps + off];
		R fv = g.fixedv[face < 0 ? 0 : face][f];
		v = face >= 0 ? fv : v;
		if (lastFixed < 0 && mx.mirror && E::reflects(f, 0)) v = -v;
		if (lastFixed < 1 && my.mirror && E::reflects(f, 1)) v = -v;
		if (lastFixed < 2 && mz.mirror && E::reflects(f, 2)) v = -v;
		c.v[f] = v;
	}
}

// LDS of one block: the block's rows of the current plane plus two halo rows
// on each side, one slot per y interface quantity, and the y fluxes.
template <typename R, class E, int W>
struct Lds3 {
	static constexpr int NV = Cell<R, E>::NV;
	R u[W + 4][NV][WAVE];      // U of rows y0-2 .. y0+W+1
	R e[W + 2][NV][WAVE];      // PLM: R edges of rows y0-1 .. y0+W-1; flux limiter: dW at y0-3/2 .. y0+W-1/2
	R f[W + 1][E::Q][WAVE];    // area-weighted y fluxes at y0-1/2 .. y0+W-1/2
};

template <typename R, class E, int NV>
__device__ __forceinline__ void lds_put(R (&s)[NV][WAVE], const R *v, int lane) {
#pragma unroll
	for (int j = 0; j < NV; ++j) s[j][lane] = v[j];
}
template <typename R, class E, int NV>
__device__ __forceinline__ void lds_get(const R (&s)[NV][WAVE], R *v, int lane) {
#pragma unroll
	for (int j = 0; j < NV; ++j) v[j] = s[j][lane];
}

// the chunk's output planes z0 .. z1-1 of block item bz
template <typename R>
__device__ __forceinline__ void chunk3d(const GridDesc<R> &g, int bz, int &z0, int &z1) {
	z0 = bz * g.rowsPerWave;
	z1 = min(z0 + g.rowsPerWave, g.n[2]);
}

// Row wave w of the block: row y0 + w of the strip through the chunk.
template <typename R, template <typename> class EQ, bool PLM, bool LIM, int FLIM, int SLIM, int W>
__device__ __forceinline__ double stage_rows3d(const StageArgs<R> &a, const GridDesc<R> &g, int sx, int by, int bz,
	int w, double dt, int lane, Lds3<R, EQ<R>, W> &sh, XEdge<R, EQ<R>, PLM, LIM, EDGE_PLANES> &edge)
{
	using E = EQ<R>;
	using C = Cell<R, E>;
	constexpr int Q = E::Q;
	constexpr int NV = C::NV;
	static_assert(!E::SOURCE, "3D is built for the Euler equations");
	static_assert(E::NW == NV, "dW shares the LDS slots of the states");
	const typename E::Par par = E::par(a);
	const int lim = PLM ? 0 : a.limiter;
	const int slim = a.limiter;
	double dtLocal = INFINITY;

	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const R *src = g.U[a.src];
	R *out = a.out >= 0 ? g.U[a.out] : g.ping;
	R *storeD = a.storeD >= 0 ? g.D[a.storeD] : nullptr;
	const int x = sx * STRIP_OUT + lane;
	const int y0 = by * W;
	const int y = y0 + w;
	const bool outLane = x < nx && y < ny;
	const Mapped mx = map_coord(x, nx, g.bc[0], g.bc[1], 0);
	const Mapped myOwn = map_coord(y, ny, g.bc[2], g.bc[3], 2);
	const R rdt = R(dt);
	const R dt_dx0 = rdt / g.dx[0];
	const R dt_dx1 = rdt / g.dx[1];
	const R dt_dx2 = rdt / g.dx[2];
	int z0, z1;
	chunk3d(g, bz, z0, z1);

	// the wave's x coordinates and row need no fixed values or mirror flips
	const bool xyPlain = wave_all(mx.fixedFace < 0 && !mx.mirror) && myOwn.fixedFace < 0 && !myOwn.mirror;
	auto loadPlane = [&](int zz, C &c) {
		Mapped mz = map_coord(zz, nz, g.bc[4], g.bc[5], 4);
		if (xyPlain && mz.fixedFace < 0 && !mz.mirror)
			load_plain<R, E>(g, src, ((long)mz.idx * g.n[1] + myOwn.idx) * g.step1 + mx.idx, c);
		else load_cell3<R, E>(g, src, mx, myOwn, mz, c);
	};

	// ---- the strip's edge interfaces of row y for planes z0 .. z1-1: lane j
	// (< 16) the left edge of plane z0 + j, lane 16 + j the right edge
	{
		const int bj = lane & (EDGE_PLANES - 1);
		const bool right = (lane & EDGE_PLANES) != 0;
		const int zb = z0 + (bj < z1 - z0 ? bj : 0);
		const int xb = sx * STRIP_OUT + (right ? STRIP_OUT : 0);   // interface xb - 1/2
		const Mapped mz = map_coord(zb, nz, g.bc[4], g.bc[5], 4);
		Mapped mxb[4];
		bool plain = mz.fixedFace < 0 && !mz.mirror && myOwn.fixedFace < 0 && !myOwn.mirror;
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			mxb[k] = map_coord(xb - 2 + k, nx, g.bc[0], g.bc[1], 0);
			plain = plain && mxb[k].fixedFace < 0 && !mxb[k].mirror;
		}
		C cb[4];
		if (wave_all(plain)) {
#pragma unroll
			for (int k = 0; k < 4; ++k)
				load_plain<R, E>(g, src, ((long)mz.idx * g.n[1] + myOwn.idx) * g.step1 + mxb[k].idx, cb[k]);
		} else {
#pragma unroll
			for (int k = 0; k < 4; ++k) load_cell3<R, E>(g, src, mxb[k], myOwn, mz, cb[k]);
		}
		edge_pass<R, E, PLM, LIM, FLIM, SLIM, EDGE_PLANES>(par, lim, slim, cb, dt_dx0, flux_on<E>(g, 0),
			flux_aov<E>(g, 0), lane, edge);
	}

	// z pipeline state (the 2D kernel's y pipeline)
	C W0, W1, W2;
	C Rz0;
	Iface<R, E> I1;        // LIM: interface r+1/2
	R dW0[E::NW];          // LIM: dW of interface r-1/2
	R FaZprev[Q];
#pragma unroll
	for (int f = 0; f < NV; ++f) { W0.v[f] = W1.v[f] = W2.v[f] = Rz0.v[f] = R(0); }
#pragma unroll
	for (int j = 0; j < Q; ++j) FaZprev[j] = R(0);
#pragma unroll
	for (int j = 0; j < E::NW; ++j) dW0[j] = R(0);
	I1 = Iface<R, E>{};

	constexpr int WARM = PLM ? 2 : (LIM ? 3 : 1);
	if constexpr (PLM) { loadPlane(z0 - 2, W0); loadPlane(z0 - 1, W1); }
	else if constexpr (LIM) { loadPlane(z0 - 2, W1); }
	else { loadPlane(z0 - 1, W0); }

	// two passes: the warm-up planes (z pipeline only), then the output
	// planes, so that the output planes' body is one block with all the work
#pragma unroll
	for (int pass = 0; pass < 2; ++pass) {
	const int rBegin = pass == 0 ? z0 - WARM : z0;
	const int rEnd = pass == 0 ? z0 : z1;
	for (int r = rBegin; r < rEnd; ++r) {
		// plane r of this wave's row (the window's first plane) to the block
		if (pass == 1) {
			lds_put<R, E, NV>(sh.u[w + 2], W0.v, lane);
			__syncthreads();
		}

		// ---- z direction: interface r+1/2, dzterm of plane r
		R dzterm[Q];
		{
			R F[Q];
			// warm-up interfaces below z0-1/2 are never used
			const bool wantFlux = r >= z0 - 1;
			if constexpr (PLM) {
				loadPlane(r + 2, W2);
				C L1, R1;
				plm_cons<R, E, SLIM, 2>(par, slim, W0, W1, W2, L1, R1, dt_dx2);
				if (wantFlux) {
					iface_solve<R, E, 2>(par, Rz0, L1, dt_dx2, lim, F);
				} else {
#pragma unroll
					for (int j = 0; j < Q; ++j) F[j] = R(0);
				}
				Rz0 = R1;
			} else if constexpr (LIM) {
				loadPlane(r + 2, W2);
				Iface<R, E> I2;
				iface_prep<R, E, 2>(par, W1, W2, I2);
				if (wantFlux) iface_flux<R, E, 2, true, FLIM>(par, I1, dW0, I2.dW, dt_dx2, lim, F);
				else {
#pragma unroll
					for (int j = 0; j < Q; ++j) F[j] = R(0);
				}
#pragma unroll
				for (int j = 0; j < E::NW; ++j) dW0[j] = I1.dW[j];
				I1 = I2;
			} else {
				loadPlane(r + 1, W1);
				iface_solve<R, E, 2>(par, W0, W1, dt_dx2, lim, F);
			}
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				R FaZ = flux_on<E>(g, 2) ? F[j] * flux_aov<E>(g, 2) : R(0);
				dzterm[j] = FaZ - FaZprev[j];
				FaZprev[j] = FaZ;
			}
		}

		if (pass == 1) {
			const C &U = W0;     // plane r of this wave's row

			// ---- y direction through LDS: the interface below the row
			R Fy[Q];
			if constexpr (PLM) {
				C Um, Up, Ly, Ry;
				lds_get<R, E, NV>(sh.u[w + 1], Um.v, lane);
				lds_get<R, E, NV>(sh.u[w + 3], Up.v, lane);
				plm_cons<R, E, SLIM, 1>(par, slim, Um, U, Up, Ly, Ry, dt_dx1);
				lds_put<R, E, NV>(sh.e[w + 1], Ry.v, lane);
				__syncthreads();
				C Rm;
				lds_get<R, E, NV>(sh.e[w], Rm.v, lane);
				iface_solve<R, E, 1>(par, Rm, Ly, dt_dx1, lim, Fy);
			} else if constexpr (LIM) {
				// slot k of sh.e holds dW of interface y0 + k - 3/2
				C Um;
				lds_get<R, E, NV>(sh.u[w + 1], Um.v, lane);
				Iface<R, E> I;
				iface_prep<R, E, 1>(par, Um, U, I);
				lds_put<R, E, NV>(sh.e[w + 1], I.dW, lane);
				__syncthreads();
				R dWL[E::NW], dWR[E::NW];
				lds_get<R, E, NV>(sh.e[w], dWL, lane);
				lds_get<R, E, NV>(sh.e[w + 2], dWR, lane);
				iface_flux<R, E, 1, true, FLIM>(par, I, dWL, dWR, dt_dx1, lim, Fy);
			} else {
				C Um;
				lds_get<R, E, NV>(sh.u[w + 1], Um.v, lane);
				iface_solve<R, E, 1>(par, Um, U, dt_dx1, lim, Fy);
				__syncthreads();
			}
			R FaY[Q];
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				FaY[j] = flux_on<E>(g, 1) ? Fy[j] * flux_aov<E>(g, 1) : R(0);
				sh.f[w][j][lane] = FaY[j];
			}
			__syncthreads();
			R dyterm[Q];
#pragma unroll
			for (int j = 0; j < Q; ++j) dyterm[j] = sh.f[w + 1][j][lane] - FaY[j];

			// ---- x direction (lanes)
			C Um, Up;
			R deriv[Q];
			bool okx;   // unused: the 3D kernel takes the general Roe form where an interface is irregular (ROE_INLINE)
			strip_x<R, E, PLM, LIM, FLIM, SLIM, EDGE_PLANES>(par, lim, slim, U, dt_dx0, flux_on<E>(g, 0),
				flux_aov<E>(g, 0), lane, r - z0, EDGE_PLANES + (r - z0), edge, edge, Um, Up, deriv, okx);
#pragma unroll
			for (int j = 0; j < Q; ++j) {
				deriv[j] = deriv[j] - dyterm[j];
				deriv[j] = deriv[j] - dzterm[j];
				if (!E::FASTM && !g.derivOn) deriv[j] = R(0);
			}

			// ---- integrator combine, constrainU, new state's dt
			if (outLane) {
				long off = ((long)r * ny + y) * g.step1 + x;
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
					double ud[NV];
#pragma unroll
					for (int j = 0; j < Q; ++j) ud[j] = (double)o[j];
#pragma unroll
					for (int j = Q; j < NV; ++j) ud[j] = (double)U.v[j];
					double d = E::template dt_cell<3>(a.pd, g.dxD, ud);
					dtLocal = fmin(dtLocal, d);
				}
			}
		}

		// ---- advance the plane window
		if constexpr (PLM || LIM) { W0 = W1; W1 = W2; }
		else { W0 = W1; }
	}
	}
	return dtLocal;
}

// The halo wave of the block: per output plane, the rows y0-2, y0-1 and
// y0+W, y0+W+1 to LDS (loaded one plane ahead), then the y interface
// quantities no row wave owns: the R edge of row y0-1 (PLM) or dW of
// interface y0-3/2 (flux limiter) for row wave 0, and the flux of the top
// interface y0+W-1/2. The same barriers as the row waves.
template <typename R, template <typename> class EQ, bool PLM, bool LIM, int FLIM, int SLIM, int W>
__device__ __forceinline__ void stage_halo3d(const StageArgs<R> &a, const GridDesc<R> &g, int sx, int by, int bz,
	double dt, int lane, Lds3<R, EQ<R>, W> &sh)
{
	using E = EQ<R>;
	using C = Cell<R, E>;
	constexpr int Q = E::Q;
	constexpr int NV = C::NV;
	constexpr bool TWO = PLM || LIM;    // two halo rows on each side, else one
	const typename E::Par par = E::par(a);
	const int lim = PLM ? 0 : a.limiter;
	const int slim = a.limiter;
	const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
	const R *src = g.U[a.src];
	const int x = sx * STRIP_OUT + lane;
	const int y0 = by * W;
	const Mapped mx = map_coord(x, nx, g.bc[0], g.bc[1], 0);
	const R dt_dx1 = R(dt) / g.dx[1];
	int z0, z1;
	chunk3d(g, bz, z0, z1);

	// halo row k: rows y0-2, y0-1, y0+W, y0+W+1 (one halo row per side:
	// y0-1 and y0+W in slots 1 and 2)
	const int rowOf[4] = {y0 - 2, y0 - 1, y0 + W, y0 + W + 1};
	const int slotOf[4] = {0, 1, W + 2, W + 3};
	C h[4];
	auto loadHalo = [&](int zz) {
		const Mapped mz = map_coord(zz, nz, g.bc[4], g.bc[5], 4);
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			if (!TWO && (k == 0 || k == 3)) continue;
			const Mapped my = map_coord(rowOf[k], ny, g.bc[2], g.bc[3], 2);
			load_cell3<R, E>(g, src, mx, my, mz, h[k]);
		}
	};
	if (z0 < z1) loadHalo(z0);
	for (int r = z0; r < z1; ++r) {
#pragma unroll
		for (int k = 0; k < 4; ++k) {
			if (!TWO && (k == 0 || k == 3)) continue;
			lds_put<R, E, NV>(sh.u[slotOf[k]], h[k].v, lane);
		}
		if (r + 1 < z1) loadHalo(r + 1);
		__syncthreads();

		R F[Q];
		if constexpr (PLM) {
			C a0, a1, a2, Le, Re, Ltop;
			lds_get<R, E, NV>(sh.u[0], a0.v, lane);
			lds_get<R, E, NV>(sh.u[1], a1.v, lane);
			lds_get<R, E, NV>(sh.u[2], a2.v, lane);
			plm_cons<R, E, SLIM, 1>(par, slim, a0, a1, a2, Le, Re, dt_dx1);
			lds_put<R, E, NV>(sh.e[0], Re.v, lane);
			lds_get<R, E, NV>(sh.u[W + 1], a0.v, lane);
			lds_get<R, E, NV>(sh.u[W + 2], a1.v, lane);
			lds_get<R, E, NV>(sh.u[W + 3], a2.v, lane);
			plm_cons<R, E, SLIM, 1>(par, slim, a0, a1, a2, Ltop, Re, dt_dx1);
			__syncthreads();
			C Rtop;
			lds_get<R, E, NV>(sh.e[W], Rtop.v, lane);     // row y0+W-1's R edge
			iface_solve<R, E, 1>(par, Rtop, Ltop, dt_dx1, lim, F);
		} else if constexpr (LIM) {
			C a0, a1;
			Iface<R, E> Ie, It, Iu;
			lds_get<R, E, NV>(sh.u[0], a0.v, lane);
			lds_get<R, E, NV>(sh.u[1], a1.v, lane);
			iface_prep<R, E, 1>(par, a0, a1, Ie);              // interface y0-3/2
			lds_put<R, E, NV>(sh.e[0], Ie.dW, lane);
			lds_get<R, E, NV>(sh.u[W + 1], a0.v, lane);
			lds_get<R, E, NV>(sh.u[W + 2], a1.v, lane);
			iface_prep<R, E, 1>(par, a0, a1, It);              // interface y0+W-1/2
			lds_put<R, E, NV>(sh.e[W + 1], It.dW, lane);
			lds_get<R, E, NV>(sh.u[W + 3], a0.v, lane);
			iface_prep<R, E, 1>(par, a1, a0, Iu);              // interface y0+W+1/2
			__syncthreads();
			R dWL[E::NW];
			lds_get<R, E, NV>(sh.e[W], dWL, lane);            // interface y0+W-3/2
			iface_flux<R, E, 1, true, FLIM>(par, It, dWL, Iu.dW, dt_dx1, lim, F);
		} else {
			C a0, a1;
			lds_get<R, E, NV>(sh.u[W + 1], a0.v, lane);
			lds_get<R, E, NV>(sh.u[W + 2], a1.v, lane);
			iface_solve<R, E, 1>(par, a0, a1, dt_dx1, lim, F);
			__syncthreads();
		}
#pragma unroll
		for (int j = 0; j < Q; ++j) sh.f[W][j][lane] = flux_on<E>(g, 1) ? F[j] * flux_aov<E>(g, 1) : R(0);
		__syncthreads();
	}
}

// One block of Block3::WAVES waves per (strip, row block, plane chunk): the
// row waves 0 .. W-1 and the halo wave W; grid = items.
template <typename R, template <typename> class EQ, bool PLM, bool LIM, int FLIM, int SLIM>
__global__ void __launch_bounds__(WAVE * Block3<R>::WAVES, Block3<R>::MIN_BLOCKS) stage_kernel3d(const StageArgs<R> a)
{
	constexpr int W = Block3<R>::W;
	__shared__ Lds3<R, EQ<R>, W> sh;
	__shared__ XEdge<R, EQ<R>, PLM, LIM, EDGE_PLANES> edges[W];
	const int lane = threadIdx.x & 63;
	const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int blk = a.xcdRemap ? xcd_block(blockIdx.x, a.nBlocks) : (int)blockIdx.x;
	const GridDesc<R> &g = a.g;
	double dt = stage_dt(a, g);
	const int sx = blk % g.stripsX;
	const int rest = blk / g.stripsX;
	const int by = rest % g.stripsY;
	const int bz = rest / g.stripsY;
	if (blk == 0 && w == 0) stage_first_strip(a, g, dt, lane);
	if (w == W) {
		stage_halo3d<R, EQ, PLM, LIM, FLIM, SLIM, W>(a, g, sx, by, bz, dt, lane, sh);
		return;
	}
	double dtLocal = stage_rows3d<R, EQ, PLM, LIM, FLIM, SLIM, W>(a, g, sx, by, bz, w, dt, lane, sh, edges[w]);
	stage_reduce(a, g, dtLocal, (long)blk * W + w, lane);
}

} // namespace hydro
