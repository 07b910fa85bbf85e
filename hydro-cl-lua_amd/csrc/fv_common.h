// fv_common.h — device-side building blocks shared by every stage kernel:
// the kernel argument block, boundary-mapped loads of the structure-of-arrays
// state, and 64-lane DPP neighbour exchange.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hydro_hip.h"

namespace hydro {

constexpr int NG = 2;              // numGhost (gridsolver.lua:41)
constexpr int WAVE = 64;           // CDNA wavefront
// Flux id of the kernels of the Roe flux with Harten's entropy fix
// (roe.cl:83-106); a template choice of its own, so that the kernels without
// the fix carry no test for it.
constexpr int FLUX_ROE_ENTROPY_FIX = 16;
constexpr int STRIP_OUT = WAVE;    // output columns per wave strip with edge passes
// Strips whose edge interfaces come from an edge pass output 64 columns; the
// others load a 2-cell halo in lanes 0-1 and 62-63 and output 60: 1D strips
// (a single row, too short to repay the pass) and equations with a source
// term, which reads the x neighbours' static fields (GLM-Maxwell: 12 state
// reals per cell make the pass's gathers cost more than the halo lanes,
// measured; shallow water: the depth).
__host__ __device__ constexpr int strip_out(bool edges) { return edges ? STRIP_OUT : WAVE - 2 * NG; }
__host__ __device__ constexpr bool strip_edges(int dim, bool source) { return dim >= 2 && !source; }
constexpr int EDGE_ROWS = WAVE / 2;// rows per wave item at most (2D): the strip-edge
                                   // interfaces of each row take a lane per side
constexpr int EDGE_PLANES = 16;    // planes per block chunk at most (3D)

// 3D stage kernel: a block is W row waves (rows of a strip that share one
// plane through LDS) and one halo wave that loads the rows outside the block
// and computes the y interfaces no row owns, so that the row waves do equal
// work between barriers. Four waves, so that several blocks fit per CU and
// one block's barrier waits are covered by the others' work. MIN_BLOCKS
// bounds the registers (FP32: three blocks per CU, at most 168 VGPRs; FP64
// takes what it needs: capped, its kernels spill).
template <typename R>
struct Block3 {
	static constexpr int W = 3;
	static constexpr int WAVES = W + 1;
	static constexpr int MIN_BLOCKS = sizeof(R) == 4 ? 3 : 1;
};

// Device-resident per-solver scalars: dt minima rotate through three slots so
// a step can read one, reduce into the next and clear the third.
constexpr int DT_LANES = 64;       // dt minima are spread over 64 addresses
struct DevScalars {
	double dtmin[3][DT_LANES];
	double t;
	double dt;
	unsigned redoCount;       // entries of the redo list of the running stage
	unsigned redoDone;        // blocks of the redo kernel that have finished
	unsigned nonFinite;       // set by nonfinite_kernel
	unsigned pad0;
	double pad[4];
};

// Boundary description of one axis, resolved on the host.
struct AxisBC {
	int n;            // interior cells
	int kind[2];      // HYDRO_BC_* at min / max
};

// Everything that is specific to one grid: its buffers, shape, boundaries,
// metric factors, launch decomposition and device scalars. The plain solver
// passes one by value; a batch keeps an array of them in device memory.
template <typename R>
struct GridDesc {
	R *U[4];                  // stage states U^(0..order-1); U[0] holds the solver state
	R *D[4];                  // stored stage derivatives (tables whose later rows use them)
	R *ping;                  // forward-Euler output (swapped with U[0] after a step)
	const R *stat;            // static (non-integrated) planes: GLM-Maxwell, shallow water
	int n[3];                 // interior cells
	long ps;                  // plane stride (reals)
	long step1;               // row stride (= n[0])
	int bc[6];                // HYDRO_BC_* per face
	R fixedv[6][HYDRO_MAX_STATES];   // HYDRO_BC_FIXED ghost values
	R areaOverVol[3];         // areaL * (1 / volume) per side (fvsolver.cl:64-102)
	int derivOn;              // volume > 1e-7 (fvsolver.cl:97)
	int fluxOn[3];            // cell_area > 1e-7 (fvsolver.lua:105)
	R aovFast[3];             // FAST kernels: areaOverVol where flux and derivative are on, else 0
	R dx[3];
	R invdx[3];               // 1 / dx in the stored precision (glm-maxwell.cl:351)
	R inv2dx[3];              // 1 / (2 dx) in the stored precision (makepartial.lua:105-108)
	double dxD[3];
	double cfl;
	int stripsX, stripsY, rowsPerWave;   // 3D: stripsY counts blocks of Block3::W rows, rowsPerWave planes per chunk
	int active;               // batch: grid takes part in the launch
	long itemBase;            // batch: first work item (wave strip) of this grid
	long nItems;              // stripsX * stripsY
	DevScalars *dev;          // dt slots and time of this grid
	int2 *redo;               // redo list: (strip item, first row) per strip, at most one each
	int stepsLeft;            // batch: steps still to run in the current tick
	int stripsZ;              // 3D: plane chunks of rowsPerWave planes
};

// Equation parameters in double: constrainU and the per-cell dt run in double
// in every stored precision.
struct EqnParD {
	double gamma, rhoMin, PMin;   // Euler (euler.lua:192-197)
	double divPhi, divPsi;        // GLM-Maxwell (glm-maxwell.lua:110-111)
	double gravity;               // shallow water (shallow-water.lua:167)
};

// Uniform for one launch: the stage's row of the Shu-Osher table (as buffer
// indices), the dt protocol and the equation parameters.
template <typename R>
struct StageArgs {
	int src;                  // buffer index of the stage whose derivative is taken
	int out;                  // output buffer index, -1: ping
	int accU[4];              // alpha terms, reference order
	double alpha[4];
	int nAlpha;
	int accD[3];              // stored derivatives of earlier stages (beta terms)
	double beta[3];
	int nStoredD;
	int storeD;               // store this stage's derivative into D[storeD] (-1: no)
	double betaNew;           // beta of this stage's derivative (0: not used)
	// dt: device slot (cfl * dtmin[slot]) or a host value
	int dtSlot;               // -1: use hostDt
	double hostDt;
	int resetSlot;            // slot to clear to +inf, -1 none
	int reduceSlot;           // slot to min-reduce the new state's dt into, -1 none
	int advanceTime;          // t += dt, dt stored
	int redoPass;             // this launch is the redo kernel of the stage
	int shareEdges;           // plain 2D launch: one strip-edge pass per block of 4 x-adjacent strips
	int stepInTick;           // batch: skip grids with stepsLeft <= stepInTick
	// equation parameters in the stencil type and in double (for the FP64
	// accumulation, constrainU and dt work of the FP32 form)
	R gamma, rhoMin, PMin, divPhi, divPsi, gravity;
	R roeFixFactor;           // Roe, Euler: Harten's entropy fix factor (roe.cl:83-106)
	EqnParD pd;
	int limiter;              // flux or slope limiter index (runtime switch form)
	int xcdRemap;
	int nBlocks;
	// one grid (plain launch) or a batch (persistent launch)
	GridDesc<R> g;
	const GridDesc<R> *grids;
	int nGrids;
	long totalItems;
};

// ---------------------------------------------------------------- lane moves
// Lane l receives lane l-1's value (wave_shr:1) / lane l+1's (wave_shl:1);
// the lane without a source (0 / 63) receives 0 (bound_ctrl: no old value to
// set up first). With zero fill, FP32 users of a shifted value take the shift
// as a DPP operand (no move).
__device__ __forceinline__ int dpp_from_left_i(int v) {
	return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_from_right_i(int v) {
	return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ double from_left(double v) {
	int2 p = __builtin_bit_cast(int2, v);
	p.x = dpp_from_left_i(p.x);
	p.y = dpp_from_left_i(p.y);
	return __builtin_bit_cast(double, p);
}
__device__ __forceinline__ double from_right(double v) {
	int2 p = __builtin_bit_cast(int2, v);
	p.x = dpp_from_right_i(p.x);
	p.y = dpp_from_right_i(p.y);
	return __builtin_bit_cast(double, p);
}
__device__ __forceinline__ float from_left(float v) {
	return __builtin_bit_cast(float, dpp_from_left_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float from_right(float v) {
	return __builtin_bit_cast(float, dpp_from_right_i(__builtin_bit_cast(int, v)));
}

// The same shifts with the empty lane (0 / 63) taking `edge` instead of 0:
// the DPP move's old operand, so no select is needed.
__device__ __forceinline__ double from_left_or(double v, double edge) {
	int2 p = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
	p.x = __builtin_amdgcn_update_dpp(e.x, p.x, 0x138, 0xF, 0xF, false);
	p.y = __builtin_amdgcn_update_dpp(e.y, p.y, 0x138, 0xF, 0xF, false);
	return __builtin_bit_cast(double, p);
}
__device__ __forceinline__ double from_right_or(double v, double edge) {
	int2 p = __builtin_bit_cast(int2, v), e = __builtin_bit_cast(int2, edge);
	p.x = __builtin_amdgcn_update_dpp(e.x, p.x, 0x130, 0xF, 0xF, false);
	p.y = __builtin_amdgcn_update_dpp(e.y, p.y, 0x130, 0xF, 0xF, false);
	return __builtin_bit_cast(double, p);
}
__device__ __forceinline__ float from_left_or(float v, float edge) {
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
		__builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float from_right_or(float v, float edge) {
	return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
		__builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// ------------------------------------------------------- boundary mapping
// A ghost coordinate maps to an interior source cell, a sign for mirror
// reflection, or a fixed value (gridsolver.lua:630-780). The reference fills
// ghosts axis by axis (x, then y, then z), so a later axis wins at corners:
// callers resolve z first, then y, then x.
struct Mapped {
	int idx;      // interior coordinate
	int mirror;   // 1 if the normal vector component flips
	int fixedFace;// face index when a fixed value applies, else -1
};

__device__ __forceinline__ Mapped map_coord(int c, int n, int kmin, int kmax, int faceMin) {
	Mapped m;
	m.idx = c;
	m.mirror = 0;
	m.fixedFace = -1;
	if (c < 0) {
		switch (kmin) {
		case HYDRO_BC_PERIODIC: m.idx = ((c % n) + n) % n; break;
		case HYDRO_BC_MIRROR: m.idx = -1 - c; m.mirror = 1; break;
		case HYDRO_BC_FREEFLOW: m.idx = 0; break;
		default: m.idx = 0; m.fixedFace = faceMin; break;
		}
	} else if (c >= n) {
		switch (kmax) {
		case HYDRO_BC_PERIODIC: m.idx = c % n; break;
		case HYDRO_BC_MIRROR: m.idx = 2 * n - 1 - c; m.mirror = 1; break;
		case HYDRO_BC_FREEFLOW: m.idx = n - 1; break;
		default: m.idx = n - 1; m.fixedFace = faceMin + 1; break;
		}
	}
	// guard tiny grids (n < ghost width): keep every access in bounds
	if (m.idx < 0) m.idx = 0;
	if (m.idx >= n) m.idx = n - 1;
	return m;
}

__device__ __forceinline__ double wave_min(double v) {
	for (int off = 32; off >= 1; off >>= 1) {
		double o = __shfl_xor(v, off, 64);
		v = o < v ? o : v;
	}
	return v;
}

// min over the 64 sub-slots of a dt slot: one coalesced load per wave
__device__ __forceinline__ double read_dt_slot(const DevScalars *d, int slot) {
	double v = d->dtmin[slot][threadIdx.x & 63];
	return wave_min(v);
}

__device__ __forceinline__ void atomic_min_pos(double *p, double v) {
	// positive doubles (and +inf) order like their bit patterns
	atomicMin(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v));
}

// Logical block id with blocks spread so that each XCD (blocks are dealt
// round-robin over the 8 XCDs) gets a contiguous range of logical ids, so
// neighbouring strips share an L2. A bijection for any block count.
__device__ __forceinline__ int xcd_block(int b, int nBlocks) {
	int q = nBlocks >> 3, r = nBlocks & 7;
	int x = b & 7;
	int start = x * q + (x < r ? x : r);
	return start + (b >> 3);
}

} // namespace hydro
