// fv_host.h — host-side internals shared by the C ABI translation units:
// error reporting, Shu-Osher tables, the Grid record (one solver's device
// state), its layout conversion to and from the reference's ghosted cons_t
// buffers, and the argument blocks of the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hydro_hip.h"
#include "fv_common.h"
#include "fv_launch.h"
#include "fv_aux.h"

namespace hydro {

// ------------------------------------------------------------------ errors
std::string &hydro_err_buf();   // thread-local last error (hydro_hip.cpp)

static inline int fail(const char *fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	hydro_err_buf() = buf;
	return -1;
}

#define HIPCHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); } while (0)

static inline void ignore(hipError_t) {}

static inline int num_states(int eqn) {
	return eqn == HYDRO_EQN_EULER ? 6 : eqn == HYDRO_EQN_GLM_MAXWELL ? 12 : eqn == HYDRO_EQN_SHALLOW_WATER ? 5 : -1;
}
static inline int num_int_states(int eqn) {
	return eqn == HYDRO_EQN_EULER ? 5 : eqn == HYDRO_EQN_GLM_MAXWELL ? 8 : eqn == HYDRO_EQN_SHALLOW_WATER ? 4 : -1;
}
// static (non-integrated) planes on the device: GLM-Maxwell's rhoCharge,
// sigma, sqrt_1_eps, sqrt_1_mu; the shallow water depth (Euler's ePot stays on
// the host)
static inline int num_static_states(int eqn) {
	return eqn == HYDRO_EQN_GLM_MAXWELL ? 4 : eqn == HYDRO_EQN_SHALLOW_WATER ? 1 : 0;
}
// equations with an addSource step, whose strips load a halo instead of
// taking an edge pass (the source reads the x neighbours' static fields)
static inline bool eqn_has_source(int eqn) { return eqn == HYDRO_EQN_GLM_MAXWELL || eqn == HYDRO_EQN_SHALLOW_WATER; }

static inline EqnParD eqn_par_d(const hydro_cfg &c) {
	EqnParD p;
	p.gamma = c.gamma;
	p.rhoMin = c.rho_min;
	p.PMin = c.P_min;
	p.divPhi = c.div_phi_wavespeed;
	p.divPsi = c.div_psi_wavespeed;
	p.gravity = c.gravity;
	return p;
}

static inline void cfg_default(hydro_cfg *c) {
	std::memset(c, 0, sizeof *c);
	c->dim = 1;
	for (int i = 0; i < 3; ++i) {
		c->size[i] = 1;
		c->mins[i] = -1;
		c->maxs[i] = 1;
	}
	c->eqn = HYDRO_EQN_EULER;
	c->flux = HYDRO_FLUX_ROE;
	c->flux_limiter = HYDRO_LIMITER_DONOR_CELL;
	c->plm = HYDRO_PLM_NONE;
	c->slope_limiter = HYDRO_LIMITER_DONOR_CELL;
	c->integrator = HYDRO_INT_FORWARD_EULER;
	c->cfl = .5;
	c->roe_entropy_fix_factor = .2;
	c->gravity = 1;
	for (int i = 0; i < 6; ++i) c->boundary[i] = HYDRO_BC_FREEFLOW;
	c->precision = HYDRO_REAL_F64;
	c->gamma = 7. / 5.;
	c->rho_min = 1e-7;
	c->P_min = 1e-7;
	c->div_phi_wavespeed = 1;
	c->div_psi_wavespeed = 1;
}

// --------------------------------------------------------- Shu-Osher tables
// hydro/int/all.lua:10-160 (rows = stages, columns = k); forward Euler is a
// one-stage table here (fe.lua:33-49 computes U + dt L(U) as well).
struct RKTable {
	int order;
	double a[4][4], b[4][4];
};

static bool rk_table(int integ, RKTable &t) {
	std::memset(&t, 0, sizeof t);
	switch (integ) {
	case HYDRO_INT_FORWARD_EULER: t.order = 1; t.a[0][0] = 1; t.b[0][0] = 1; return true;
	case HYDRO_INT_RK2: t.order = 2;
		t.a[0][0] = 1; t.a[1][0] = 1; t.b[0][0] = .5; t.b[1][1] = 1; return true;
	case HYDRO_INT_RK2_HEUN: t.order = 2;
		t.a[0][0] = 1; t.a[1][0] = 1; t.b[0][0] = 1; t.b[1][0] = .5; t.b[1][1] = .5; return true;
	case HYDRO_INT_RK2_RALSTON: t.order = 2;
		t.a[0][0] = 1; t.a[1][0] = 1; t.b[0][0] = 2. / 3.; t.b[1][0] = 1. / 4.; t.b[1][1] = 3. / 4.; return true;
	case HYDRO_INT_RK3: t.order = 3;
		t.a[0][0] = t.a[1][0] = t.a[2][0] = 1;
		t.b[0][0] = .5; t.b[1][0] = -1; t.b[1][1] = 2;
		t.b[2][0] = 1. / 6.; t.b[2][1] = 2. / 6.; t.b[2][2] = 1. / 6.; return true;
	case HYDRO_INT_RK4: t.order = 4;
		for (int i = 0; i < 4; ++i) t.a[i][0] = 1;
		t.b[0][0] = .5; t.b[1][1] = .5; t.b[2][2] = 1;
		t.b[3][0] = 1. / 6.; t.b[3][1] = 2. / 6.; t.b[3][2] = 2. / 6.; t.b[3][3] = 1. / 6.; return true;
	case HYDRO_INT_RK4_3_8THS: t.order = 4;
		for (int i = 0; i < 4; ++i) t.a[i][0] = 1;
		t.b[0][0] = 1. / 3.; t.b[1][0] = -1. / 3.;
		t.b[2][0] = 1; t.b[2][1] = -1; t.b[2][2] = 1;
		t.b[3][0] = 1. / 8.; t.b[3][1] = 3. / 8.; t.b[3][2] = 3. / 8.; t.b[3][3] = 1. / 8.; return true;
	case HYDRO_INT_RK2_TVD: t.order = 2;
		t.a[0][0] = 1; t.a[1][0] = .5; t.a[1][1] = .5; t.b[0][0] = 1; t.b[1][1] = .5; return true;
	case HYDRO_INT_RK2_NON_TVD: t.order = 2;
		t.a[0][0] = 1; t.a[1][0] = 1; t.b[0][0] = -20; t.b[1][0] = 41. / 40.; t.b[1][1] = -1. / 40.; return true;
	case HYDRO_INT_RK3_TVD: t.order = 3;
		t.a[0][0] = 1; t.a[1][0] = 3. / 4.; t.a[1][1] = 1. / 4.; t.a[2][0] = 1. / 3.; t.a[2][2] = 2. / 3.;
		t.b[0][0] = 1; t.b[1][1] = 1. / 4.; t.b[2][2] = 2. / 3.; return true;
	case HYDRO_INT_RK4_TVD: t.order = 4;
		t.a[0][0] = 1;
		t.a[1][0] = 649. / 1600.; t.a[1][1] = 951. / 1600.;
		t.a[2][0] = 53989. / 2500000.; t.a[2][1] = 4806213. / 20000000.; t.a[2][2] = 23619. / 32000.;
		t.a[3][0] = 1. / 5.; t.a[3][1] = 6127. / 30000.; t.a[3][2] = 7873. / 30000.; t.a[3][3] = 1. / 3.;
		t.b[0][0] = .5;
		t.b[1][0] = -10890423. / 25193600.; t.b[1][1] = 5000. / 7873;
		t.b[2][0] = -102261. / 5000000.; t.b[2][1] = -5121. / 20000.; t.b[2][2] = 7873. / 10000.;
		t.b[3][0] = 1. / 10.; t.b[3][1] = 1. / 6.; t.b[3][3] = 1. / 6.; return true;
	case HYDRO_INT_RK4_NON_TVD: t.order = 4;
		for (int i = 0; i < 3; ++i) t.a[i][0] = 1;
		t.a[3][0] = -1. / 3.; t.a[3][1] = 1. / 3.; t.a[3][2] = 2. / 3.; t.a[3][3] = 1. / 3.;
		t.b[0][0] = .5; t.b[1][1] = .5; t.b[2][2] = 1; t.b[3][3] = 1. / 6.; return true;
	}
	return false;
}

static bool needs_deriv(const RKTable &t, int i) {
	for (int m = i + 1; m < t.order; ++m)
		if (t.b[m][i] != 0) return true;
	return false;
}

// ------------------------------------------------------------------ Grid
struct Grid {
	hydro_cfg cfg;
	int dim, Q, NSTAT, S;
	int n[3];
	long ncell;
	long ps;
	size_t elem;
	RKTable rk;
	void *U[4];
	void *D[4];
	void *ping;
	void *stat;
	DevScalars *dev;
	int slot;           // dt slot of the next step
	int dtValid;        // slot holds the min for the current state
	int rowsPerWave;
	int nvel;           // Euler: velocity components that are not identically zero
	std::vector<double> ePot;   // Euler ePot of the last upload (dynamics-neutral)
	void *redo;         // redo list of the stage kernels (see redo_kernels)
};

// Stage kernels (plain and batched) that leave strips with an irregular Roe
// interface to a redo kernel (fv_stage.h, ROE_FLAGGED, redo_form): FAST,
// Euler, Roe without the entropy fix and without a flux limiter, 2D
static inline bool redo_kernels(const hydro_cfg &c) {
	const bool lim = c.plm == HYDRO_PLM_NONE && c.flux_limiter != HYDRO_LIMITER_DONOR_CELL;
	return c.fast_math && c.eqn == HYDRO_EQN_EULER && c.flux == HYDRO_FLUX_ROE && !c.roe_entropy_fix && !lim
		&& c.dim == 2;
}

// Blocks of the redo kernel, whose waves stride over the redo lists.
static constexpr int REDO_BLOCKS = 64;

// capacity of the redo list: one entry per strip at one row per strip, the
// most strips any launch shape makes
static inline size_t redo_capacity(const Grid &g) {
	if (!redo_kernels(g.cfg)) return 0;
	const long so = strip_out(strip_edges(g.dim, eqn_has_source(g.cfg.eqn)));
	return (size_t)((g.n[0] + so - 1) / so) * (size_t)(g.dim >= 2 ? g.n[1] : 1);
}

static size_t elem_size(const hydro_cfg &c) { return c.precision == HYDRO_REAL_F32 ? 4 : 8; }

static int validate(const hydro_cfg *c) {
	if (!c) return fail("cfg is NULL");
	if (c->dim < 1 || c->dim > 3) return fail("dim %d not supported (1, 2 or 3)", c->dim);
	if (c->dim == 3 && c->eqn != HYDRO_EQN_EULER) return fail("3D is built for the Euler equations only");
	if (c->eqn == HYDRO_EQN_SHALLOW_WATER && !(c->gravity > 0)) return fail("gravity must be > 0");
	for (int i = 0; i < c->dim; ++i)
		if (c->size[i] < 1) return fail("size[%d] = %d must be >= 1", i, c->size[i]);
	if (c->eqn != HYDRO_EQN_EULER && c->eqn != HYDRO_EQN_GLM_MAXWELL && c->eqn != HYDRO_EQN_SHALLOW_WATER)
		return fail("eqn %d not supported", c->eqn);
	if (c->flux < HYDRO_FLUX_ROE || c->flux > HYDRO_FLUX_MARQUINA)
		return fail("flux %d not supported", c->flux);
	if (c->flux != HYDRO_FLUX_ROE && c->eqn != HYDRO_EQN_EULER)
		return fail("the HLL, Rusanov, HLLC and Marquina fluxes take the Euler equations");
	if (c->eqn == HYDRO_EQN_SHALLOW_WATER && c->plm > HYDRO_PLM_CONS)
		return fail("the shallow water equations take 'plm cons' or no reconstruction");
	if (c->roe_entropy_fix && (c->flux != HYDRO_FLUX_ROE || c->eqn != HYDRO_EQN_EULER))
		return fail("the entropy fix belongs to the Roe flux of the Euler equations (roe.cl:83-106)");
	if (c->roe_entropy_fix && !(c->roe_entropy_fix_factor > 0))
		return fail("roe_entropy_fix_factor must be > 0");
	if (c->flux_limiter < 0 || c->flux_limiter >= HYDRO_NUM_LIMITERS) return fail("flux_limiter %d out of range", c->flux_limiter);
	if (c->slope_limiter < 0 || c->slope_limiter >= HYDRO_NUM_LIMITERS) return fail("slope_limiter %d out of range", c->slope_limiter);
	if (c->plm < HYDRO_PLM_NONE || c->plm > HYDRO_PLM_EIG_PRIM_REF)
		return fail("plm %d not supported", c->plm);
	const bool glmPpm = c->plm == HYDRO_PLM_PPM && c->eqn == HYDRO_EQN_GLM_MAXWELL;
	if (c->plm > HYDRO_PLM_CONS && !glmPpm && (c->eqn != HYDRO_EQN_EULER || c->flux != HYDRO_FLUX_ROE || c->roe_entropy_fix))
		return fail("%s is built for the Euler equations with the Roe flux (no entropy fix)",
			c->plm == HYDRO_PLM_PRIM ? "plm prim" : c->plm == HYDRO_PLM_EIG ? "plm eig"
			: c->plm == HYDRO_PLM_CONS_FLUX ? "plm cons with flux" : c->plm == HYDRO_PLM_ATHENA ? "plm athena"
			: c->plm == HYDRO_PLM_PPM ? "ppm" : c->plm == HYDRO_PLM_EIG_PRIM ? "plm eig prim" : "plm eig prim ref");
	// gridsolver.lua:119
	if (c->plm && c->flux_limiter != HYDRO_LIMITER_DONOR_CELL)
		return fail("are you sure you want to use flux and slope limiters at the same time?");
	RKTable t;
	if (!rk_table(c->integrator, t)) return fail("integrator %d not supported", c->integrator);
	for (int f = 0; f < 2 * c->dim; ++f)
		if (c->boundary[f] < 0 || c->boundary[f] > HYDRO_BC_FIXED) return fail("boundary[%d] = %d invalid", f, c->boundary[f]);
	if (c->precision != HYDRO_REAL_F64 && c->precision != HYDRO_REAL_F32) return fail("precision %d invalid", c->precision);
	if (!(c->cfl > 0) && !c->use_fixed_dt) return fail("cfl must be > 0");
	return 0;
}

// Compute units of the current device (256 on MI355X; also when none is found).
static long device_cus() {
	int dev = 0, cus = 0;
	if (hipGetDevice(&dev) != hipSuccess ||
		hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
		return 256;
	return cus;
}

static StageKey stage_key_cfg(const hydro_cfg &cfg, int nvel);

// Blocks of the plain FP64 stage kernel of cfg that a CU holds at once (its
// registers and LDS decide: the heavier reconstructions and fluxes hold 2 or
// 3 where 'plm cons' with Roe holds 4); 4 if the query fails. The velocity
// components are taken as the dimension's, the usual Euler state.
static long plain_occupancy(const hydro_cfg &cfg) {
	StageArgs<double> a{};
	const StageKey k = stage_key_cfg(cfg, std::min(cfg.dim, 3));
	stage_occupancy = 0;
	if (launch_stage<double>(k, a, dim3(0), dim3(0), nullptr) != hipSuccess || stage_occupancy < 1) {
		(void)hipGetLastError();
		return 4;
	}
	return stage_occupancy;
}

// Rows per wave of the plain stage kernel (2D), from KH sweeps (rows 3..32 at
// 512^2 .. 4096^2, fast math). Each wave marches `rows` rows of a strip after a
// warm-up of 1-3 rows and, in the 64-column strips, its strip-edge pass (about
// 1.5 rows of work; about 0.5 when a block's four x-adjacent strips share one
// pass); halo strips (equations with a source) have no edge pass.
// FP64: blocks of 4 waves are dealt to the CUs, a CU runs L = ceil(blocks /
// CUs) of them, at most occ at once (the kernel's resident blocks per CU, from
// an occupancy query), and the launch takes about
// L * (rows + warm + edge) / eff(min(L, occ)), eff (fitted: .7, .77, .9, 1 for
// 1, 2, 3, 4+) being how well that many waves per SIMD hide latency; the rows
// in 4..8 with the least time are taken: for the kernels of 4 waves per SIMD 4
// at 512^2 and 1024^2 (one round), 8 at 2048^2 and 4096^2. A halo-strip kernel
// of at most 2 blocks per CU takes rows up to 16, so that its launch can end in
// fewer rounds (GLM-Maxwell: 10 rows at 1024^2, 12 at 2048^2, each the fastest
// of a sweep of 3..16). FP32 is latency-bound and gains from many short waves:
// rows = clamp(strips * ny / 4096, 4, 8) (4 at 512^2 .. 1024^2, 8 from 2048^2;
// longer waves leave a costly partial last round).
// 3D: planes per block chunk (at most 16). Blocks are small (4 waves) and many resident
// per CU, so the tail matters less than in 2D: keep at least 5000 blocks and,
// among those chunk sizes, the fewest plane iterations per column (chunks that
// divide the depth evenly, each with its 1-3 plane warm-up).
static int default_rows(const hydro_cfg &cfg, int nx, int ny, int nz) {
	if (cfg.dim < 2) return 1;
	const int so = strip_out(strip_edges(cfg.dim, eqn_has_source(cfg.eqn)));
	long stripsX = (nx + so - 1) / so;
	int warm = cfg.plm ? 2 : (cfg.flux_limiter ? 3 : 1);
	if (cfg.dim == 3) {
		long W = cfg.precision == HYDRO_REAL_F32 ? Block3<float>::W : Block3<double>::W;
		long perPlane = stripsX * ((ny + W - 1) / W);
		long best = 8, bestCost = -1;
		for (long r = 8; r <= EDGE_PLANES; ++r) {
			long chunks = (nz + r - 1) / r;
			if (r > 8 && perPlane * chunks < 5000) continue;
			long cost = chunks * (r + warm);
			if (bestCost < 0 || cost < bestCost) { best = r; bestCost = cost; }
		}
		return (int)best;
	}
	if (cfg.precision == HYDRO_REAL_F32)
		return (int)std::max(4L, std::min(8L, (stripsX * (long)ny + 4095) / 4096));
	static const double eff[5] = {0, .7, .77, .9, 1.};
	const long cus = device_cus();
	const long occ = plain_occupancy(cfg);
	const bool edges = strip_edges(cfg.dim, eqn_has_source(cfg.eqn));
	int best = 4;
	double bestT = -1;
	// at most 8 rows for 64-column strips and for kernels of 3 or more blocks
	// per CU: longer strips measured no faster at 2048^2 and 4096^2
	const int rMax = !edges && occ <= 2 ? 16 : 8;
	for (int r = 4; r <= rMax; ++r) {
		long blocks = (stripsX * ((ny + r - 1) / r) + 3) / 4;
		long L = (blocks + cus - 1) / cus;
		// the edge pass: a wave's own, or a quarter of its block's shared one
		const double edge = !edges ? 0. : stripsX % 4 == 0 && 5 * r <= WAVE ? .5 : 1.5;
		double t = (double)L * (r + warm + edge) / eff[std::min(std::min(L, occ), 4L)];
		if (bestT < 0 || t < bestT * (1 - 1e-9)) { best = r; bestT = t; }
	}
	return best;
}

static void grid_init(Grid &g, const hydro_cfg &cfg) {
	g.cfg = cfg;
	g.dim = cfg.dim;
	g.Q = num_int_states(cfg.eqn);
	g.S = num_states(cfg.eqn);
	g.NSTAT = num_static_states(cfg.eqn);
	g.ncell = 1;
	for (int i = 0; i < 3; ++i) {
		g.n[i] = i < g.dim ? cfg.size[i] : 1;
		g.ncell *= g.n[i];
	}
	g.ps = g.ncell;
	g.elem = elem_size(cfg);
	rk_table(cfg.integrator, g.rk);
	std::memset(g.U, 0, sizeof g.U);
	std::memset(g.D, 0, sizeof g.D);
	g.ping = g.stat = nullptr;
	g.dev = nullptr;
	g.slot = 0;
	g.dtValid = 0;
	g.nvel = 3;
	g.rowsPerWave = default_rows(cfg, g.n[0], g.n[1], g.n[2]);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// device bytes of a grid, in the order grid_bind lays them out
static size_t grid_bytes(const Grid &g) {
	size_t plane = (size_t)g.ps * g.elem, state = align256(plane * g.Q);
	size_t b = align256(sizeof(DevScalars));
	int o = g.rk.order;
	for (int i = 0; i < o; ++i) {
		b += state;
		if (needs_deriv(g.rk, i)) b += state;
	}
	if (o == 1) b += state;
	if (g.NSTAT) b += align256(plane * g.NSTAT);
	b += align256(redo_capacity(g) * sizeof(int2));
	return b;
}

static void grid_bind(Grid &g, char *base) {
	size_t plane = (size_t)g.ps * g.elem, state = align256(plane * g.Q);
	char *p = base;
	g.dev = (DevScalars *)p;
	p += align256(sizeof(DevScalars));
	int o = g.rk.order;
	for (int i = 0; i < o; ++i) {
		g.U[i] = p;
		p += state;
		if (needs_deriv(g.rk, i)) {
			g.D[i] = p;
			p += state;
		}
	}
	if (o == 1) {
		g.ping = p;
		p += state;
	}
	if (g.NSTAT) {
		g.stat = p;
		p += align256(plane * g.NSTAT);
	}
	g.redo = redo_capacity(g) ? p : nullptr;
}

// zero every state buffer, +inf every dt slot (asynchronous on s)
[[maybe_unused]] static int grid_clear(Grid &g, char *base, hipStream_t s) {
	HIPCHECK(hipMemsetAsync(base, 0, grid_bytes(g), s));
	DevScalars z;
	std::memset(&z, 0, sizeof z);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < DT_LANES; ++j) z.dtmin[i][j] = INFINITY;
	HIPCHECK(hipMemcpyAsync(g.dev, &z, sizeof z, hipMemcpyHostToDevice, s));
	return 0;
}

template <typename R>
static void grid_desc(const Grid &g, GridDesc<R> &d) {
	std::memset(&d, 0, sizeof d);
	for (int i = 0; i < 4; ++i) {
		d.U[i] = (R *)g.U[i];
		d.D[i] = (R *)g.D[i];
	}
	d.ping = (R *)g.ping;
	d.stat = (const R *)g.stat;
	d.redo = (int2 *)g.redo;
	for (int i = 0; i < 3; ++i) d.n[i] = g.n[i];
	d.ps = g.ps;
	d.step1 = g.n[0];
	const hydro_cfg &c = g.cfg;
	for (int f = 0; f < 6; ++f) {
		d.bc[f] = c.boundary[f];
		for (int j = 0; j < g.Q + g.NSTAT; ++j) d.fixedv[f][j] = R(c.boundary_fixed[f][j]);
	}
	// fvsolver.cl:37-102 in the stored precision: volume = 1 * dx0 * dx1...,
	// areaL = 1 * prod(dx_other), areaL *= 1 / volume
	R dxr[3];
	for (int i = 0; i < 3; ++i) {
		dxr[i] = R((c.maxs[i] - c.mins[i]) / (double)g.n[i]);
		d.dx[i] = dxr[i];
		d.invdx[i] = R(1) / dxr[i];
		d.inv2dx[i] = R(1) / (dxr[i] * R(2));
		d.dxD[i] = (c.maxs[i] - c.mins[i]) / (double)g.n[i];
	}
	R volume = R(1);
	for (int i = 0; i < g.dim; ++i) volume = volume * dxr[i];
	d.derivOn = volume > R(1e-7);
	R invVolume = R(1) / volume;
	for (int side = 0; side < 3; ++side) {
		R area = R(1);
		for (int i = 0; i < g.dim; ++i)
			if (i != side) area = area * dxr[i];
		R aL = area;
		if (aL <= R(1e-7)) aL = R(0);
		d.areaOverVol[side] = aL * invVolume;
		d.fluxOn[side] = area > R(1e-7);
		// both guards as one factor: the flux term is then exactly 0 where a
		// guard is off (for finite fluxes; the sign of a zero may differ)
		d.aovFast[side] = d.fluxOn[side] && d.derivOn ? d.areaOverVol[side] : R(0);
	}
	d.cfl = c.cfl;
	d.rowsPerWave = g.rowsPerWave;
	const int so = strip_out(strip_edges(g.dim, eqn_has_source(g.cfg.eqn)));
	d.stripsX = (g.n[0] + so - 1) / so;
	if (g.dim == 3) {
		d.stripsY = (g.n[1] + Block3<R>::W - 1) / Block3<R>::W;
		d.stripsZ = (g.n[2] + g.rowsPerWave - 1) / g.rowsPerWave;
	} else {
		d.stripsY = g.dim >= 2 ? (g.n[1] + g.rowsPerWave - 1) / g.rowsPerWave : 1;
		d.stripsZ = 1;
	}
	d.nItems = (long)d.stripsX * d.stripsY * d.stripsZ;
	d.active = 1;
	d.itemBase = 0;
	d.dev = g.dev;
	d.stepsLeft = 1 << 30;
}

// Uniform arguments of stage s (0-based) of the table.
template <typename R>
static void stage_args(const Grid &g, int s, StageArgs<R> &a) {
	std::memset(&a, 0, sizeof a);
	const RKTable &t = g.rk;
	int o = t.order;
	a.src = s;
	bool last = s == o - 1;
	a.out = !last ? s + 1 : (o == 1 ? -1 : 0);
	a.nAlpha = 0;
	for (int k = 0; k <= s; ++k) {
		if (t.a[s][k] != 0) {
			a.accU[a.nAlpha] = k;
			a.alpha[a.nAlpha] = t.a[s][k];
			a.nAlpha++;
		}
	}
	a.nStoredD = 0;
	for (int k = 0; k < s; ++k) {
		if (t.b[s][k] != 0) {
			a.accD[a.nStoredD] = k;
			a.beta[a.nStoredD] = t.b[s][k];
			a.nStoredD++;
		}
	}
	a.storeD = needs_deriv(t, s) ? s : -1;
	a.betaNew = t.b[s][s];
	const hydro_cfg &c = g.cfg;
	a.gamma = R(c.gamma);
	a.rhoMin = R(c.rho_min);
	a.PMin = R(c.P_min);
	a.divPhi = R(c.div_phi_wavespeed);
	a.divPsi = R(c.div_psi_wavespeed);
	a.gravity = R(c.gravity);
	a.pd = eqn_par_d(c);
	a.roeFixFactor = R(c.roe_entropy_fix_factor);
	a.limiter = c.plm ? c.slope_limiter : c.flux_limiter;
	a.resetSlot = -1;
	a.reduceSlot = -1;
	a.dtSlot = -1;
	a.advanceTime = 0;
	a.xcdRemap = 1;
	a.redoPass = 0;
	a.shareEdges = 0;
}

static StageKey stage_key_cfg(const hydro_cfg &cfg, int nvel) {
	StageKey k;
	k.eqn = cfg.eqn;
	k.dim = cfg.dim;
	k.plm = cfg.plm != HYDRO_PLM_NONE;
	k.flux = cfg.flux;
	k.roeFix = cfg.roe_entropy_fix != 0;
	k.plmKind = cfg.plm;
	// only the Roe flux uses a flux limiter (fvsolver.lua:63)
	k.lim = !k.plm && k.flux == HYDRO_FLUX_ROE && cfg.flux_limiter != HYDRO_LIMITER_DONOR_CELL;
	k.limiter = k.plm ? cfg.slope_limiter : cfg.flux_limiter;
	k.fast = cfg.fast_math != 0;
	k.nvel = nvel;
	return k;
}

static StageKey stage_key(const Grid &g) { return stage_key_cfg(g.cfg, g.nvel); }

// ------------------------------------------------------- layout conversion
// Ghost coordinate -> interior source, following the reference's x, y, z
// boundary passes (gridsolver.lua:1272-1314).
struct HostMap {
	int idx, mirror, fixedFace;
};
static HostMap host_map(int c, int n, int kmin, int kmax, int faceMin) {
	HostMap m{c, 0, -1};
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
	m.idx = std::max(0, std::min(n - 1, m.idx));
	return m;
}

static bool reflects(int eqn, int field, int axis) {
	if (eqn == HYDRO_EQN_EULER || eqn == HYDRO_EQN_SHALLOW_WATER) return field == 1 + axis;
	return field == axis || field == 3 + axis;
}

static size_t ghosted_cells(const Grid &g) {
	size_t c = 1;
	for (int i = 0; i < 3; ++i) c *= (size_t)(i < g.dim ? g.n[i] + 2 * NG : 1);
	return c;
}

// Runs fn(r0, r1) over [0, n) split into contiguous ranges on up to 8 host
// threads (one range, on the caller's thread, below `grain` items).
template <typename F>
static void parallel_ranges(long n, long grain, F fn) {
	const long hw = std::max(1L, (long)std::thread::hardware_concurrency());
	const long nt = std::max(1L, std::min({8L, hw, n / std::max(1L, grain)}));
	if (nt <= 1) {
		fn(0L, n);
		return;
	}
	std::vector<std::thread> th;
	th.reserve(nt - 1);
	for (long t = 1; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
	fn(0L, n / nt);
	for (auto &x : th) x.join();
}

// AoS ghosted cons_t -> the device layout (q planes and nstat static planes of
// ps values each, interior cells dense, the rest zero), into host memory the
// caller provides; sets g.nvel and g.ePot from the state. Rows are packed on
// several host threads.
template <typename R>
static void pack_interior(Grid &g, const R *aos, R *planes, R *stat) {
	int gs[3], o[3];
	for (int i = 0; i < 3; ++i) {
		gs[i] = i < g.dim ? g.n[i] + 2 * NG : 1;
		o[i] = i < g.dim ? NG : 0;
	}
	const bool euler = g.cfg.eqn == HYDRO_EQN_EULER;
	if (euler) g.ePot.assign(g.ncell, 0.);
	const long rows = (long)g.n[1] * g.n[2];
	const int nx = g.n[0], Q = g.Q, NS = g.NSTAT, S = g.S, dim = g.dim;
	const long ps = g.ps;
	double *ePot = euler ? g.ePot.data() : nullptr;
	// per-row flags (each row written by one thread), reduced afterwards
	std::vector<char> rowE(euler ? rows : 0, 0), rowV(euler ? rows : 0, 0);
	parallel_ranges(rows, 64, [&](long r0, long r1) {
		for (long row = r0; row < r1; ++row) {
			const int j = (int)(row % g.n[1]), k = (int)(row / g.n[1]);
			const R *u = aos + ((size_t)o[0] + (size_t)gs[0] * ((j + o[1]) + (size_t)gs[1] * (k + o[2]))) * S;
			const size_t c0 = (size_t)nx * (size_t)row;
			bool anyE = false, nzv = false;
			for (int i = 0; i < nx; ++i, u += S) {
				const size_t c = c0 + i;
				for (int f = 0; f < Q; ++f) planes[(size_t)f * ps + c] = u[f];
				for (int f = 0; f < NS; ++f) stat[(size_t)f * ps + c] = u[Q + f];
				if (euler) {
					ePot[c] = (double)u[5];
					anyE = anyE || u[5] != R(0);
					for (int f = 1 + dim; f <= 3; ++f) nzv = nzv || u[f] != R(0);
				}
			}
			if (euler) {
				rowE[row] = anyE;
				rowV[row] = nzv;
			}
		}
	});
	// each plane's cells past the interior (none while the planes are dense)
	for (int f = 0; f < Q; ++f) std::fill(planes + (size_t)f * ps + g.ncell, planes + (size_t)(f + 1) * ps, R(0));
	for (int f = 0; f < NS; ++f) std::fill(stat + (size_t)f * ps + g.ncell, stat + (size_t)(f + 1) * ps, R(0));
	bool anyE = false, nzv = false;
	for (long r = 0; r < (long)rowE.size(); ++r) {
		anyE = anyE || rowE[r];
		nzv = nzv || rowV[r];
	}
	if (!anyE) g.ePot.clear();
	// Euler: momentum components beyond dim that are zero everywhere stay zero
	// (their fluxes and wave amplitudes are exact zeros), so the kernels skip them
	g.nvel = euler && dim < 3 && !nzv ? dim : 3;
}

template <typename R>
static void unpack_with_ghosts(const Grid &g, const R *planes, const R *stat, R *aos) {
	int gs[3];
	for (int i = 0; i < 3; ++i) gs[i] = i < g.dim ? g.n[i] + 2 * NG : 1;
	const hydro_cfg &c = g.cfg;
	parallel_ranges((long)gs[1] * gs[2], 64, [&](long r0, long r1) {
		for (long row = r0; row < r1; ++row) {
			const int j = (int)(row % gs[1]), k = (int)(row / gs[1]);
			for (int i = 0; i < gs[0]; ++i) {
				int cc[3] = {i - (g.dim > 0 ? NG : 0), j - (g.dim > 1 ? NG : 0), k - (g.dim > 2 ? NG : 0)};
				HostMap m[3];
				for (int a = 0; a < 3; ++a) {
					if (a < g.dim) m[a] = host_map(cc[a], g.n[a], c.boundary[2 * a], c.boundary[2 * a + 1], 2 * a);
					else m[a] = HostMap{0, 0, -1};
				}
				size_t src = (size_t)m[0].idx + (size_t)g.n[0] * (m[1].idx + (size_t)g.n[1] * m[2].idx);
				R *u = aos + ((size_t)i + (size_t)gs[0] * (j + (size_t)gs[1] * k)) * g.S;
				for (int f = 0; f < g.S; ++f) {
					R v;
					if (f < g.Q) v = planes[(size_t)f * g.ps + src];
					else if (f - g.Q < g.NSTAT) v = stat[(size_t)(f - g.Q) * g.ps + src];
					else v = (g.cfg.eqn == HYDRO_EQN_EULER && !g.ePot.empty()) ? R(g.ePot[src]) : R(0);
					for (int a = 0; a < g.dim; ++a) {
						if (m[a].fixedFace >= 0) v = R(c.boundary_fixed[m[a].fixedFace][f]);
						else if (m[a].mirror && reflects(c.eqn, f, a)) v = -v;
					}
					u[f] = v;
				}
			}
		}
	});
}

template <typename R>
static int grid_upload_t(Grid &g, const void *U, hipStream_t s) {
	std::vector<R> planes((size_t)g.Q * g.ps), stat((size_t)g.NSTAT * g.ps);
	pack_interior<R>(g, (const R *)U, planes.data(), stat.data());
	HIPCHECK(hipMemcpyAsync(g.U[0], planes.data(), planes.size() * sizeof(R), hipMemcpyHostToDevice, s));
	if (g.NSTAT) HIPCHECK(hipMemcpyAsync(g.stat, stat.data(), stat.size() * sizeof(R), hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	g.dtValid = 0;
	return 0;
}

template <typename R>
static int grid_download_t(const Grid &g, void *U, hipStream_t s) {
	std::vector<R> planes((size_t)g.Q * g.ps), stat((size_t)g.NSTAT * g.ps);
	HIPCHECK(hipMemcpyAsync(planes.data(), g.U[0], planes.size() * sizeof(R), hipMemcpyDeviceToHost, s));
	if (g.NSTAT) HIPCHECK(hipMemcpyAsync(stat.data(), g.stat, stat.size() * sizeof(R), hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	unpack_with_ghosts<R>(g, planes.data(), stat.data(), (R *)U);
	return 0;
}

[[maybe_unused]] static int grid_upload(Grid &g, const void *U, size_t bytes, hipStream_t s) {
	if (!U) return fail("NULL argument");
	size_t need = ghosted_cells(g) * g.S * g.elem;
	if (bytes != need) return fail("upload size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	return g.elem == 8 ? grid_upload_t<double>(g, U, s) : grid_upload_t<float>(g, U, s);
}

[[maybe_unused]] static int grid_download(const Grid &g, void *U, size_t bytes, hipStream_t s) {
	if (!U) return fail("NULL argument");
	size_t need = ghosted_cells(g) * g.S * g.elem;
	if (bytes != need) return fail("download size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	return g.elem == 8 ? grid_download_t<double>(g, U, s) : grid_download_t<float>(g, U, s);
}

template <typename R>
static GridArgs<R> grid_args(const Grid &g, int slot) {
	GridArgs<R> a;
	std::memset(&a, 0, sizeof a);
	a.U = (R *)g.U[0];
	a.stat = (const R *)g.stat;
	a.ps = g.ps;
	a.ncell = g.ncell;
	a.dim = g.dim;
	for (int i = 0; i < 3; ++i) a.dx[i] = (g.cfg.maxs[i] - g.cfg.mins[i]) / (double)g.n[i];
	a.pd = eqn_par_d(g.cfg);
	a.dev = g.dev;
	a.slot = slot;
	return a;
}

static int grid_dt_kernel(Grid &g, int slot, hipStream_t s) {
	HIPCHECK(launch_reset_slots(g.dev, 1 << slot, s));
	hipError_t e = g.elem == 8 ? launch_dt<double>(g.cfg.eqn, g.dim, g.cfg.fast_math, grid_args<double>(g, slot), s)
	                           : launch_dt<float>(g.cfg.eqn, g.dim, g.cfg.fast_math, grid_args<float>(g, slot), s);
	if (e != hipSuccess) return fail("dt kernel: %s", hipGetErrorString(e));
	return 0;
}

[[maybe_unused]] static int grid_constrain(Grid &g, hipStream_t s) {
	hipError_t e = g.elem == 8 ? launch_constrain<double>(g.cfg.eqn, g.cfg.fast_math, grid_args<double>(g, 0), s)
	                           : launch_constrain<float>(g.cfg.eqn, g.cfg.fast_math, grid_args<float>(g, 0), s);
	if (e != hipSuccess) return fail("constrainU kernel: %s", hipGetErrorString(e));
	g.dtValid = 0;
	return 0;
}

[[maybe_unused]] static int grid_calc_dt(Grid &g, double *dt, hipStream_t s) {
	if (g.cfg.use_fixed_dt) { *dt = g.cfg.fixed_dt; return 0; }
	if (grid_dt_kernel(g, g.slot, s)) return -1;
	HIPCHECK(launch_fold_slot(g.dev, g.slot, s));
	double m;
	HIPCHECK(hipMemcpyAsync(&m, &g.dev->dtmin[g.slot][0], sizeof m, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	g.dtValid = 1;
	*dt = g.cfg.cfl * m;
	return 0;
}

[[maybe_unused]] static int grid_get_time(const Grid &g, double *t, double *dt, hipStream_t s) {
	double v[2];
	HIPCHECK(hipMemcpyAsync(v, &g.dev->t, sizeof v, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	if (t) *t = v[0];
	if (dt) *dt = v[1];
	return 0;
}

[[maybe_unused]] static int grid_set_time(Grid &g, double t, hipStream_t s) {
	HIPCHECK(hipMemcpyAsync(&g.dev->t, &t, sizeof t, hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	return 0;
}


// One stage of every grid described by descs (device array of nGrids
// GridDesc), as one persistent launch. rep supplies the shared configuration.
template <typename R>
static int launch_batched_stage(const Grid &rep, int nvel, int s, int cur, const GridDesc<R> *descs, int nGrids,
	long totalItems, int stepInTick, int wavesPerBlock, int xcdRemap, hipStream_t stream)
{
	int o = rep.rk.order;
	int fixed = rep.cfg.use_fixed_dt;
	bool last = s == o - 1;
	StageArgs<R> a;
	stage_args<R>(rep, s, a);
	if (!fixed) a.dtSlot = cur;
	else a.hostDt = rep.cfg.fixed_dt;
	a.resetSlot = (!fixed && s == 0) ? (cur + 2) % 3 : -1;
	a.reduceSlot = (!fixed && last) ? (cur + 1) % 3 : -1;
	a.advanceTime = last;
	a.stepInTick = stepInTick;
	a.grids = descs;
	a.nGrids = nGrids;
	a.totalItems = totalItems;
	a.xcdRemap = xcdRemap;
	StageKey k = stage_key(rep);
	k.nvel = nvel;
	hipError_t e = launch_stage<R>(k, a, dim3(0), dim3(64 * wavesPerBlock), stream);
	if (e != hipSuccess) return fail("batched stage launch: %s", hipGetErrorString(e));
	if (redo_kernels(rep.cfg)) {
		// the strips the stage kernel left to the general Roe form (usually none)
		a.redoPass = 1;
		a.nBlocks = REDO_BLOCKS;
		e = launch_stage<R>(k, a, dim3(REDO_BLOCKS), dim3(256), stream);
		if (e != hipSuccess) return fail("batched redo kernel launch: %s", hipGetErrorString(e));
	}
	return 0;
}

} // namespace hydro
