// hydro_hip.cpp — the C ABI (include/hydro_hip.h) over the HIP kernels.
//
// A Grid is one solver's device state: the integrator's stage buffers
// (structure-of-arrays, interior cells only, one plane per integrated field),
// stored derivative buffers when the Shu-Osher table needs them, static planes
// (GLM-Maxwell's non-integrated fields) and a small block of device scalars
// (dt slots, t). Ghost cells exist only at the ABI boundary: uploads drop them,
// downloads rebuild them from the boundary conditions.
//
// hydro_solver owns one Grid and launches one wave per strip; hydro_batch owns
// many Grids carved out of one arena and advances them all with one persistent
// launch per stage, optionally replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hydro_hip.h"
#include "fv_common.h"
#include "fv_launch.h"
#include "fv_aux.h"

using namespace hydro;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

static int fail(const char *fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_err = buf;
	return -1;
}

#define HIPCHECK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) \
	return fail("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); } while (0)

static void ignore(hipError_t) {}

extern "C" const char *hydro_last_error(void) { return g_err.c_str(); }

extern "C" int hydro_eqn_num_states(int eqn) { return eqn == HYDRO_EQN_EULER ? 6 : eqn == HYDRO_EQN_GLM_MAXWELL ? 12 : -1; }
extern "C" int hydro_eqn_num_int_states(int eqn) { return eqn == HYDRO_EQN_EULER ? 5 : eqn == HYDRO_EQN_GLM_MAXWELL ? 8 : -1; }

extern "C" int hydro_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

extern "C" void hydro_cfg_default(hydro_cfg *c) {
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
};

static size_t elem_size(const hydro_cfg &c) { return c.precision == HYDRO_REAL_F32 ? 4 : 8; }

static int validate(const hydro_cfg *c) {
	if (!c) return fail("cfg is NULL");
	if (c->dim < 1 || c->dim > 2) return fail("dim %d not supported (1 or 2)", c->dim);
	for (int i = 0; i < c->dim; ++i)
		if (c->size[i] < 1) return fail("size[%d] = %d must be >= 1", i, c->size[i]);
	if (c->eqn != HYDRO_EQN_EULER && c->eqn != HYDRO_EQN_GLM_MAXWELL) return fail("eqn %d not supported", c->eqn);
	if (c->flux != HYDRO_FLUX_ROE) return fail("flux %d not supported", c->flux);
	if (c->flux_limiter < 0 || c->flux_limiter >= HYDRO_NUM_LIMITERS) return fail("flux_limiter %d out of range", c->flux_limiter);
	if (c->slope_limiter < 0 || c->slope_limiter >= HYDRO_NUM_LIMITERS) return fail("slope_limiter %d out of range", c->slope_limiter);
	if (c->plm != HYDRO_PLM_NONE && c->plm != HYDRO_PLM_CONS) return fail("plm %d not supported", c->plm);
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

static int default_rows(int dim, int nx, int ny, long targetWaves) {
	// enough waves to fill the device without letting the 2-3 row warm-up of
	// the y pipeline dominate
	if (dim < 2) return 1;
	long stripsX = (nx + STRIP_OUT - 1) / STRIP_OUT;
	long rows = (stripsX * (long)ny + targetWaves - 1) / targetWaves;
	return (int)std::max(4L, std::min(64L, rows));
}

static void grid_init(Grid &g, const hydro_cfg &cfg) {
	g.cfg = cfg;
	g.dim = cfg.dim;
	g.Q = hydro_eqn_num_int_states(cfg.eqn);
	g.S = hydro_eqn_num_states(cfg.eqn);
	g.NSTAT = cfg.eqn == HYDRO_EQN_GLM_MAXWELL ? 4 : 0;
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
	g.rowsPerWave = default_rows(g.dim, g.n[0], g.n[1], 4096);
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
	if (g.NSTAT) g.stat = p;
}

// zero every state buffer, +inf every dt slot (asynchronous on s)
static int grid_clear(Grid &g, char *base, hipStream_t s) {
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
	}
	d.cfl = c.cfl;
	d.rowsPerWave = g.rowsPerWave;
	d.stripsX = (g.n[0] + STRIP_OUT - 1) / STRIP_OUT;
	d.stripsY = g.dim >= 2 ? (g.n[1] + g.rowsPerWave - 1) / g.rowsPerWave : 1;
	d.nItems = (long)d.stripsX * d.stripsY;
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
	a.gammaD = c.gamma;
	a.rhoMinD = c.rho_min;
	a.PMinD = c.P_min;
	a.divPhiD = c.div_phi_wavespeed;
	a.divPsiD = c.div_psi_wavespeed;
	a.limiter = c.plm ? c.slope_limiter : c.flux_limiter;
	a.resetSlot = -1;
	a.reduceSlot = -1;
	a.dtSlot = -1;
	a.advanceTime = 0;
	a.xcdRemap = 1;
}

static StageKey stage_key(const Grid &g) {
	StageKey k;
	k.eqn = g.cfg.eqn;
	k.dim = g.dim;
	k.plm = g.cfg.plm != HYDRO_PLM_NONE;
	k.lim = !k.plm && g.cfg.flux_limiter != HYDRO_LIMITER_DONOR_CELL;
	k.limiter = k.plm ? g.cfg.slope_limiter : g.cfg.flux_limiter;
	k.fast = g.cfg.fast_math != 0;
	k.nvel = g.nvel;
	return k;
}

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
	if (eqn == HYDRO_EQN_EULER) return field == 1 + axis;
	return field == axis || field == 3 + axis;
}

static size_t ghosted_cells(const Grid &g) {
	size_t c = 1;
	for (int i = 0; i < 3; ++i) c *= (size_t)(i < g.dim ? g.n[i] + 2 * NG : 1);
	return c;
}

template <typename R>
static void pack_interior(Grid &g, const R *aos, std::vector<R> &planes, std::vector<R> &stat) {
	int gs[3], o[3];
	for (int i = 0; i < 3; ++i) {
		gs[i] = i < g.dim ? g.n[i] + 2 * NG : 1;
		o[i] = i < g.dim ? NG : 0;
	}
	planes.assign((size_t)g.Q * g.ps, R(0));
	stat.assign((size_t)g.NSTAT * g.ps, R(0));
	bool anyE = false;
	if (g.cfg.eqn == HYDRO_EQN_EULER) g.ePot.assign(g.ncell, 0.);
	for (int k = 0; k < g.n[2]; ++k)
		for (int j = 0; j < g.n[1]; ++j)
			for (int i = 0; i < g.n[0]; ++i) {
				size_t gi = (size_t)(i + o[0]) + (size_t)gs[0] * ((j + o[1]) + (size_t)gs[1] * (k + o[2]));
				size_t c = (size_t)i + (size_t)g.n[0] * (j + (size_t)g.n[1] * k);
				const R *u = aos + gi * g.S;
				for (int f = 0; f < g.Q; ++f) planes[(size_t)f * g.ps + c] = u[f];
				for (int f = 0; f < g.NSTAT; ++f) stat[(size_t)f * g.ps + c] = u[g.Q + f];
				if (g.cfg.eqn == HYDRO_EQN_EULER) {
					g.ePot[c] = (double)u[5];
					anyE = anyE || u[5] != R(0);
				}
			}
	if (!anyE) g.ePot.clear();
	// Euler: momentum components beyond dim that are zero everywhere stay zero
	// (their fluxes and wave amplitudes are exact zeros), so the kernels skip them
	g.nvel = 3;
	if (g.cfg.eqn == HYDRO_EQN_EULER && g.dim < 3) {
		bool zero = true;
		for (int f = 1 + g.dim; f <= 3 && zero; ++f)
			for (long c = 0; c < g.ps; ++c)
				if (planes[(size_t)f * g.ps + c] != R(0)) { zero = false; break; }
		if (zero) g.nvel = g.dim;
	}
}

template <typename R>
static void unpack_with_ghosts(const Grid &g, const std::vector<R> &planes, const std::vector<R> &stat, R *aos) {
	int gs[3];
	for (int i = 0; i < 3; ++i) gs[i] = i < g.dim ? g.n[i] + 2 * NG : 1;
	const hydro_cfg &c = g.cfg;
	for (int k = 0; k < gs[2]; ++k)
		for (int j = 0; j < gs[1]; ++j)
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

template <typename R>
static int grid_upload_t(Grid &g, const void *U, hipStream_t s) {
	std::vector<R> planes, stat;
	pack_interior<R>(g, (const R *)U, planes, stat);
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
	unpack_with_ghosts<R>(g, planes, stat, (R *)U);
	return 0;
}

static int grid_upload(Grid &g, const void *U, size_t bytes, hipStream_t s) {
	if (!U) return fail("NULL argument");
	size_t need = ghosted_cells(g) * g.S * g.elem;
	if (bytes != need) return fail("upload size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	return g.elem == 8 ? grid_upload_t<double>(g, U, s) : grid_upload_t<float>(g, U, s);
}

static int grid_download(const Grid &g, void *U, size_t bytes, hipStream_t s) {
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
	a.gamma = g.cfg.gamma;
	a.rhoMin = g.cfg.rho_min;
	a.PMin = g.cfg.P_min;
	a.divPhi = g.cfg.div_phi_wavespeed;
	a.divPsi = g.cfg.div_psi_wavespeed;
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

static int grid_constrain(Grid &g, hipStream_t s) {
	hipError_t e = g.elem == 8 ? launch_constrain<double>(g.cfg.eqn, g.cfg.fast_math, grid_args<double>(g, 0), s)
	                           : launch_constrain<float>(g.cfg.eqn, g.cfg.fast_math, grid_args<float>(g, 0), s);
	if (e != hipSuccess) return fail("constrainU kernel: %s", hipGetErrorString(e));
	g.dtValid = 0;
	return 0;
}

static int grid_calc_dt(Grid &g, double *dt, hipStream_t s) {
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

static int grid_get_time(const Grid &g, double *t, double *dt, hipStream_t s) {
	double v[2];
	HIPCHECK(hipMemcpyAsync(v, &g.dev->t, sizeof v, hipMemcpyDeviceToHost, s));
	HIPCHECK(hipStreamSynchronize(s));
	if (t) *t = v[0];
	if (dt) *dt = v[1];
	return 0;
}

static int grid_set_time(Grid &g, double t, hipStream_t s) {
	HIPCHECK(hipMemcpyAsync(&g.dev->t, &t, sizeof t, hipMemcpyHostToDevice, s));
	HIPCHECK(hipStreamSynchronize(s));
	return 0;
}

// ------------------------------------------------------------------ solver
struct hydro_solver {
	Grid g;
	void *mem;
	KSum *partial;
	double *totalsDev;
	int nPartial;
	hipStream_t stream;
	int xcdRemap;
	size_t devBytes;
};

extern "C" int hydro_create(const hydro_cfg *cfg, hydro_solver **out) {
	if (!out) return fail("out is NULL");
	*out = nullptr;
	if (validate(cfg)) return -1;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
	hydro_solver *h = new hydro_solver();
	grid_init(h->g, *cfg);
	h->xcdRemap = 1;
	auto cleanup = [&](int rc) { hydro_destroy(h); return rc; };
	if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
		delete h;
		return fail("hipStreamCreate failed");
	}
	size_t bytes = grid_bytes(h->g);
	if (hipMalloc(&h->mem, bytes) != hipSuccess) return cleanup(fail("hipMalloc %zu B of solver state", bytes));
	h->devBytes = bytes;
	grid_bind(h->g, (char *)h->mem);
	h->nPartial = 1024;
	if (hipMalloc((void **)&h->partial, sizeof(KSum) * h->nPartial * HYDRO_MAX_STATES) != hipSuccess) return cleanup(fail("hipMalloc partials"));
	if (hipMalloc((void **)&h->totalsDev, sizeof(double) * HYDRO_MAX_STATES) != hipSuccess) return cleanup(fail("hipMalloc totals"));
	if (grid_clear(h->g, (char *)h->mem, h->stream)) return cleanup(-1);
	if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail("sync"));
	*out = h;
	return 0;
}

extern "C" int hydro_destroy(hydro_solver *h) {
	if (!h) return 0;
	if (h->stream) ignore(hipStreamSynchronize(h->stream));
	if (h->mem) ignore(hipFree(h->mem));
	if (h->partial) ignore(hipFree(h->partial));
	if (h->totalsDev) ignore(hipFree(h->totalsDev));
	if (h->stream) ignore(hipStreamDestroy(h->stream));
	delete h;
	return 0;
}

extern "C" size_t hydro_num_cells(const hydro_solver *h) { return h ? ghosted_cells(h->g) : 0; }
extern "C" size_t hydro_device_bytes(const hydro_solver *h) { return h ? h->devBytes : 0; }

extern "C" int hydro_set_launch(hydro_solver *h, int rows_per_wave, int xcd_remap) {
	if (!h) return fail("NULL handle");
	if (rows_per_wave < 0 || rows_per_wave > 4096) return fail("rows_per_wave %d out of range", rows_per_wave);
	if (rows_per_wave > 0 && h->g.dim >= 2) h->g.rowsPerWave = rows_per_wave;
	if (xcd_remap > 0) h->xcdRemap = xcd_remap == 1;
	return 0;
}

extern "C" int hydro_get_launch(const hydro_solver *h, int *rows_per_wave, int *xcd_remap) {
	if (!h) return fail("NULL handle");
	if (rows_per_wave) *rows_per_wave = h->g.rowsPerWave;
	if (xcd_remap) *xcd_remap = h->xcdRemap ? 1 : 2;
	return 0;
}

extern "C" int hydro_upload_U(hydro_solver *h, const void *U, size_t bytes) {
	if (!h) return fail("NULL handle");
	return grid_upload(h->g, U, bytes, h->stream);
}

extern "C" int hydro_download_U(hydro_solver *h, void *U, size_t bytes) {
	if (!h) return fail("NULL handle");
	return grid_download(h->g, U, bytes, h->stream);
}

extern "C" int hydro_constrain_U(hydro_solver *h) {
	if (!h) return fail("NULL handle");
	return grid_constrain(h->g, h->stream);
}

extern "C" int hydro_boundary(hydro_solver *h) { return h ? 0 : fail("NULL handle"); }

extern "C" int hydro_calc_dt(hydro_solver *h, double *dt) {
	if (!h || !dt) return fail("NULL argument");
	return grid_calc_dt(h->g, dt, h->stream);
}

// Launch stage s of the plain solver.
template <typename R>
static int solver_stage(hydro_solver *h, int s, int useSlot, double hostDt, int resetSlot, int reduceSlot,
	int advance, hipEvent_t ev0, hipEvent_t ev1)
{
	Grid &g = h->g;
	StageArgs<R> a;
	stage_args<R>(g, s, a);
	grid_desc<R>(g, a.g);
	if (useSlot) a.dtSlot = g.slot;
	else a.hostDt = hostDt;
	a.resetSlot = resetSlot;
	a.reduceSlot = reduceSlot;
	a.advanceTime = advance;
	a.xcdRemap = h->xcdRemap;
	a.nBlocks = (int)((a.g.nItems + 3) / 4);
	if (ev0) HIPCHECK(hipEventRecord(ev0, h->stream));
	hipError_t e = launch_stage<R>(stage_key(g), a, dim3(a.nBlocks), dim3(256), h->stream);
	if (e != hipSuccess) return fail("stage kernel launch: %s", hipGetErrorString(e));
	if (ev1) HIPCHECK(hipEventRecord(ev1, h->stream));
	return 0;
}

// One integrator step. useSlot: dt = cfl * dtmin[slot] on the device.
template <typename R>
static int solver_step(hydro_solver *h, int useSlot, double hostDt, int update, std::vector<hipEvent_t> *evs) {
	Grid &g = h->g;
	int o = g.rk.order;
	int cur = g.slot;
	int fixed = g.cfg.use_fixed_dt;
	for (int s = 0; s < o; ++s) {
		bool last = s == o - 1;
		int resetSlot = (update && !fixed && s == 0) ? (cur + 2) % 3 : -1;
		int reduceSlot = (update && !fixed && last) ? (cur + 1) % 3 : -1;
		hipEvent_t e0 = nullptr, e1 = nullptr;
		if (evs) {
			HIPCHECK(hipEventCreate(&e0));
			HIPCHECK(hipEventCreate(&e1));
			evs->push_back(e0);
			evs->push_back(e1);
		}
		if (solver_stage<R>(h, s, useSlot, hostDt, resetSlot, reduceSlot, update && last, e0, e1)) return -1;
	}
	if (o == 1) std::swap(g.U[0], g.ping);
	if (update && !fixed) {
		g.slot = (cur + 1) % 3;
		g.dtValid = 1;
	}
	return 0;
}

extern "C" int hydro_step(hydro_solver *h, double dt) {
	if (!h) return fail("NULL handle");
	int rc = h->g.elem == 8 ? solver_step<double>(h, 0, dt, 0, nullptr) : solver_step<float>(h, 0, dt, 0, nullptr);
	h->g.dtValid = 0;
	return rc;
}

static int update_impl(hydro_solver *h, int nsteps, std::vector<hipEvent_t> *evs) {
	if (nsteps < 0) return fail("nsteps < 0");
	Grid &g = h->g;
	int fixed = g.cfg.use_fixed_dt;
	if (!fixed && !g.dtValid) {
		HIPCHECK(launch_reset_slots(g.dev, 7, h->stream));
		if (grid_dt_kernel(g, g.slot, h->stream)) return -1;
		g.dtValid = 1;
	}
	for (int i = 0; i < nsteps; ++i) {
		int rc = g.elem == 8 ? solver_step<double>(h, !fixed, g.cfg.fixed_dt, 1, evs)
		                     : solver_step<float>(h, !fixed, g.cfg.fixed_dt, 1, evs);
		if (rc) return rc;
	}
	return 0;
}

extern "C" int hydro_update(hydro_solver *h, int nsteps) {
	if (!h) return fail("NULL handle");
	return update_impl(h, nsteps, nullptr);
}

extern "C" int hydro_get_time(hydro_solver *h, double *t, double *dt) {
	if (!h) return fail("NULL handle");
	return grid_get_time(h->g, t, dt, h->stream);
}

extern "C" int hydro_set_time(hydro_solver *h, double t) {
	if (!h) return fail("NULL handle");
	return grid_set_time(h->g, t, h->stream);
}

extern "C" int hydro_conserved_totals(hydro_solver *h, double *out) {
	if (!h || !out) return fail("NULL argument");
	Grid &g = h->g;
	hipError_t e = g.elem == 8
		? launch_totals<double>((const double *)g.U[0], g.ps, g.ncell, g.Q, h->partial, h->nPartial, h->totalsDev, h->stream)
		: launch_totals<float>((const float *)g.U[0], g.ps, g.ncell, g.Q, h->partial, h->nPartial, h->totalsDev, h->stream);
	if (e != hipSuccess) return fail("totals kernel: %s", hipGetErrorString(e));
	HIPCHECK(hipMemcpyAsync(out, h->totalsDev, sizeof(double) * g.Q, hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	return 0;
}

extern "C" int hydro_sync(hydro_solver *h) {
	if (!h) return fail("NULL handle");
	HIPCHECK(hipStreamSynchronize(h->stream));
	return 0;
}

extern "C" int hydro_time_updates(hydro_solver *h, int nsteps, int per_kernel_events, hydro_timing *out) {
	if (!h || !out) return fail("NULL argument");
	std::memset(out, 0, sizeof *out);
	Grid &g = h->g;
	hipEvent_t t0, t1;
	HIPCHECK(hipEventCreate(&t0));
	HIPCHECK(hipEventCreate(&t1));
	std::vector<hipEvent_t> evs;
	if (!g.cfg.use_fixed_dt && !g.dtValid) {
		hipEvent_t d0, d1;
		HIPCHECK(hipEventCreate(&d0));
		HIPCHECK(hipEventCreate(&d1));
		HIPCHECK(hipEventRecord(d0, h->stream));
		HIPCHECK(launch_reset_slots(g.dev, 7, h->stream));
		if (grid_dt_kernel(g, g.slot, h->stream)) return -1;
		HIPCHECK(hipEventRecord(d1, h->stream));
		g.dtValid = 1;
		HIPCHECK(hipEventSynchronize(d1));
		float ms = 0;
		HIPCHECK(hipEventElapsedTime(&ms, d0, d1));
		out->dt_ms = ms;
		out->n_dt_launches = 1;
		ignore(hipEventDestroy(d0));
		ignore(hipEventDestroy(d1));
	}
	HIPCHECK(hipEventRecord(t0, h->stream));
	int rc = update_impl(h, nsteps, per_kernel_events ? &evs : nullptr);
	if (rc) return rc;
	HIPCHECK(hipEventRecord(t1, h->stream));
	HIPCHECK(hipEventSynchronize(t1));
	float ms = 0;
	HIPCHECK(hipEventElapsedTime(&ms, t0, t1));
	out->total_ms = ms;
	for (size_t i = 0; i + 1 < evs.size(); i += 2) {
		float k = 0;
		HIPCHECK(hipEventElapsedTime(&k, evs[i], evs[i + 1]));
		out->stage_ms += k;
		out->n_stage_launches++;
	}
	for (auto e : evs) ignore(hipEventDestroy(e));
	ignore(hipEventDestroy(t0));
	ignore(hipEventDestroy(t1));
	return 0;
}

// =================================================================== batch
// Many independent grids sharing one kernel configuration, advanced together
// by one persistent launch per stage. Each grid keeps its own dt (the
// reference's per-solver calcDT) and time; dt slots rotate in lock step.

struct hydro_batch {
	std::vector<Grid> grids;
	std::vector<size_t> offsets;   // of each grid in the arena
	char *arena;
	size_t arenaBytes;
	void *descs;                   // device GridDesc array
	size_t descBytes;
	hipStream_t stream;
	int slot;
	int dtValid;
	int rowsPerWave;
	int xcdRemap;
	int wavesPerBlock;
	std::vector<int> active;
	hipGraph_t graph;
	hipGraphExec_t graphExec;
	int descsDirty;
	int primed;                    // a direct launch has run (host-side caches filled)
};

static int same_kernel_config(const hydro_cfg &a, const hydro_cfg &b) {
	return a.dim == b.dim && a.eqn == b.eqn && a.flux == b.flux && a.flux_limiter == b.flux_limiter
		&& a.plm == b.plm && a.slope_limiter == b.slope_limiter && a.integrator == b.integrator
		&& a.precision == b.precision && a.fast_math == b.fast_math && a.gamma == b.gamma
		&& a.rho_min == b.rho_min && a.P_min == b.P_min && a.div_phi_wavespeed == b.div_phi_wavespeed
		&& a.div_psi_wavespeed == b.div_psi_wavespeed && a.use_fixed_dt == b.use_fixed_dt
		&& (!a.use_fixed_dt || a.fixed_dt == b.fixed_dt);
}

static void batch_drop_graph(hydro_batch *b) {
	if (b->graphExec) ignore(hipGraphExecDestroy(b->graphExec));
	if (b->graph) ignore(hipGraphDestroy(b->graph));
	b->graphExec = nullptr;
	b->graph = nullptr;
}

static long batch_items(const hydro_batch *b) {
	long n = 0;
	for (auto &g : b->grids) {
		int rows = g.dim >= 2 ? b->rowsPerWave : 1;
		long sx = (g.n[0] + STRIP_OUT - 1) / STRIP_OUT;
		long sy = g.dim >= 2 ? (g.n[1] + rows - 1) / rows : 1;
		n += sx * sy;
	}
	return n;
}

template <typename R>
static int batch_write_descs(hydro_batch *b) {
	std::vector<GridDesc<R>> d(b->grids.size());
	long base = 0;
	for (size_t i = 0; i < b->grids.size(); ++i) {
		Grid &g = b->grids[i];
		g.rowsPerWave = g.dim >= 2 ? b->rowsPerWave : 1;
		grid_desc<R>(g, d[i]);
		d[i].itemBase = base;
		d[i].active = b->active[i];
		base += d[i].nItems;
	}
	HIPCHECK(hipMemcpyAsync(b->descs, d.data(), d.size() * sizeof(GridDesc<R>), hipMemcpyHostToDevice, b->stream));
	HIPCHECK(hipStreamSynchronize(b->stream));
	b->descsDirty = 0;
	return 0;
}

template <typename R>
static int batch_stage(hydro_batch *b, int s, int cur) {
	const Grid &g0 = b->grids[0];
	int o = g0.rk.order;
	int fixed = g0.cfg.use_fixed_dt;
	bool last = s == o - 1;
	StageArgs<R> a;
	stage_args<R>(g0, s, a);
	if (!fixed) a.dtSlot = cur;
	else a.hostDt = g0.cfg.fixed_dt;
	a.resetSlot = (!fixed && s == 0) ? (cur + 2) % 3 : -1;
	a.reduceSlot = (!fixed && last) ? (cur + 1) % 3 : -1;
	a.advanceTime = last;
	a.stepInTick = 0;
	a.grids = (const GridDesc<R> *)b->descs;
	a.nGrids = (int)b->grids.size();
	a.totalItems = batch_items(b);
	a.xcdRemap = b->xcdRemap;
	StageKey k = stage_key(g0);
	// one kernel for every grid: take the widest velocity set
	int nv = 0;
	for (auto &g : b->grids) nv = std::max(nv, g.nvel);
	k.nvel = nv;
	hipError_t e = launch_stage<R>(k, a, dim3(0), dim3(64 * b->wavesPerBlock), b->stream);
	if (e != hipSuccess) return fail("batched stage launch: %s", hipGetErrorString(e));
	return 0;
}

// One step of every active grid with the slots rotated to cur.
template <typename R>
static int batch_step(hydro_batch *b, int cur) {
	int o = b->grids[0].rk.order;
	for (int s = 0; s < o; ++s)
		if (batch_stage<R>(b, s, cur)) return -1;
	if (o == 1) {
		// forward Euler wrote into ping: swap every grid's buffers, which
		// changes the descriptors of the next step
		for (auto &g : b->grids) std::swap(g.U[0], g.ping);
		if (batch_write_descs<R>(b)) return -1;
	}
	return 0;
}

extern "C" int hydro_batch_create(const hydro_cfg *cfgs, int n, hydro_batch **out) {
	if (!out) return fail("out is NULL");
	*out = nullptr;
	if (!cfgs || n < 1) return fail("need at least one grid");
	for (int i = 0; i < n; ++i) {
		if (validate(&cfgs[i])) return fail("grid %d: %s", i, g_err.c_str());
		if (!same_kernel_config(cfgs[0], cfgs[i]))
			return fail("grid %d: a batch shares eqn, dim, flux, limiters, plm, integrator, precision, math and "
			            "equation parameters", i);
	}
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
	hydro_batch *b = new hydro_batch();
	b->grids.resize(n);
	b->xcdRemap = 1;
	b->wavesPerBlock = 4;
	long cells = 0;
	for (int i = 0; i < n; ++i) {
		grid_init(b->grids[i], cfgs[i]);
		cells += b->grids[i].ncell;
	}
	b->rowsPerWave = cfgs[0].dim >= 2 ? (int)std::max(4L, std::min(32L, cells / (60L * 4096) + 1)) : 1;
	b->active.assign(n, 1);
	auto cleanup = [&](int rc) { hydro_batch_destroy(b); return rc; };
	if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
		delete b;
		return fail("hipStreamCreate failed");
	}
	size_t total = 0;
	for (auto &g : b->grids) {
		b->offsets.push_back(total);
		total += grid_bytes(g);
	}
	b->arenaBytes = total;
	if (hipMalloc((void **)&b->arena, total) != hipSuccess) return cleanup(fail("hipMalloc %zu B of batch state", total));
	for (int i = 0; i < n; ++i) {
		grid_bind(b->grids[i], b->arena + b->offsets[i]);
		if (grid_clear(b->grids[i], b->arena + b->offsets[i], b->stream)) return cleanup(-1);
	}
	size_t descSize = cfgs[0].precision == HYDRO_REAL_F32 ? sizeof(GridDesc<float>) : sizeof(GridDesc<double>);
	b->descBytes = descSize * n;
	if (hipMalloc(&b->descs, b->descBytes) != hipSuccess) return cleanup(fail("hipMalloc descriptors"));
	b->descsDirty = 1;
	if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(fail("sync"));
	*out = b;
	return 0;
}

extern "C" int hydro_batch_destroy(hydro_batch *b) {
	if (!b) return 0;
	if (b->stream) ignore(hipStreamSynchronize(b->stream));
	batch_drop_graph(b);
	if (b->arena) ignore(hipFree(b->arena));
	if (b->descs) ignore(hipFree(b->descs));
	if (b->stream) ignore(hipStreamDestroy(b->stream));
	delete b;
	return 0;
}

extern "C" int hydro_batch_size(const hydro_batch *b) { return b ? (int)b->grids.size() : 0; }

extern "C" size_t hydro_batch_device_bytes(const hydro_batch *b) { return b ? b->arenaBytes + b->descBytes : 0; }

static int batch_index(hydro_batch *b, int i) {
	if (!b) return fail("NULL handle");
	if (i < 0 || i >= (int)b->grids.size()) return fail("grid index %d out of range", i);
	return 0;
}

extern "C" int hydro_batch_upload_U(hydro_batch *b, int i, const void *U, size_t bytes) {
	if (batch_index(b, i)) return -1;
	int nv = b->grids[i].nvel;
	int rc = grid_upload(b->grids[i], U, bytes, b->stream);
	if (b->grids[i].nvel != nv) batch_drop_graph(b);
	b->dtValid = 0;
	return rc;
}

extern "C" int hydro_batch_download_U(hydro_batch *b, int i, void *U, size_t bytes) {
	if (batch_index(b, i)) return -1;
	return grid_download(b->grids[i], U, bytes, b->stream);
}

extern "C" int hydro_batch_constrain_U(hydro_batch *b) {
	if (!b) return fail("NULL handle");
	for (auto &g : b->grids)
		if (grid_constrain(g, b->stream)) return -1;
	b->dtValid = 0;
	return 0;
}

extern "C" int hydro_batch_get_time(hydro_batch *b, int i, double *t, double *dt) {
	if (batch_index(b, i)) return -1;
	return grid_get_time(b->grids[i], t, dt, b->stream);
}

extern "C" int hydro_batch_set_active(hydro_batch *b, int i, int active) {
	if (batch_index(b, i)) return -1;
	b->active[i] = active ? 1 : 0;
	b->descsDirty = 1;
	return 0;
}

extern "C" int hydro_batch_set_launch(hydro_batch *b, int rows_per_wave, int waves_per_block) {
	if (!b) return fail("NULL handle");
	if (rows_per_wave < 0 || rows_per_wave > 4096) return fail("rows_per_wave %d out of range", rows_per_wave);
	if (waves_per_block != 0 && waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
		return fail("waves_per_block must be 1, 2 or 4");
	if (rows_per_wave > 0) b->rowsPerWave = rows_per_wave;
	if (waves_per_block > 0) b->wavesPerBlock = waves_per_block;
	b->descsDirty = 1;
	batch_drop_graph(b);
	return 0;
}

extern "C" int hydro_batch_get_launch(const hydro_batch *b, int *rows_per_wave, int *waves_per_block) {
	if (!b) return fail("NULL handle");
	if (rows_per_wave) *rows_per_wave = b->rowsPerWave;
	if (waves_per_block) *waves_per_block = b->wavesPerBlock;
	return 0;
}

// descriptors current, and every active grid's dt in the current slot
static int batch_prepare(hydro_batch *b) {
	bool f64 = b->grids[0].elem == 8;
	if (b->descsDirty) {
		batch_drop_graph(b);
		if (f64 ? batch_write_descs<double>(b) : batch_write_descs<float>(b)) return -1;
	}
	if (!b->grids[0].cfg.use_fixed_dt && !b->dtValid) {
		for (size_t i = 0; i < b->grids.size(); ++i) {
			Grid &g = b->grids[i];
			HIPCHECK(launch_reset_slots(g.dev, 7, b->stream));
			if (b->active[i] && grid_dt_kernel(g, b->slot, b->stream)) return -1;
		}
		b->dtValid = 1;
	}
	return 0;
}

static int batch_run_direct(hydro_batch *b, int nsteps) {
	bool f64 = b->grids[0].elem == 8;
	for (int i = 0; i < nsteps; ++i) {
		int rc = f64 ? batch_step<double>(b, b->slot) : batch_step<float>(b, b->slot);
		if (rc) return rc;
		b->slot = (b->slot + 1) % 3;
		b->primed = 1;
	}
	return 0;
}

// Advance every active grid nsteps. With use_graph the steps are replayed
// from a captured graph of three steps (one turn of the dt-slot rotation);
// forward Euler (buffer swap per step) always launches directly.
extern "C" int hydro_batch_update(hydro_batch *b, int nsteps, int use_graph) {
	if (!b) return fail("NULL handle");
	if (nsteps < 0) return fail("nsteps < 0");
	if (batch_prepare(b)) return -1;
	bool f64 = b->grids[0].elem == 8;
	int o = b->grids[0].rk.order;
	if (!use_graph || o == 1) return batch_run_direct(b, nsteps);
	if (!b->primed && nsteps > 0) {
		// the first launch queries occupancy; keep that out of the capture
		if (batch_run_direct(b, 1)) return -1;
		--nsteps;
	}
	b->primed = 1;
	// align to the rotation, then whole graph turns, then the remainder
	while (nsteps > 0 && b->slot != 0) {
		if (batch_run_direct(b, 1)) return -1;
		--nsteps;
	}
	if (nsteps >= 3) {
		if (!b->graphExec) {
			HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
			int rc = 0;
			for (int k = 0; k < 3 && !rc; ++k)
				rc = f64 ? batch_step<double>(b, k) : batch_step<float>(b, k);
			hipGraph_t graph = nullptr;
			hipError_t e = hipStreamEndCapture(b->stream, &graph);
			if (rc) {
				if (graph) ignore(hipGraphDestroy(graph));
				return rc;
			}
			if (e != hipSuccess) return fail("graph capture: %s", hipGetErrorString(e));
			b->graph = graph;
			HIPCHECK(hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0));
		}
		while (nsteps >= 3) {
			HIPCHECK(hipGraphLaunch(b->graphExec, b->stream));
			nsteps -= 3;
		}
	}
	return batch_run_direct(b, nsteps);
}

extern "C" int hydro_batch_sync(hydro_batch *b) {
	if (!b) return fail("NULL handle");
	HIPCHECK(hipStreamSynchronize(b->stream));
	return 0;
}

extern "C" int hydro_batch_time_updates(hydro_batch *b, int nsteps, int use_graph, double *ms) {
	if (!b || !ms) return fail("NULL argument");
	if (batch_prepare(b)) return -1;
	hipEvent_t t0, t1;
	HIPCHECK(hipEventCreate(&t0));
	HIPCHECK(hipEventCreate(&t1));
	HIPCHECK(hipEventRecord(t0, b->stream));
	int rc = hydro_batch_update(b, nsteps, use_graph);
	if (rc) return rc;
	HIPCHECK(hipEventRecord(t1, b->stream));
	HIPCHECK(hipEventSynchronize(t1));
	float f = 0;
	HIPCHECK(hipEventElapsedTime(&f, t0, t1));
	*ms = f;
	ignore(hipEventDestroy(t0));
	ignore(hipEventDestroy(t1));
	return 0;
}

// Tries rows-per-wave / waves-per-block candidates on trial_steps steps each
// (the state advances) and keeps the fastest.
extern "C" int hydro_batch_autotune(hydro_batch *b, int trial_steps, int *rows_out, int *waves_out, double *ms_out) {
	if (!b) return fail("NULL handle");
	if (trial_steps < 1) trial_steps = 3;
	static const int rowsCand[] = {2, 3, 4, 6, 8, 12, 16};
	static const int wavesCand[] = {4, 2, 1};
	int bestR = b->rowsPerWave, bestW = b->wavesPerBlock;
	double best = 1e300;
	bool twoD = b->grids[0].dim >= 2;
	for (int w : wavesCand) {
		for (int r : rowsCand) {
			if (!twoD && r != rowsCand[0]) continue;
			if (hydro_batch_set_launch(b, r, w)) return -1;
			double ms = 0;
			if (hydro_batch_time_updates(b, 1, 0, &ms)) return -1;   // warm
			if (hydro_batch_time_updates(b, trial_steps, 0, &ms)) return -1;
			if (ms < best) {
				best = ms;
				bestR = r;
				bestW = w;
			}
		}
	}
	if (hydro_batch_set_launch(b, bestR, bestW)) return -1;
	if (rows_out) *rows_out = bestR;
	if (waves_out) *waves_out = bestW;
	if (ms_out) *ms_out = best / trial_steps;
	return 0;
}
