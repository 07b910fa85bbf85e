// hydro_hip.cpp — the C ABI (include/hydro_hip.h) over the HIP kernels.
//
// A solver handle owns: the integrator's stage buffers (structure-of-arrays,
// interior cells only, one plane per integrated field), stored derivative
// buffers when the Shu-Osher table needs them, static planes (GLM-Maxwell's
// non-integrated fields), a small block of device scalars (dt slots, t) and a
// HIP stream. Ghost cells exist only at the ABI boundary: uploads drop them,
// downloads rebuild them from the boundary conditions.
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

// ------------------------------------------------------------------ solver
struct hydro_solver {
	hydro_cfg cfg;
	int dim, Q, NSTAT, S;
	int n[3];
	long ncell;
	long ps;
	size_t elem;
	RKTable rk;
	void *Ubuf[4];      // U^(0..order-1); Ubuf[0] holds the solver state
	void *Dbuf[4];      // stored derivatives D^(k) when later stages use them
	void *stat;         // static planes
	void *ping;         // forward-Euler output buffer (ping-pong with Ubuf[0])
	DevScalars *dev;
	KSum *partial;
	double *totalsDev;
	int nPartial;
	hipStream_t stream;
	int slot;           // dt slot of the next step
	int dtValid;        // slot holds the min for the current state
	int rowsPerWave;
	int xcdRemap;
	int nvel;           // Euler: velocity components that are not identically zero
	std::vector<double> ePot;   // Euler ePot of the last upload (dynamics-neutral)
	size_t devBytes;
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

static void choose_launch(hydro_solver *h) {
	// rows per wave: enough waves for ~4 per SIMD (1024 SIMDs) without making
	// the y-pipeline warm-up (2-3 rows) dominate
	if (h->dim < 2) { h->rowsPerWave = 1; return; }
	int nx = h->n[0], ny = h->n[1];
	long stripsX = (nx + STRIP_OUT - 1) / STRIP_OUT;
	long target = 4096;
	long rows = (stripsX * (long)ny + target - 1) / target;
	rows = std::max(4L, std::min(64L, rows));
	h->rowsPerWave = (int)rows;
}

extern "C" int hydro_create(const hydro_cfg *cfg, hydro_solver **out) {
	if (!out) return fail("out is NULL");
	*out = nullptr;
	if (validate(cfg)) return -1;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
	hydro_solver *h = new hydro_solver();
	h->cfg = *cfg;
	h->dim = cfg->dim;
	h->Q = hydro_eqn_num_int_states(cfg->eqn);
	h->S = hydro_eqn_num_states(cfg->eqn);
	h->NSTAT = cfg->eqn == HYDRO_EQN_GLM_MAXWELL ? 4 : 0;
	h->ncell = 1;
	for (int i = 0; i < 3; ++i) {
		h->n[i] = i < h->dim ? cfg->size[i] : 1;
		h->ncell *= h->n[i];
	}
	h->ps = h->ncell;
	h->elem = elem_size(*cfg);
	rk_table(cfg->integrator, h->rk);
	h->xcdRemap = 1;
	h->nvel = 3;
	choose_launch(h);
	auto cleanup = [&](int rc) { hydro_destroy(h); return rc; };
	if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
		delete h;
		return fail("hipStreamCreate failed");
	}
	size_t planeBytes = (size_t)h->ps * h->elem;
	size_t stateBytes = planeBytes * h->Q;
	int o = h->rk.order;
	for (int i = 0; i < o; ++i) {
		if (hipMalloc(&h->Ubuf[i], stateBytes) != hipSuccess) return cleanup(fail("hipMalloc state (%zu B)", stateBytes));
		h->devBytes += stateBytes;
		bool needD = false;
		for (int m = i + 1; m < o; ++m) needD = needD || h->rk.b[m][i] != 0;
		if (needD) {
			if (hipMalloc(&h->Dbuf[i], stateBytes) != hipSuccess) return cleanup(fail("hipMalloc deriv"));
			h->devBytes += stateBytes;
		}
	}
	if (o == 1) {
		if (hipMalloc(&h->ping, stateBytes) != hipSuccess) return cleanup(fail("hipMalloc ping"));
		if (hipMemsetAsync(h->ping, 0, stateBytes, h->stream) != hipSuccess) return cleanup(fail("clear ping"));
		h->devBytes += stateBytes;
	}
	for (int i = 0; i < o; ++i)
		if (h->Dbuf[i] && hipMemsetAsync(h->Dbuf[i], 0, stateBytes, h->stream) != hipSuccess) return cleanup(fail("clear deriv"));
	if (h->NSTAT) {
		if (hipMalloc(&h->stat, planeBytes * h->NSTAT) != hipSuccess) return cleanup(fail("hipMalloc static"));
		h->devBytes += planeBytes * h->NSTAT;
	}
	if (hipMalloc((void **)&h->dev, sizeof(DevScalars)) != hipSuccess) return cleanup(fail("hipMalloc scalars"));
	h->nPartial = 1024;
	if (hipMalloc((void **)&h->partial, sizeof(KSum) * h->nPartial * HYDRO_MAX_STATES) != hipSuccess) return cleanup(fail("hipMalloc partials"));
	if (hipMalloc((void **)&h->totalsDev, sizeof(double) * HYDRO_MAX_STATES) != hipSuccess) return cleanup(fail("hipMalloc totals"));
	DevScalars z;
	std::memset(&z, 0, sizeof z);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < DT_LANES; ++j) z.dtmin[i][j] = INFINITY;
	if (hipMemcpyAsync(h->dev, &z, sizeof z, hipMemcpyHostToDevice, h->stream) != hipSuccess) return cleanup(fail("init scalars"));
	for (int i = 0; i < o; ++i)
		if (hipMemsetAsync(h->Ubuf[i], 0, stateBytes, h->stream) != hipSuccess) return cleanup(fail("clear state"));
	if (hipStreamSynchronize(h->stream) != hipSuccess) return cleanup(fail("sync"));
	*out = h;
	return 0;
}

extern "C" int hydro_destroy(hydro_solver *h) {
	if (!h) return 0;
	if (h->stream) hipStreamSynchronize(h->stream);
	for (int i = 0; i < 4; ++i) {
		if (h->Ubuf[i]) hipFree(h->Ubuf[i]);
		if (h->Dbuf[i]) hipFree(h->Dbuf[i]);
	}
	if (h->ping) hipFree(h->ping);
	if (h->stat) hipFree(h->stat);
	if (h->dev) hipFree(h->dev);
	if (h->partial) hipFree(h->partial);
	if (h->totalsDev) hipFree(h->totalsDev);
	if (h->stream) hipStreamDestroy(h->stream);
	delete h;
	return 0;
}

extern "C" size_t hydro_num_cells(const hydro_solver *h) {
	size_t g = 1;
	for (int i = 0; i < 3; ++i) g *= (size_t)(i < h->dim ? h->n[i] + 2 * NG : 1);
	return g;
}

extern "C" size_t hydro_device_bytes(const hydro_solver *h) { return h ? h->devBytes : 0; }

extern "C" int hydro_set_launch(hydro_solver *h, int rows_per_wave, int xcd_remap) {
	if (!h) return fail("NULL handle");
	if (rows_per_wave < 0 || rows_per_wave > 4096) return fail("rows_per_wave %d out of range", rows_per_wave);
	if (rows_per_wave > 0 && h->dim >= 2) h->rowsPerWave = rows_per_wave;
	if (xcd_remap > 0) h->xcdRemap = xcd_remap == 1;
	return 0;
}

extern "C" int hydro_get_launch(const hydro_solver *h, int *rows_per_wave, int *xcd_remap) {
	if (!h) return fail("NULL handle");
	if (rows_per_wave) *rows_per_wave = h->rowsPerWave;
	if (xcd_remap) *xcd_remap = h->xcdRemap ? 1 : 2;
	return 0;
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

template <typename R>
static void pack_interior(const hydro_solver *h, const R *aos, std::vector<R> &planes, std::vector<R> &stat,
	std::vector<double> &epot)
{
	int gs[3];
	for (int i = 0; i < 3; ++i) gs[i] = i < h->dim ? h->n[i] + 2 * NG : 1;
	int o[3];
	for (int i = 0; i < 3; ++i) o[i] = i < h->dim ? NG : 0;
	planes.assign((size_t)h->Q * h->ps, R(0));
	stat.assign((size_t)h->NSTAT * h->ps, R(0));
	bool anyE = false;
	if (h->cfg.eqn == HYDRO_EQN_EULER) epot.assign(h->ncell, 0.);
	for (int k = 0; k < h->n[2]; ++k)
		for (int j = 0; j < h->n[1]; ++j)
			for (int i = 0; i < h->n[0]; ++i) {
				size_t g = (size_t)(i + o[0]) + (size_t)gs[0] * ((j + o[1]) + (size_t)gs[1] * (k + o[2]));
				size_t c = (size_t)i + (size_t)h->n[0] * (j + (size_t)h->n[1] * k);
				const R *u = aos + g * h->S;
				for (int f = 0; f < h->Q; ++f) planes[(size_t)f * h->ps + c] = u[f];
				for (int f = 0; f < h->NSTAT; ++f) stat[(size_t)f * h->ps + c] = u[h->Q + f];
				if (h->cfg.eqn == HYDRO_EQN_EULER) {
					epot[c] = (double)u[5];
					anyE = anyE || u[5] != R(0);
				}
			}
	if (!anyE) epot.clear();
}

template <typename R>
static void unpack_with_ghosts(const hydro_solver *h, const std::vector<R> &planes, const std::vector<R> &stat, R *aos) {
	int gs[3];
	for (int i = 0; i < 3; ++i) gs[i] = i < h->dim ? h->n[i] + 2 * NG : 1;
	const hydro_cfg &c = h->cfg;
	for (int k = 0; k < gs[2]; ++k)
		for (int j = 0; j < gs[1]; ++j)
			for (int i = 0; i < gs[0]; ++i) {
				int cc[3] = {i - (h->dim > 0 ? NG : 0), j - (h->dim > 1 ? NG : 0), k - (h->dim > 2 ? NG : 0)};
				HostMap m[3];
				for (int a = 0; a < 3; ++a) {
					if (a < h->dim) m[a] = host_map(cc[a], h->n[a], c.boundary[2 * a], c.boundary[2 * a + 1], 2 * a);
					else m[a] = HostMap{0, 0, -1};
				}
				size_t src = (size_t)m[0].idx + (size_t)h->n[0] * (m[1].idx + (size_t)h->n[1] * m[2].idx);
				R *u = aos + ((size_t)i + (size_t)gs[0] * (j + (size_t)gs[1] * k)) * h->S;
				for (int f = 0; f < h->S; ++f) {
					R v;
					if (f < h->Q) v = planes[(size_t)f * h->ps + src];
					else if (f - h->Q < h->NSTAT) v = stat[(size_t)(f - h->Q) * h->ps + src];
					else v = (h->cfg.eqn == HYDRO_EQN_EULER && !h->ePot.empty()) ? R(h->ePot[src]) : R(0);
					for (int a = 0; a < h->dim; ++a) {
						if (m[a].fixedFace >= 0) v = R(c.boundary_fixed[m[a].fixedFace][f]);
						else if (m[a].mirror && reflects(c.eqn, f, a)) v = -v;
					}
					u[f] = v;
				}
			}
}

template <typename R>
static int upload_t(hydro_solver *h, const void *U) {
	std::vector<R> planes, stat;
	pack_interior<R>(h, (const R *)U, planes, stat, h->ePot);
	// Euler: momentum components beyond dim that are zero everywhere stay zero
	// (their fluxes and wave amplitudes are exact zeros), so the kernels can
	// skip them
	h->nvel = 3;
	if (h->cfg.eqn == HYDRO_EQN_EULER && h->dim < 3) {
		bool zero = true;
		for (int f = 1 + h->dim; f <= 3 && zero; ++f)
			for (long c = 0; c < h->ps; ++c)
				if (planes[(size_t)f * h->ps + c] != R(0)) { zero = false; break; }
		if (zero) h->nvel = h->dim;
	}
	HIPCHECK(hipMemcpyAsync(h->Ubuf[0], planes.data(), planes.size() * sizeof(R), hipMemcpyHostToDevice, h->stream));
	if (h->NSTAT) HIPCHECK(hipMemcpyAsync(h->stat, stat.data(), stat.size() * sizeof(R), hipMemcpyHostToDevice, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	h->dtValid = 0;
	return 0;
}

template <typename R>
static int download_t(hydro_solver *h, void *U) {
	std::vector<R> planes((size_t)h->Q * h->ps), stat((size_t)h->NSTAT * h->ps);
	HIPCHECK(hipMemcpyAsync(planes.data(), h->Ubuf[0], planes.size() * sizeof(R), hipMemcpyDeviceToHost, h->stream));
	if (h->NSTAT) HIPCHECK(hipMemcpyAsync(stat.data(), h->stat, stat.size() * sizeof(R), hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	unpack_with_ghosts<R>(h, planes, stat, (R *)U);
	return 0;
}

extern "C" int hydro_upload_U(hydro_solver *h, const void *U, size_t bytes) {
	if (!h || !U) return fail("NULL argument");
	size_t need = hydro_num_cells(h) * h->S * h->elem;
	if (bytes != need) return fail("upload size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	return h->elem == 8 ? upload_t<double>(h, U) : upload_t<float>(h, U);
}

extern "C" int hydro_download_U(hydro_solver *h, void *U, size_t bytes) {
	if (!h || !U) return fail("NULL argument");
	size_t need = hydro_num_cells(h) * h->S * h->elem;
	if (bytes != need) return fail("download size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	return h->elem == 8 ? download_t<double>(h, U) : download_t<float>(h, U);
}

// ------------------------------------------------------------ kernel calls
template <typename R>
static GridArgs<R> grid_args(hydro_solver *h, int slot) {
	GridArgs<R> g;
	std::memset(&g, 0, sizeof g);
	g.U = (R *)h->Ubuf[0];
	g.stat = (const R *)h->stat;
	g.ps = h->ps;
	g.ncell = h->ncell;
	g.dim = h->dim;
	for (int i = 0; i < 3; ++i) g.dx[i] = (h->cfg.maxs[i] - h->cfg.mins[i]) / (double)h->n[i];
	g.gamma = h->cfg.gamma;
	g.rhoMin = h->cfg.rho_min;
	g.PMin = h->cfg.P_min;
	g.divPhi = h->cfg.div_phi_wavespeed;
	g.divPsi = h->cfg.div_psi_wavespeed;
	g.dev = h->dev;
	g.slot = slot;
	return g;
}

static int run_dt_kernel(hydro_solver *h, int slot) {
	HIPCHECK(launch_reset_slots(h->dev, 1 << slot, h->stream));
	hipError_t e = h->elem == 8 ? launch_dt<double>(h->cfg.eqn, h->dim, h->cfg.fast_math, grid_args<double>(h, slot), h->stream)
	                            : launch_dt<float>(h->cfg.eqn, h->dim, h->cfg.fast_math, grid_args<float>(h, slot), h->stream);
	if (e != hipSuccess) return fail("dt kernel: %s", hipGetErrorString(e));
	return 0;
}

extern "C" int hydro_constrain_U(hydro_solver *h) {
	if (!h) return fail("NULL handle");
	hipError_t e = h->elem == 8 ? launch_constrain<double>(h->cfg.eqn, h->cfg.fast_math, grid_args<double>(h, 0), h->stream)
	                            : launch_constrain<float>(h->cfg.eqn, h->cfg.fast_math, grid_args<float>(h, 0), h->stream);
	if (e != hipSuccess) return fail("constrainU kernel: %s", hipGetErrorString(e));
	h->dtValid = 0;
	return 0;
}

extern "C" int hydro_boundary(hydro_solver *h) { return h ? 0 : fail("NULL handle"); }

extern "C" int hydro_calc_dt(hydro_solver *h, double *dt) {
	if (!h || !dt) return fail("NULL argument");
	if (h->cfg.use_fixed_dt) { *dt = h->cfg.fixed_dt; return 0; }
	if (run_dt_kernel(h, h->slot)) return -1;
	HIPCHECK(launch_fold_slot(h->dev, h->slot, h->stream));
	double m;
	HIPCHECK(hipMemcpyAsync(&m, &h->dev->dtmin[h->slot][0], sizeof m, hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	h->dtValid = 1;
	*dt = h->cfg.cfl * m;
	return 0;
}

// Fills the argument block of stage s (0-based) of the solver's table.
template <typename R>
static void stage_args(hydro_solver *h, int s, StageArgs<R> &a, void *out) {
	std::memset(&a, 0, sizeof a);
	const RKTable &t = h->rk;
	a.src = (const R *)(s == 0 ? h->Ubuf[0] : h->Ubuf[s]);
	a.out = (R *)out;
	a.stat = (const R *)h->stat;
	a.nAlpha = 0;
	for (int k = 0; k <= s; ++k) {
		if (t.a[s][k] != 0) {
			a.accU[a.nAlpha] = (const R *)h->Ubuf[k];
			a.alpha[a.nAlpha] = t.a[s][k];
			a.nAlpha++;
		}
	}
	a.nStoredD = 0;
	for (int k = 0; k < s; ++k) {
		if (t.b[s][k] != 0) {
			a.accD[a.nStoredD] = (const R *)h->Dbuf[k];
			a.beta[a.nStoredD] = t.b[s][k];
			a.nStoredD++;
		}
	}
	a.storeD = (R *)h->Dbuf[s];
	a.betaNew = t.b[s][s];
	for (int i = 0; i < 3; ++i) a.n[i] = h->n[i];
	a.ps = h->ps;
	a.step[0] = 1;
	a.step[1] = h->n[0];
	a.step[2] = (long)h->n[0] * h->n[1];
	const hydro_cfg &c = h->cfg;
	for (int f = 0; f < 6; ++f) {
		a.bc[f] = c.boundary[f];
		// fixed ghost values: integrated fields, then static fields
		for (int j = 0; j < h->Q; ++j) a.fixedv[f][j] = R(c.boundary_fixed[f][j]);
		for (int j = 0; j < h->NSTAT; ++j) a.fixedv[f][h->Q + j] = R(c.boundary_fixed[f][h->Q + j]);
	}
	// fvsolver.cl:37-102 in the stored precision: volume = 1 * dx0 * dx1...,
	// areaL = 1 * prod(dx_other), areaL *= 1 / volume
	R dxr[3];
	for (int i = 0; i < 3; ++i) {
		dxr[i] = R((c.maxs[i] - c.mins[i]) / (double)h->n[i]);
		a.dx[i] = dxr[i];
		a.invdx[i] = R(1) / dxr[i];
		a.dxD[i] = (c.maxs[i] - c.mins[i]) / (double)h->n[i];
	}
	R volume = R(1);
	for (int i = 0; i < h->dim; ++i) volume = volume * dxr[i];
	a.derivOn = volume > R(1e-7);
	R invVolume = R(1) / volume;
	for (int side = 0; side < 3; ++side) {
		R area = R(1);
		for (int i = 0; i < h->dim; ++i)
			if (i != side) area = area * dxr[i];
		R aL = area;
		if (aL <= R(1e-7)) aL = R(0);
		a.areaOverVol[side] = aL * invVolume;
		a.fluxOn[side] = area > R(1e-7);
	}
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
	a.dev = h->dev;
	a.cfl = c.cfl;
	a.resetSlot = -1;
	a.reduceSlot = -1;
	a.dtSlot = -1;
	a.advanceTime = 0;
	a.rowsPerWave = h->rowsPerWave;
	a.stripsX = (h->n[0] + STRIP_OUT - 1) / STRIP_OUT;
	a.stripsY = h->dim >= 2 ? (h->n[1] + h->rowsPerWave - 1) / h->rowsPerWave : 1;
	long waves = (long)a.stripsX * a.stripsY;
	a.nBlocks = (int)((waves + 3) / 4);
	a.xcdRemap = h->xcdRemap;
}

static StageKey stage_key(const hydro_solver *h) {
	StageKey k;
	k.eqn = h->cfg.eqn;
	k.dim = h->dim;
	k.plm = h->cfg.plm != HYDRO_PLM_NONE;
	k.lim = !k.plm && h->cfg.flux_limiter != HYDRO_LIMITER_DONOR_CELL;
	k.limiter = k.plm ? h->cfg.slope_limiter : h->cfg.flux_limiter;
	k.fast = h->cfg.fast_math != 0;
	k.nvel = h->nvel;
	return k;
}

// Launch one stage; dtMode: 0 host dt, 1 device slot.
template <typename R>
static int launch_one_stage(hydro_solver *h, int s, int useSlot, double hostDt, int resetSlot, int reduceSlot,
	int advance, void *out, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr)
{
	StageArgs<R> a;
	stage_args<R>(h, s, a, out);
	if (useSlot) a.dtSlot = h->slot;
	else a.hostDt = hostDt;
	a.resetSlot = resetSlot;
	a.reduceSlot = reduceSlot;
	a.advanceTime = advance;
	if (ev0) HIPCHECK(hipEventRecord(ev0, h->stream));
	hipError_t e = launch_stage<R>(stage_key(h), a, dim3(a.nBlocks), dim3(256), h->stream);
	if (e != hipSuccess) return fail("stage kernel launch: %s", hipGetErrorString(e));
	if (ev1) HIPCHECK(hipEventRecord(ev1, h->stream));
	return 0;
}

// One integrator step. useSlot: dt = cfl * dtmin[slot] from the device.
template <typename R>
static int do_step(hydro_solver *h, int useSlot, double hostDt, int update, std::vector<hipEvent_t> *evs) {
	int o = h->rk.order;
	int cur = h->slot;
	int fixed = h->cfg.use_fixed_dt;
	for (int s = 0; s < o; ++s) {
		bool last = s == o - 1;
		void *out;
		if (!last) out = h->Ubuf[s + 1];
		else out = o == 1 ? h->ping : h->Ubuf[0];
		int resetSlot = (update && !fixed && s == 0) ? (cur + 2) % 3 : -1;
		int reduceSlot = (update && !fixed && last) ? (cur + 1) % 3 : -1;
		hipEvent_t e0 = nullptr, e1 = nullptr;
		if (evs) {
			HIPCHECK(hipEventCreate(&e0));
			HIPCHECK(hipEventCreate(&e1));
			evs->push_back(e0);
			evs->push_back(e1);
		}
		if (launch_one_stage<R>(h, s, useSlot, hostDt, resetSlot, reduceSlot, update && last, out, e0, e1)) return -1;
	}
	if (o == 1) std::swap(h->Ubuf[0], h->ping);
	if (update && !fixed) {
		h->slot = (cur + 1) % 3;
		h->dtValid = 1;
	}
	return 0;
}

extern "C" int hydro_step(hydro_solver *h, double dt) {
	if (!h) return fail("NULL handle");
	int rc = h->elem == 8 ? do_step<double>(h, 0, dt, 0, nullptr) : do_step<float>(h, 0, dt, 0, nullptr);
	h->dtValid = 0;
	return rc;
}

static int update_impl(hydro_solver *h, int nsteps, std::vector<hipEvent_t> *evs) {
	if (nsteps < 0) return fail("nsteps < 0");
	int fixed = h->cfg.use_fixed_dt;
	if (!fixed && !h->dtValid) {
		HIPCHECK(launch_reset_slots(h->dev, 7, h->stream));
		if (run_dt_kernel(h, h->slot)) return -1;
		h->dtValid = 1;
	}
	for (int i = 0; i < nsteps; ++i) {
		int rc = h->elem == 8 ? do_step<double>(h, !fixed, h->cfg.fixed_dt, 1, evs)
		                      : do_step<float>(h, !fixed, h->cfg.fixed_dt, 1, evs);
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
	DevScalars d;
	HIPCHECK(hipMemcpyAsync(&d, h->dev, sizeof d, hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	if (t) *t = d.t;
	if (dt) *dt = d.dt;
	return 0;
}

extern "C" int hydro_set_time(hydro_solver *h, double t) {
	if (!h) return fail("NULL handle");
	DevScalars d;
	HIPCHECK(hipMemcpyAsync(&d, h->dev, sizeof d, hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	d.t = t;
	HIPCHECK(hipMemcpyAsync(h->dev, &d, sizeof d, hipMemcpyHostToDevice, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	return 0;
}

extern "C" int hydro_conserved_totals(hydro_solver *h, double *out) {
	if (!h || !out) return fail("NULL argument");
	hipError_t e = h->elem == 8
		? launch_totals<double>((const double *)h->Ubuf[0], h->ps, h->ncell, h->Q, h->partial, h->nPartial, h->totalsDev, h->stream)
		: launch_totals<float>((const float *)h->Ubuf[0], h->ps, h->ncell, h->Q, h->partial, h->nPartial, h->totalsDev, h->stream);
	if (e != hipSuccess) return fail("totals kernel: %s", hipGetErrorString(e));
	HIPCHECK(hipMemcpyAsync(out, h->totalsDev, sizeof(double) * h->Q, hipMemcpyDeviceToHost, h->stream));
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
	hipEvent_t t0, t1;
	HIPCHECK(hipEventCreate(&t0));
	HIPCHECK(hipEventCreate(&t1));
	std::vector<hipEvent_t> evs;
	if (!h->cfg.use_fixed_dt && !h->dtValid) {
		hipEvent_t d0, d1;
		HIPCHECK(hipEventCreate(&d0));
		HIPCHECK(hipEventCreate(&d1));
		HIPCHECK(hipEventRecord(d0, h->stream));
		HIPCHECK(launch_reset_slots(h->dev, 7, h->stream));
		if (run_dt_kernel(h, h->slot)) return -1;
		HIPCHECK(hipEventRecord(d1, h->stream));
		h->dtValid = 1;
		HIPCHECK(hipEventSynchronize(d1));
		float ms = 0;
		HIPCHECK(hipEventElapsedTime(&ms, d0, d1));
		out->dt_ms = ms;
		out->n_dt_launches = 1;
		hipEventDestroy(d0);
		hipEventDestroy(d1);
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
	for (auto e : evs) hipEventDestroy(e);
	hipEventDestroy(t0);
	hipEventDestroy(t1);
	return 0;
}
