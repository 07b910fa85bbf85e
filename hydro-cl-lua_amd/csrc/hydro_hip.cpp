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

#include "fv_host.h"

using namespace hydro;

thread_local int hydro::stage_occupancy = 0;

// ------------------------------------------------------------------ errors
std::string &hydro::hydro_err_buf() {
	static thread_local std::string e;
	return e;
}

extern "C" const char *hydro_last_error(void) { return hydro_err_buf().c_str(); }
extern "C" int hydro_eqn_num_states(int eqn) { return num_states(eqn); }
extern "C" int hydro_eqn_num_int_states(int eqn) { return num_int_states(eqn); }
extern "C" void hydro_cfg_default(hydro_cfg *c) { cfg_default(c); }

extern "C" int hydro_integrator_table(int integrator, int *stages, double *alpha, double *beta) {
	RKTable t;
	if (!rk_table(integrator, t)) return fail("integrator %d not supported", integrator);
	if (stages) *stages = t.order;
	for (int i = 0; i < 4; ++i)
		for (int k = 0; k < 4; ++k) {
			if (alpha) alpha[4 * i + k] = t.a[i][k];
			if (beta) beta[4 * i + k] = t.b[i][k];
		}
	return 0;
}

extern "C" int hydro_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
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
	const int maxRows = h->g.dim == 3 ? EDGE_PLANES : EDGE_ROWS;
	if (rows_per_wave < 0 || rows_per_wave > maxRows)
		return fail("rows_per_wave %d out of range 0..%d", rows_per_wave, maxRows);
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
	a.nBlocks = (int)(g.dim == 3 ? a.g.nItems : (a.g.nItems + 3) / 4);
	// a block's four waves are x-adjacent strips of one row band when the
	// strips per row are a multiple of 4: one edge pass computes the five
	// edge columns they need (5 rows-per-wave interfaces in one wave)
	a.shareEdges = g.dim == 2 && strip_edges(2, eqn_has_source(g.cfg.eqn)) && a.g.stripsX % 4 == 0 &&
		5 * a.g.rowsPerWave <= WAVE;
	const bool redo = redo_kernels(g.cfg);
	if (ev0) HIPCHECK(hipEventRecord(ev0, h->stream));
	const StageKey k = stage_key(g);
	hipError_t e = launch_stage<R>(k, a, dim3(a.nBlocks), dim3(256), h->stream);
	if (e != hipSuccess) return fail("stage kernel launch: %s", hipGetErrorString(e));
	if (redo) {
		// the strips the stage kernel left to the general Roe form (usually none)
		a.redoPass = 1;
		a.shareEdges = 0;
		a.nBlocks = REDO_BLOCKS;
		e = launch_stage<R>(k, a, dim3(REDO_BLOCKS), dim3(256), h->stream);
		if (e != hipSuccess) return fail("redo kernel launch: %s", hipGetErrorString(e));
	}
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

extern "C" int hydro_var_stats(hydro_solver *h, int field, double *out) {
	if (!h || !out) return fail("NULL argument");
	Grid &g = h->g;
	const bool derived = field >= HYDRO_VAR_P && field < HYDRO_VAR_END;
	if (derived && g.cfg.eqn != HYDRO_EQN_EULER) return fail("the HYDRO_VAR_* display variables are the Euler equations'");
	if (!derived && (field < 0 || field >= g.Q + g.NSTAT))
		return fail("field %d out of range 0..%d or HYDRO_VAR_*", field, g.Q + g.NSTAT - 1);
	const double eqnPar[3] = {g.cfg.gamma, g.cfg.rho_min, g.cfg.P_min};
	double nInterior = (double)g.ncell;
	hipError_t e;
	if (g.elem == 8) {
		GridDesc<double> d;
		grid_desc<double>(g, d);
		e = launch_stats<double>(g.cfg.eqn, g.dim, d, field, h->partial, h->nPartial, nInterior, h->totalsDev, eqnPar,
			h->stream);
	} else {
		GridDesc<float> d;
		grid_desc<float>(g, d);
		e = launch_stats<float>(g.cfg.eqn, g.dim, d, field, h->partial, h->nPartial, nInterior, h->totalsDev, eqnPar,
			h->stream);
	}
	if (e != hipSuccess) return fail("stats kernel: %s", hipGetErrorString(e));
	HIPCHECK(hipMemcpyAsync(out, h->totalsDev, sizeof(double) * 4, hipMemcpyDeviceToHost, h->stream));
	HIPCHECK(hipStreamSynchronize(h->stream));
	return 0;
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
		&& (!a.use_fixed_dt || a.fixed_dt == b.fixed_dt) && a.roe_entropy_fix == b.roe_entropy_fix
		&& (!a.roe_entropy_fix || a.roe_entropy_fix_factor == b.roe_entropy_fix_factor)
		&& a.gravity == b.gravity;
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
		const int so = strip_out(strip_edges(g.dim, eqn_has_source(g.cfg.eqn)));
		long sx = (g.n[0] + so - 1) / so;
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
	int nv = 0;
	for (auto &g : b->grids) nv = std::max(nv, g.nvel);
	return launch_batched_stage<R>(b->grids[0], nv, s, cur, (const GridDesc<R> *)b->descs, (int)b->grids.size(),
		batch_items(b), 0, b->wavesPerBlock, b->xcdRemap, b->stream);
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
		if (validate(&cfgs[i])) return fail("grid %d: %s", i, hydro_err_buf().c_str());
		if (cfgs[i].dim == 3) return fail("grid %d: the batched form runs 1D and 2D grids", i);
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
	b->rowsPerWave = cfgs[0].dim >= 2 ? (int)std::max(4L, std::min(32L, cells / (64L * 4096) + 1)) : 1;
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
	const int on = active ? 1 : 0;
	// A grid that sat out kept none of the dt slots of the rotation current
	// (inactive grids neither reset nor reduce them), so a grid switched back
	// on recomputes dt from its state. batch_prepare then resets every grid's
	// slots and takes each active grid's dt from its state: the same minimum
	// its last stage reduced, since min is exact in any order.
	if (on && !b->active[i]) b->dtValid = 0;
	b->active[i] = on;
	b->descsDirty = 1;
	return 0;
}

extern "C" int hydro_batch_set_launch(hydro_batch *b, int rows_per_wave, int waves_per_block) {
	if (!b) return fail("NULL handle");
	if (rows_per_wave < 0 || rows_per_wave > EDGE_ROWS)
		return fail("rows_per_wave %d out of range 0..%d", rows_per_wave, EDGE_ROWS);
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
