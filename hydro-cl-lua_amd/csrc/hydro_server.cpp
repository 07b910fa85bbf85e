// hydro_server.cpp — continuous-batching solve server.
//
// Requests (a solver configuration, an initial state and a step count) queue on
// submit. Each tick admits queued requests in arrival order while their device
// state fits the HBM budget (requests that do not fit wait; smaller ones behind
// them may pass), uploads them into a first-fit arena, advances every resident
// request by up to steps_per_tick steps, and retires the finished ones (their
// final state is copied out and their memory freed). Requests are grouped in
// pools by kernel configuration; each pool's steps are persistent batched
// launches on the pool's own stream, so pools run concurrently on the device.
// A request's result does not depend on what else is resident: per-grid dt
// and time, no arithmetic across grids.
#include <deque>
#include <map>

#include "fv_host.h"

using namespace hydro;

namespace {

// first-fit allocator over one device allocation, with coalescing
struct Arena {
	char *base = nullptr;
	size_t size = 0;
	std::map<size_t, size_t> freeBlocks;   // offset -> bytes
	size_t used = 0;

	void init(char *b, size_t n) {
		base = b;
		size = n;
		freeBlocks.clear();
		freeBlocks[0] = n;
		used = 0;
	}
	long alloc(size_t n) {
		for (auto it = freeBlocks.begin(); it != freeBlocks.end(); ++it) {
			if (it->second >= n) {
				size_t off = it->first, rem = it->second - n;
				freeBlocks.erase(it);
				if (rem) freeBlocks[off + n] = rem;
				used += n;
				return (long)off;
			}
		}
		return -1;
	}
	void release(size_t off, size_t n) {
		used -= n;
		auto it = freeBlocks.emplace(off, n).first;
		auto nx = std::next(it);
		if (nx != freeBlocks.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			freeBlocks.erase(nx);
		}
		if (it != freeBlocks.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				freeBlocks.erase(it);
			}
		}
	}
};

struct Request {
	long long id = 0;
	hydro_cfg cfg;
	std::vector<char> U0;
	int nsteps = 0, remaining = 0;
	int state = 0;             // 0 queued, 1 resident, 2 finished
	Grid g;
	size_t off = 0, bytes = 0;
	int pool = -1;
	std::vector<char> result;
	int status = 0;            // 0 ok, 1 non-finite values in the result
	double t = 0;
};

struct Pool {
	hydro_cfg cfg;
	int nvel = 3;
	std::vector<long long> resident;
	void *descs = nullptr;
	size_t descCap = 0;
	std::vector<char> hostDescs;
	hipStream_t stream = nullptr;
	int slot = 0;
	int rowsPerWave = 12;
	int wavesPerBlock = 1;
};

bool same_pool(const hydro_cfg &a, const hydro_cfg &b) {
	return a.dim == b.dim && a.eqn == b.eqn && a.flux == b.flux && a.flux_limiter == b.flux_limiter
		&& a.plm == b.plm && a.slope_limiter == b.slope_limiter && a.integrator == b.integrator
		&& a.precision == b.precision && a.fast_math == b.fast_math && a.gamma == b.gamma
		&& a.rho_min == b.rho_min && a.P_min == b.P_min && a.div_phi_wavespeed == b.div_phi_wavespeed
		&& a.div_psi_wavespeed == b.div_psi_wavespeed && a.use_fixed_dt == b.use_fixed_dt
		&& (!a.use_fixed_dt || a.fixed_dt == b.fixed_dt) && a.roe_entropy_fix == b.roe_entropy_fix
		&& (!a.roe_entropy_fix || a.roe_entropy_fix_factor == b.roe_entropy_fix_factor)
		&& a.gravity == b.gravity;
}

template <typename R>
bool all_finite(const std::vector<char> &buf) {
	const R *p = (const R *)buf.data();
	size_t n = buf.size() / sizeof(R);
	for (size_t i = 0; i < n; ++i)
		if (!std::isfinite(p[i])) return false;
	return true;
}

} // namespace

struct hydro_server {
	Arena arena;
	char *mem = nullptr;
	int stepsPerTick = 10;
	long long nextId = 0;
	long long ticks = 0;
	std::map<long long, Request> reqs;
	std::deque<long long> queued;
	std::vector<long long> finished;   // not yet fetched
	std::vector<Pool> pools;
	hipStream_t upload = nullptr;
	// launch shape of 2D pools: the batch autotuner's choice for an ensemble
	// of mixed 64..512 grids (12 rows per wave, one wave per block)
	int rowsPerWave = 12;
	int wavesPerBlock = 1;
};

extern "C" int hydro_server_create(size_t hbm_budget_bytes, int steps_per_tick, hydro_server **out) {
	if (!out) return fail("out is NULL");
	*out = nullptr;
	if (hbm_budget_bytes < (1u << 20)) return fail("HBM budget must be at least 1 MiB");
	if (steps_per_tick < 1) return fail("steps_per_tick must be >= 1");
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail("no HIP device available");
	hydro_server *s = new hydro_server();
	s->stepsPerTick = steps_per_tick;
	if (hipMalloc((void **)&s->mem, hbm_budget_bytes) != hipSuccess) {
		delete s;
		return fail("hipMalloc of the %zu B budget failed", hbm_budget_bytes);
	}
	s->arena.init(s->mem, hbm_budget_bytes);
	if (hipStreamCreateWithFlags(&s->upload, hipStreamNonBlocking) != hipSuccess) {
		hydro_server_destroy(s);
		return fail("hipStreamCreate failed");
	}
	*out = s;
	return 0;
}

extern "C" int hydro_server_destroy(hydro_server *s) {
	if (!s) return 0;
	for (auto &p : s->pools) {
		if (p.stream) {
			ignore(hipStreamSynchronize(p.stream));
			ignore(hipStreamDestroy(p.stream));
		}
		if (p.descs) ignore(hipFree(p.descs));
	}
	if (s->upload) {
		ignore(hipStreamSynchronize(s->upload));
		ignore(hipStreamDestroy(s->upload));
	}
	if (s->mem) ignore(hipFree(s->mem));
	delete s;
	return 0;
}

extern "C" int hydro_server_submit(hydro_server *s, const hydro_cfg *cfg, const void *U, size_t bytes, int nsteps,
	long long *id)
{
	if (!s || !cfg || !U || !id) return fail("NULL argument");
	if (validate(cfg)) return -1;
	if (cfg->dim == 3) return fail("the server runs 1D and 2D requests");
	RKTable t;
	rk_table(cfg->integrator, t);
	if (t.order < 2) return fail("the server runs the Runge-Kutta integrators (forward Euler swaps buffers every step)");
	if (nsteps < 1) return fail("nsteps must be >= 1");
	Request r;
	r.id = s->nextId++;
	r.cfg = *cfg;
	grid_init(r.g, *cfg);
	size_t need = ghosted_cells(r.g) * r.g.S * r.g.elem;
	if (bytes != need) return fail("state size %zu != %zu (numCells * numStates * sizeof(real))", bytes, need);
	r.bytes = grid_bytes(r.g);
	if (r.bytes > s->arena.size)
		return fail("request needs %zu B of device state, more than the whole %zu B budget", r.bytes, s->arena.size);
	r.U0.assign((const char *)U, (const char *)U + bytes);
	r.nsteps = r.remaining = nsteps;
	s->queued.push_back(r.id);
	*id = r.id;
	s->reqs.emplace(r.id, std::move(r));
	return 0;
}

static int pool_for(hydro_server *s, const Request &r) {
	for (size_t i = 0; i < s->pools.size(); ++i)
		if (same_pool(s->pools[i].cfg, r.cfg) && s->pools[i].nvel == r.g.nvel) return (int)i;
	Pool p;
	p.cfg = r.cfg;
	p.nvel = r.g.nvel;
	p.rowsPerWave = r.g.dim >= 2 ? s->rowsPerWave : 1;
	p.wavesPerBlock = s->wavesPerBlock;
	if (hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking) != hipSuccess) return fail("hipStreamCreate failed");
	s->pools.push_back(std::move(p));
	return (int)s->pools.size() - 1;
}

// Moves queued requests that fit the arena to their pools: state upload, then
// their dt into the pool's current slot. A request that fails any step gets its
// arena block back and stays queued, ahead of those not yet looked at and
// behind those that did not fit (the queue order is kept).
static int admit(hydro_server *s) {
	std::deque<long long> still;
	while (!s->queued.empty()) {
		long long id = s->queued.front();
		Request &r = s->reqs.at(id);
		long off = s->arena.alloc(r.bytes);
		if (off < 0) {
			s->queued.pop_front();
			still.push_back(id);
			continue;
		}
		r.off = (size_t)off;
		grid_bind(r.g, s->arena.base + r.off);
		int p = -1;
		bool ok = !grid_clear(r.g, s->arena.base + r.off, s->upload)
			&& !grid_upload(r.g, r.U0.data(), r.U0.size(), s->upload)      // synchronizes
			&& (p = pool_for(s, r)) >= 0;
		if (ok) {
			hipError_t e = launch_reset_slots(r.g.dev, 7, s->upload);
			if (e != hipSuccess) ok = !fail("reset of the dt slots: %s", hipGetErrorString(e));
		}
		ok = ok && (r.cfg.use_fixed_dt || !grid_dt_kernel(r.g, s->pools[p].slot, s->upload));
		if (!ok) {
			s->arena.release(r.off, r.bytes);
			for (auto it = still.rbegin(); it != still.rend(); ++it) s->queued.push_front(*it);
			return -1;
		}
		s->queued.pop_front();
		r.U0.clear();
		r.U0.shrink_to_fit();
		r.pool = p;
		r.state = 1;
		s->pools[p].resident.push_back(id);
	}
	s->queued.swap(still);
	HIPCHECK(hipStreamSynchronize(s->upload));
	return 0;
}

template <typename R>
static int pool_tick(hydro_server *s, Pool &p, int steps) {
	int n = (int)p.resident.size();
	size_t descBytes = sizeof(GridDesc<R>) * n;
	if (descBytes > p.descCap) {
		if (p.descs) {
			HIPCHECK(hipStreamSynchronize(p.stream));
			HIPCHECK(hipFree(p.descs));
			p.descs = nullptr;
		}
		size_t cap = std::max(descBytes * 2, (size_t)4096);
		HIPCHECK(hipMalloc(&p.descs, cap));
		p.descCap = cap;
	}
	p.hostDescs.resize(descBytes);
	GridDesc<R> *d = (GridDesc<R> *)p.hostDescs.data();
	long base = 0;
	int maxSteps = 0;
	for (int i = 0; i < n; ++i) {
		Request &r = s->reqs.at(p.resident[i]);
		r.g.rowsPerWave = p.rowsPerWave;
		grid_desc<R>(r.g, d[i]);
		d[i].itemBase = base;
		d[i].stepsLeft = std::min(r.remaining, steps);
		maxSteps = std::max(maxSteps, d[i].stepsLeft);
		base += d[i].nItems;
	}
	HIPCHECK(hipMemcpyAsync(p.descs, d, descBytes, hipMemcpyHostToDevice, p.stream));
	const Grid &rep = s->reqs.at(p.resident[0]).g;
	int o = rep.rk.order;
	for (int j = 0; j < maxSteps; ++j) {
		int cur = (p.slot + j) % 3;
		for (int st = 0; st < o; ++st)
			if (launch_batched_stage<R>(rep, p.nvel, st, cur, (const GridDesc<R> *)p.descs, n, base, j,
				p.wavesPerBlock, 1, p.stream)) return -1;
	}
	p.slot = (p.slot + maxSteps) % 3;
	for (int i = 0; i < n; ++i) {
		Request &r = s->reqs.at(p.resident[i]);
		r.remaining -= d[i].stepsLeft;
	}
	return 0;
}

static int retire(hydro_server *s, Pool &p, int *nDone) {
	std::vector<long long> keep;
	for (long long id : p.resident) {
		Request &r = s->reqs.at(id);
		if (r.remaining > 0) {
			keep.push_back(id);
			continue;
		}
		r.result.resize(ghosted_cells(r.g) * r.g.S * r.g.elem);
		if (grid_download(r.g, r.result.data(), r.result.size(), p.stream)) return -1;
		if (grid_get_time(r.g, &r.t, nullptr, p.stream)) return -1;
		bool ok = r.g.elem == 8 ? all_finite<double>(r.result) : all_finite<float>(r.result);
		r.status = ok ? 0 : 1;
		s->arena.release(r.off, r.bytes);
		r.state = 2;
		s->finished.push_back(id);
		++*nDone;
	}
	p.resident.swap(keep);
	return 0;
}

extern "C" int hydro_server_tick(hydro_server *s, int *n_completed) {
	if (!s) return fail("NULL handle");
	int done = 0;
	if (admit(s)) return -1;
	for (auto &p : s->pools) {
		if (p.resident.empty()) continue;
		int rc = p.cfg.precision == HYDRO_REAL_F32 ? pool_tick<float>(s, p, s->stepsPerTick)
		                                           : pool_tick<double>(s, p, s->stepsPerTick);
		if (rc) return rc;
	}
	for (auto &p : s->pools) {
		HIPCHECK(hipStreamSynchronize(p.stream));
		if (retire(s, p, &done)) return -1;
	}
	s->ticks++;
	if (n_completed) *n_completed = done;
	return 0;
}

extern "C" int hydro_server_poll(hydro_server *s, long long *ids, int max, int *n) {
	if (!s || !n) return fail("NULL argument");
	int k = 0;
	for (long long id : s->finished) {
		if (k >= max) break;
		if (ids) ids[k] = id;
		++k;
	}
	*n = k;
	return 0;
}

extern "C" int hydro_server_result(hydro_server *s, long long id, void *U, size_t bytes, double *t, int *status) {
	if (!s) return fail("NULL handle");
	auto it = s->reqs.find(id);
	if (it == s->reqs.end()) return fail("unknown request %lld", id);
	Request &r = it->second;
	if (r.state != 2) return fail("request %lld has not finished", id);
	if (U) {
		if (bytes != r.result.size()) return fail("result size %zu != %zu", bytes, r.result.size());
		std::memcpy(U, r.result.data(), bytes);
	}
	if (t) *t = r.t;
	if (status) *status = r.status;
	s->finished.erase(std::remove(s->finished.begin(), s->finished.end(), id), s->finished.end());
	s->reqs.erase(it);
	return 0;
}

extern "C" int hydro_server_info_get(hydro_server *s, hydro_server_info *info) {
	if (!s || !info) return fail("NULL argument");
	std::memset(info, 0, sizeof *info);
	info->queued = (int)s->queued.size();
	for (auto &p : s->pools) info->resident += (int)p.resident.size();
	info->finished = (int)s->finished.size();
	info->pools = (int)s->pools.size();
	info->bytes_in_use = s->arena.used;
	info->budget = s->arena.size;
	info->ticks = s->ticks;
	return 0;
}

extern "C" int hydro_server_set_launch(hydro_server *s, int rows_per_wave, int waves_per_block) {
	if (!s) return fail("NULL handle");
	if (rows_per_wave < 0 || rows_per_wave > EDGE_ROWS)
		return fail("rows_per_wave %d out of range 0..%d", rows_per_wave, EDGE_ROWS);
	if (waves_per_block != 0 && waves_per_block != 1 && waves_per_block != 2 && waves_per_block != 4)
		return fail("waves_per_block must be 1, 2 or 4");
	if (rows_per_wave > 0) s->rowsPerWave = rows_per_wave;
	if (waves_per_block > 0) s->wavesPerBlock = waves_per_block;
	for (auto &p : s->pools) {
		if (rows_per_wave > 0 && p.cfg.dim >= 2) p.rowsPerWave = rows_per_wave;
		if (waves_per_block > 0) p.wavesPerBlock = waves_per_block;
	}
	return 0;
}
