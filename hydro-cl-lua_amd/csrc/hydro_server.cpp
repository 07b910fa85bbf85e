// hydro_server.cpp — continuous-batching solve server.
//
// Requests (a solver configuration, an initial state and a step count) queue on
// submit, their state packed into the device layout. Admission takes queued
// requests in arrival order while their device state fits the HBM budget and
// their host staging fits the pinned staging area (requests that do not fit
// wait; smaller ones behind them may pass): the packed state is copied into
// pinned memory and the request joins the pool of its kernel configuration. Each pool advances on its own
// stream in chunks of up to steps_per_tick steps: a chunk uploads the pool's
// joining requests, runs the persistent batched stage launches over every
// resident grid, and, for each request the chunk finishes, a non-finite check
// on the device and the copy of the final state into pinned memory; an event
// closes the chunk. No call waits for the whole device: a tick completes the
// chunks whose events have fired (results unpacked, memory released), admits,
// and starts the next chunk of every idle pool, so one pool's host work
// overlaps the others' kernels; only when nothing can move does it wait, for
// the chunk started first. A request's result does not depend on what else is
// resident: per-grid dt and time, no arithmetic across grids.
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

enum { QUEUED = 0, JOINING = 1, RESIDENT = 2, RETIRING = 3, FINISHED = 4 };

struct Request {
	long long id = 0;
	hydro_cfg cfg;
	std::vector<char> packed;  // the state in the device layout, from submit to admission
	int nsteps = 0;
	int remaining = 0;         // steps not yet launched
	int state = QUEUED;
	Grid g;
	size_t off = 0, bytes = 0; // device arena block
	long upOff = -1;           // pinned staging of the packed state, until its upload has run
	long outOff = -1;          // pinned staging of the result, from its copy-out on
	size_t stageBytes = 0;     // bytes of either block
	int pool = -1;
	std::vector<char> result;
	int status = 0;            // 0 ok, 1 non-finite values in the result
	double t = 0;
};

struct Pool {
	hydro_cfg cfg;
	int nvel = 3;
	std::vector<long long> joining;    // packed, uploaded by the pool's next chunk
	std::vector<long long> resident;   // on the device, advanced by the next chunk
	std::vector<long long> landing;    // uploaded by the chunk in flight (staging still held)
	std::vector<long long> retiring;   // copied out by the chunk in flight
	void *descs = nullptr;             // device GridDesc array
	size_t descCap = 0;
	char *hostDescs = nullptr;         // pinned, written only while the pool is idle
	size_t hostDescCap = 0;
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;         // closes the chunk in flight
	bool inflight = false;
	long long started = 0;             // launch order of the chunk in flight
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

// pinned staging of one request: its device-layout planes (q integrated, then
// static), then 64 bytes for t, dt and the non-finite flag
size_t plane_bytes(const Grid &g) { return (size_t)(g.Q + g.NSTAT) * g.ps * g.elem; }
size_t staging_bytes(const Grid &g) { return ((plane_bytes(g) + 255) & ~(size_t)255) + 256; }

// at most this much pinned host memory for staging (the budget when smaller)
constexpr size_t PINNED_MAX = (size_t)256 << 20;

} // namespace

struct hydro_server {
	Arena arena;               // device state
	Arena pinned;              // host staging
	char *mem = nullptr;
	char *pinnedMem = nullptr;
	int stepsPerTick = 10;
	long long nextId = 0;
	long long ticks = 0;
	long long chunks = 0;
	std::map<long long, Request> reqs;
	std::deque<long long> queued;
	std::vector<long long> finished;   // not yet fetched
	std::vector<Pool> pools;
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
	const size_t pinnedBytes = std::min(hbm_budget_bytes, PINNED_MAX);
	if (hipHostMalloc((void **)&s->pinnedMem, pinnedBytes, hipHostMallocDefault) != hipSuccess) {
		s->pinnedMem = nullptr;
		hydro_server_destroy(s);
		return fail("hipHostMalloc of %zu B of staging failed", pinnedBytes);
	}
	s->pinned.init(s->pinnedMem, pinnedBytes);
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
		if (p.done) ignore(hipEventDestroy(p.done));
		if (p.descs) ignore(hipFree(p.descs));
		if (p.hostDescs) ignore(hipHostFree(p.hostDescs));
	}
	if (s->pinnedMem) ignore(hipHostFree(s->pinnedMem));
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
	r.stageBytes = staging_bytes(r.g);
	if (r.stageBytes > s->pinned.size)
		return fail("request needs %zu B of host staging, more than the server's %zu B", r.stageBytes, s->pinned.size);
	// packed once, here, into the device layout (this also sets nvel, which
	// selects the pool)
	r.packed.resize(plane_bytes(r.g));
	char *planes = r.packed.data();
	char *stat = planes + (size_t)r.g.Q * r.g.ps * r.g.elem;
	if (r.g.elem == 8) pack_interior<double>(r.g, (const double *)U, (double *)planes, (double *)stat);
	else pack_interior<float>(r.g, (const float *)U, (float *)planes, (float *)stat);
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
	if (hipEventCreateWithFlags(&p.done, hipEventDisableTiming) != hipSuccess) {
		ignore(hipStreamDestroy(p.stream));
		return fail("hipEventCreate failed");
	}
	s->pools.push_back(std::move(p));
	return (int)s->pools.size() - 1;
}

// Queued requests that fit the device arena and the pinned staging, in arrival
// order (ones that do not fit wait, smaller ones may pass): the packed state
// is copied into pinned memory (host work only) and the request joins its
// pool. One call copies at most about ADMIT_BYTES (at least one
// request), so that pools start while later ticks copy the rest. A
// request that fails a step gets its blocks back and stays queued, ahead of
// those not yet looked at (the queue order is kept).
constexpr size_t ADMIT_BYTES = (size_t)64 << 20;

static int admit(hydro_server *s, bool *moved) {
	std::deque<long long> still;
	size_t copied = 0;
	while (!s->queued.empty() && copied < ADMIT_BYTES) {
		long long id = s->queued.front();
		Request &r = s->reqs.at(id);
		long off = s->arena.alloc(r.bytes);
		long st = off < 0 ? -1 : s->pinned.alloc(r.stageBytes);
		if (off < 0 || st < 0) {
			if (off >= 0) s->arena.release((size_t)off, r.bytes);
			s->queued.pop_front();
			still.push_back(id);
			continue;
		}
		r.off = (size_t)off;
		r.upOff = st;
		grid_bind(r.g, s->arena.base + r.off);
		char *stage = s->pinned.base + st;
		const char *src = r.packed.data();
		parallel_ranges((long)r.packed.size(), (long)1 << 20, [&](long b0, long b1) {
			std::memcpy(stage + b0, src + b0, (size_t)(b1 - b0));
		});
		int p = pool_for(s, r);
		if (p < 0) {
			s->arena.release(r.off, r.bytes);
			s->pinned.release((size_t)st, r.stageBytes);
			r.upOff = -1;
			for (auto it = still.rbegin(); it != still.rend(); ++it) s->queued.push_front(*it);
			return -1;
		}
		s->queued.pop_front();
		copied += r.stageBytes;
		r.packed.clear();
		r.packed.shrink_to_fit();
		r.pool = p;
		r.state = JOINING;
		s->pools[p].joining.push_back(id);
		*moved = true;
	}
	// those that did not fit go back ahead of the ones not looked at
	for (auto it = still.rbegin(); it != still.rend(); ++it) s->queued.push_front(*it);
	return 0;
}

// Enqueues the upload of a joining request on the pool's stream: zeroed
// buffers, dt slots at +inf, the packed state (from its staging block, which
// is released when the chunk ends), and its dt into the pool's slot.
static int enqueue_join(hydro_server *s, Pool &p, Request &r) {
	HIPCHECK(hipMemsetAsync(s->arena.base + r.off, 0, r.bytes, p.stream));
	HIPCHECK(launch_reset_slots(r.g.dev, 7, p.stream));
	const char *stage = s->pinned.base + r.upOff;
	const size_t qb = (size_t)r.g.Q * r.g.ps * r.g.elem;
	HIPCHECK(hipMemcpyAsync(r.g.U[0], stage, qb, hipMemcpyHostToDevice, p.stream));
	if (r.g.NSTAT) HIPCHECK(hipMemcpyAsync(r.g.stat, stage + qb, plane_bytes(r.g) - qb, hipMemcpyHostToDevice, p.stream));
	r.g.dtValid = 0;
	if (!r.cfg.use_fixed_dt && grid_dt_kernel(r.g, p.slot, p.stream)) return -1;
	return 0;
}

// Enqueues, after the chunk's kernels, the non-finite check of a finished
// request and the copy of its state, time and flag into its staging block.
static int enqueue_retire(hydro_server *s, Pool &p, Request &r) {
	hipError_t e = r.g.elem == 8
		? launch_nonfinite<double>((const double *)r.g.U[0], (const double *)r.g.stat, r.g.ps, r.g.ncell, r.g.Q,
			r.g.NSTAT, &r.g.dev->nonFinite, p.stream)
		: launch_nonfinite<float>((const float *)r.g.U[0], (const float *)r.g.stat, r.g.ps, r.g.ncell, r.g.Q,
			r.g.NSTAT, &r.g.dev->nonFinite, p.stream);
	if (e != hipSuccess) return fail("non-finite check: %s", hipGetErrorString(e));
	char *stage = s->pinned.base + r.outOff;
	const size_t qb = (size_t)r.g.Q * r.g.ps * r.g.elem;
	HIPCHECK(hipMemcpyAsync(stage, r.g.U[0], qb, hipMemcpyDeviceToHost, p.stream));
	if (r.g.NSTAT) HIPCHECK(hipMemcpyAsync(stage + qb, r.g.stat, plane_bytes(r.g) - qb, hipMemcpyDeviceToHost, p.stream));
	char *scal = stage + staging_bytes(r.g) - 256;
	// t, dt, ..., nonFinite: the scalars from t up to the flag
	HIPCHECK(hipMemcpyAsync(scal, &r.g.dev->t, offsetof(DevScalars, pad0) - offsetof(DevScalars, t), hipMemcpyDeviceToHost,
		p.stream));
	return 0;
}

// One chunk of an idle pool on its stream: the joining requests' uploads, up
// to steps_per_tick steps of every resident grid with steps left, the
// copy-out of every request with none left (as far as staging allows: the
// others wait for a later chunk), and the event that closes the chunk.
// Returns 1 when it enqueued anything, 0 when there was nothing to do.
template <typename R>
static int launch_chunk(hydro_server *s, Pool &p) {
	bool any = false;
	for (long long id : p.joining) {
		Request &r = s->reqs.at(id);
		if (enqueue_join(s, p, r)) return -1;
		r.state = RESIDENT;
		p.resident.push_back(id);
		p.landing.push_back(id);
		any = true;
	}
	p.joining.clear();
	std::vector<long long> active;
	for (long long id : p.resident)
		if (s->reqs.at(id).remaining > 0) active.push_back(id);
	const int n = (int)active.size();
	if (n) {
		const size_t descBytes = sizeof(GridDesc<R>) * n;
		if (descBytes > p.descCap) {
			// the pool is idle: nothing of its stream reads the old array
			if (p.descs) HIPCHECK(hipFree(p.descs));
			p.descs = nullptr;
			p.descCap = 0;
			const size_t cap = std::max(descBytes * 2, (size_t)4096);
			HIPCHECK(hipMalloc(&p.descs, cap));
			p.descCap = cap;
		}
		if (descBytes > p.hostDescCap) {
			if (p.hostDescs) HIPCHECK(hipHostFree(p.hostDescs));
			p.hostDescs = nullptr;
			p.hostDescCap = 0;
			const size_t cap = std::max(descBytes * 2, (size_t)4096);
			HIPCHECK(hipHostMalloc((void **)&p.hostDescs, cap, hipHostMallocDefault));
			p.hostDescCap = cap;
		}
		GridDesc<R> *d = (GridDesc<R> *)p.hostDescs;
		long base = 0;
		int maxSteps = 0;
		for (int i = 0; i < n; ++i) {
			Request &r = s->reqs.at(active[i]);
			r.g.rowsPerWave = p.rowsPerWave;
			grid_desc<R>(r.g, d[i]);
			d[i].itemBase = base;
			d[i].stepsLeft = std::min(r.remaining, s->stepsPerTick);
			maxSteps = std::max(maxSteps, d[i].stepsLeft);
			base += d[i].nItems;
		}
		HIPCHECK(hipMemcpyAsync(p.descs, d, descBytes, hipMemcpyHostToDevice, p.stream));
		const Grid &rep = s->reqs.at(active[0]).g;
		const int o = rep.rk.order;
		for (int j = 0; j < maxSteps; ++j) {
			const int cur = (p.slot + j) % 3;
			for (int st = 0; st < o; ++st)
				if (launch_batched_stage<R>(rep, p.nvel, st, cur, (const GridDesc<R> *)p.descs, n, base, j,
					p.wavesPerBlock, 1, p.stream)) return -1;
		}
		p.slot = (p.slot + maxSteps) % 3;
		for (int i = 0; i < n; ++i) s->reqs.at(active[i]).remaining -= d[i].stepsLeft;
		any = true;
	}
	// copy-out, behind the kernels, of the requests with no steps left
	std::vector<long long> keep;
	for (long long id : p.resident) {
		Request &r = s->reqs.at(id);
		if (r.remaining > 0) {
			keep.push_back(id);
			continue;
		}
		const long st = s->pinned.alloc(r.stageBytes);
		if (st < 0) {
			keep.push_back(id);      // staging is full: a later chunk copies it out
			continue;
		}
		r.outOff = st;
		if (enqueue_retire(s, p, r)) return -1;
		r.state = RETIRING;
		p.retiring.push_back(id);
		any = true;
	}
	p.resident.swap(keep);
	if (!any) return 0;
	HIPCHECK(hipEventRecord(p.done, p.stream));
	p.inflight = true;
	p.started = ++s->chunks;
	return 1;
}

// The chunk in flight of p has ended: the uploads' staging is released, and
// finished requests are unpacked from their staging into the result and their
// device and staging blocks released.
static int complete_chunk(hydro_server *s, Pool &p, int *nDone) {
	for (long long id : p.landing) {
		Request &r = s->reqs.at(id);
		s->pinned.release((size_t)r.upOff, r.stageBytes);
		r.upOff = -1;
	}
	p.landing.clear();
	for (long long id : p.retiring) {
		Request &r = s->reqs.at(id);
		const char *stage = s->pinned.base + r.outOff;
		const char *stat = stage + (size_t)r.g.Q * r.g.ps * r.g.elem;
		r.result.resize(ghosted_cells(r.g) * r.g.S * r.g.elem);
		if (r.g.elem == 8)
			unpack_with_ghosts<double>(r.g, (const double *)stage, (const double *)stat, (double *)r.result.data());
		else
			unpack_with_ghosts<float>(r.g, (const float *)stage, (const float *)stat, (float *)r.result.data());
		const char *scal = stage + r.stageBytes - 256;
		double td[2];
		unsigned flag;
		std::memcpy(td, scal, sizeof td);
		std::memcpy(&flag, scal + offsetof(DevScalars, nonFinite) - offsetof(DevScalars, t), sizeof flag);
		r.t = td[0];
		r.status = flag ? 1 : 0;
		s->arena.release(r.off, r.bytes);
		s->pinned.release((size_t)r.outOff, r.stageBytes);
		r.outOff = -1;
		r.state = FINISHED;
		s->finished.push_back(id);
		++*nDone;
	}
	p.retiring.clear();
	p.inflight = false;
	return 0;
}

extern "C" int hydro_server_tick(hydro_server *s, int *n_completed) {
	if (!s) return fail("NULL handle");
	int done = 0;
	for (;;) {
		bool moved = false;
		for (auto &p : s->pools) {
			if (!p.inflight) continue;
			hipError_t e = hipEventQuery(p.done);
			if (e == hipErrorNotReady) continue;
			if (e != hipSuccess) return fail("chunk event: %s", hipGetErrorString(e));
			if (complete_chunk(s, p, &done)) return -1;
			moved = true;
		}
		if (admit(s, &moved)) return -1;
		for (auto &p : s->pools) {
			if (p.inflight || (p.joining.empty() && p.resident.empty())) continue;
			int rc = p.cfg.precision == HYDRO_REAL_F32 ? launch_chunk<float>(s, p) : launch_chunk<double>(s, p);
			if (rc < 0) return rc;
			moved = moved || rc > 0;
		}
		if (moved) break;
		// nothing could move: wait for the chunk started first, then go again
		Pool *first = nullptr;
		for (auto &p : s->pools)
			if (p.inflight && (!first || p.started < first->started)) first = &p;
		if (!first) break;     // idle: nothing resident, nothing that fits
		HIPCHECK(hipEventSynchronize(first->done));
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
	if (r.state != FINISHED) return fail("request %lld has not finished", id);
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
	for (auto &p : s->pools)
		info->resident += (int)(p.joining.size() + p.resident.size() + p.retiring.size());
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
