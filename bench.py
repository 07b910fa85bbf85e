"""Benchmark of the finite-volume hot path on one MI355X.

Headline (`value`): 2D Euler Kelvin-Helmholtz, 1024x1024 periodic, Roe +
plm cons (minmod), RK2-TVD, CFL 0.4, FP64, in Gcell-updates/s. A step is one
full RK2-TVD update of the grid (two fused stage launches; dt stays on the
device). The state is resident in HBM before the timed region, which holds
exactly --steps steps between two device synchronizations.

Beside it on the same JSON line:
  roofline      the stage kernel: algorithmic bytes per launch / average launch
                duration (HIP events on the solver stream around the timed
                region, which holds only stage launches, over their count)
  hbm_regime    the same KH problem on 4096^2 (1.07 GB of FP64 state, beyond the
                256 MB last-level cache), FP64 and FP32: the roofline fraction
                with the stage kernel's traffic going to HBM, beside what a
                device-to-device copy of the same footprint reaches
  fp32          the same KH problem in the FP32 bulk form, and its relative L1
                difference per conserved variable from the FP64 result
  fluxes        the same KH problem with the HLL, Rusanov, HLLC and Marquina
                fluxes (plm cons / minmod)
  reconstructions  the same KH problem, Roe flux, with 'plm prim', 'plm eig',
                'plm cons with flux', 'plm athena', 'ppm', 'plm eig prim' and
                'plm eig prim ref'
  equations     shallow water (humps on a sloped, bumpy depth) and GLM-Maxwell
                (Gaussian pulses) on the same grid and scheme
  batched       512 independent 2D Euler four-quadrant grids (sides log-uniform
                in [64, 512]), FP32, RK2-TVD, 300 steps: one persistent launch
                per stage over all grids, hipGraph replay, autotuned launch
                shape; solves/s against the same grids advanced one after the
                other by the plain solver, and the FP32 error against FP64
  blast3d       3D Euler blast 256^3, van Leer flux limiter, RK3-TVD, FP32:
                throughput, mass / energy drift and L1 against the FP64 run
  serving       Poisson stream of mixed 1D Euler / 2D Euler / 2D GLM-Maxwell
                solves at 80% of the measured closed-loop capacity: solves/s,
                p50 / p99 latency, against one-at-a-time plain solves
  cpu_baseline  the CPU oracle (FP64 reference kernel sequence, OpenMP) on a
                bounded sample of the headline workload, with the host's CPU
                model, logical / affinity CPUs and cgroup quota

    python bench.py [--gpus 1] [--steps 500] [--warmup 20]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, 'hydro-cl-lua_amd'))

HBM_PEAK_GBS = 8000.0
N = 1024
NQ = 5                      # integrated reals per cell (Euler)


def algorithmic_bytes_per_cell_update(w, stages_rw):
    """bytes of state that must cross HBM per cell-update (one RK2-TVD step):
    stage 1 reads U^n and writes U^1, stage 2 reads U^1 and U^n and writes
    U^{n+1}: 5 q w"""
    return stages_rw * NQ * w


def rel_l1(got, ref):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    ax = tuple(range(ref.ndim - 1))
    den = np.abs(ref).sum(axis=ax)
    return np.abs(got - ref).sum(axis=ax) / np.where(den == 0, 1, den)


def interior(U, dim):
    sl = [slice(None)] * 4
    for i in range(dim):
        sl[2 - i] = slice(2, -2)
    return U[tuple(sl)]


def cpu_baseline(seconds=10.0):
    """The CPU oracle (FP64, reference kernel sequence, OpenMP) on the same
    KH 1024^2 input, timed for about `seconds` of steps."""
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle
    from hydro.init import euler as init_euler
    host = host_cpus()
    threads = host['threads_used']
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(dim=2, size=(N, N), mins=(-.5, -.5), maxs=(.5, .5), plm='plm cons',
                        slope_limiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4,
                        boundary='periodic')
    U0 = init_euler.to_ghosted(o.cfg, init_euler.kelvin_helmholtz(o.cfg, seed=0))
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    o.update(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.update(1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds and steps >= 2:
            break
    rate = N * N * steps / el / 1e9
    return {'value': rate, 'unit': 'Gcell-updates/s', 'cores': threads, 'kind': 'port',
            'sample': f'{steps} RK2-TVD steps of the same KH 1024^2 FP64 input ({el:.1f} s), '
                      f'oracle = reference kernel sequence, AoS, unfused, OpenMP on {threads} threads',
            'host': host}


def host_cpus():
    """The host's CPUs as the process sees them: the model, the logical CPUs
    of the machine, those this process may run on (affinity), the cgroup CPU
    quota, and the threads the CPU baseline uses: the CPU share of one GPU on a
    shared GPU host (16), or fewer when the process may use fewer."""
    model = ''
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                model = line.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    quota = None
    try:
        q, per = open('/sys/fs/cgroup/cpu.max').read().split()[:2]
        if q != 'max':
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    share = 16
    limit = min(x for x in (affinity, quota, share) if x)
    return {'cpu_model': model, 'logical_cpus': os.cpu_count(), 'affinity_cpus': affinity,
            'cgroup_cpu_quota': quota, 'threads_used': max(1, int(limit)),
            'note': 'threads = min(affinity, cgroup quota, 16): 16 host threads is the CPU share of one GPU '
                    'on the shared GPU host this runs on'}


def kh_solver(n, real, fast, rows=0, flux='roe', plm='plm cons'):
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(dim=2, gridSize=(n, n), mins=(-.5, -.5), maxs=(.5, .5), flux=flux, usePLM=plm,
                    slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4,
                    boundary='periodic', real=real, fastMath=fast)
    if rows:
        s.setLaunch(rows, 0)
    U0 = init_euler.to_ghosted(s.cfg, init_euler.kelvin_helmholtz(s.cfg, seed=0))
    s.resetState(U0.astype(s.dtype))
    return s


def timed_steps(s, warmup, steps):
    """wall time of exactly `steps` steps between two device synchronizations,
    and the HIP-event time of the same region on the solver stream"""
    s.update(max(warmup, 1))
    s.sync()
    t0 = time.perf_counter()
    kt = s.timeUpdates(steps, perKernel=False)     # events on the solver stream around update(steps)
    s.sync()
    return time.perf_counter() - t0, kt.total_ms


def equations_leg(n, steps, warmup, fast):
    """the other equations on the headline's shape and scheme (n^2, FP64, Roe +
    plm cons/minmod, RK2-TVD, CFL 0.4): shallow water (seeded humps on a
    sloped, bumpy depth, freeflow) and GLM-Maxwell (seeded Gaussian pulses,
    sigma 0, periodic). Algorithmic bytes per cell-update: 5 q w for the q
    integrated fields, as for Euler, plus the read-only fields read once per
    stage (the depth; GLM's rhoCharge, sigma, sqrt_1_eps, sqrt_1_mu)."""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler, shallow_water as isw, glm_maxwell as iglm
    legs = {
        'shallow_water': (4, 1, 'freeflow', lambda cfg: isw.humps(cfg, seed=0)),
        'glm_maxwell': (8, 4, 'periodic', lambda cfg: iglm.gaussian_pulses(cfg, seed=0, conductivity=0.)),
    }
    out = {}
    for eqn, (q, nstat, bc, init) in legs.items():
        s = HipFVSolver(dim=2, gridSize=(n, n), mins=(-.5, -.5), maxs=(.5, .5), eqn=eqn, flux='roe',
                        usePLM='plm cons', slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4,
                        boundary=bc, real='double', fastMath=fast)
        s.resetState(init_euler.to_ghosted(s.cfg, init(s.cfg)).astype(s.dtype))
        wall, ev = timed_steps(s, warmup, steps)
        avg = ev / (2 * steps)
        per_launch = (5 * q + 2 * nstat) * 8 / 2 * n * n
        U = interior(s.UBufObj.toCPU(), 2)[..., :q]
        out[eqn] = {'value': n * n * steps / wall / 1e9, 'unit': 'Gcell-updates/s', 'steps': steps,
                    'boundary': bc, 'avg_launch_ms': avg, 'algorithmic_bytes_per_launch': per_launch,
                    'roofline_frac': per_launch / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    'rows_per_wave': s.getLaunch()[0], 't_sim': s.t, 'finite': bool(np.isfinite(U).all())}
        del s
    return out


def committed_traffic(key):
    """HBM bytes per stage launch of a workload and where they come from: the
    rocprofv3 FETCH_SIZE and WRITE_SIZE passes committed under profiles/
    (counters cannot be read inside the timed run)"""
    tpath = os.path.join(REPO, 'profiles', 'stage_kernel_traffic.json')
    try:
        tj = json.load(open(tpath))
    except (OSError, ValueError):
        return None, None
    entry = tj.get(key)
    if isinstance(entry, dict):
        return entry.get('bytes_per_launch'), entry.get('source')
    return entry, tj.get('source')


def hbm_leg(n, steps, warmup, fast, no_fp32):
    """The headline workload on a grid whose state does not fit the 256 MB
    last-level cache (4096^2: U^n and U^1 are 1.07 GB in FP64), so that the
    stage kernel's traffic goes to HBM: Gcell-updates/s and the stage
    kernel's roofline fraction, FP64 and FP32 bulk"""
    res = {'workload': f'2D Euler Kelvin-Helmholtz {n}x{n} periodic, Roe + plm cons/minmod, RK2-TVD, CFL 0.4',
           'cells': n * n, 'steps': steps}
    for real in (('double',) if no_fp32 else ('double', 'float')):
        s = kh_solver(n, real, fast)
        wall, ev = timed_steps(s, warmup, steps)
        w = 8 if real == 'double' else 4
        avg = ev / (2 * steps)
        alg = algorithmic_bytes_per_cell_update(w, 5) / 2 * n * n
        res['f64' if real == 'double' else 'f32'] = {
            'value': n * n * steps / wall / 1e9, 'unit': 'Gcell-updates/s', 'avg_launch_ms': avg,
            'rows_per_wave': s.getLaunch()[0],
            'roofline': {'bound': 'hbm', 'achieved': alg / (avg * 1e-3) / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': alg / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 'algorithmic_bytes_per_launch': alg}}
        tr, src = committed_traffic(f'kh{n}_{real}')
        res['f64' if real == 'double' else 'f32']['roofline'].update({'traffic': tr, 'traffic_source': src})
        del s
    res['device_copy_ceiling'] = copy_ceiling(n)
    return res


def copy_ceiling(n, reps=10):
    """What a plain device-to-device copy reaches on this device at the same
    footprint (the 5-field FP64 state of one n x n grid): hipMemcpy D2D through
    the HIP runtime the library is linked to (libamdhip64.so.7), timed with
    events; GB/s of bytes read + written, and as a fraction of the 8 TB/s peak"""
    import ctypes
    hip = ctypes.CDLL('libamdhip64.so.7')
    vp = ctypes.c_void_p
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    hip.hipFree.argtypes = [vp]
    hip.hipMemcpyAsync.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, vp]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    nbytes = n * n * 5 * 8
    src, dst, e0, e1 = vp(), vp(), vp(), vp()
    if hip.hipMalloc(ctypes.byref(src), nbytes) or hip.hipMalloc(ctypes.byref(dst), nbytes):
        return None
    try:
        assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
        assert hip.hipMemcpyAsync(dst, src, nbytes, 3, None) == 0          # hipMemcpyDeviceToDevice, warm
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert hip.hipMemcpyAsync(dst, src, nbytes, 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        gbs = 2 * nbytes * reps / (ms.value * 1e-3) / 1e9
        return {'bytes_each_way': nbytes, 'GB/s': gbs, 'frac_of_peak': gbs / HBM_PEAK_GBS}
    finally:
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        hip.hipFree(src)
        hip.hipFree(dst)


def batched_leg(count, steps, seed=0, accuracy_grids=8):
    """solves/s of the batched form vs the plain solver one grid at a time"""
    from hydro import workloads as wl
    from hydro.solver.batch import HipBatch
    from hydro.solver.hipfvsolver import HipFVSolver
    ens = wl.ensemble_2d(count, seed=seed)
    cells = sum(kw['size'][0] * kw['size'][1] for kw, _ in ens)
    b = HipBatch([kw for kw, _ in ens])

    def load():
        for i, (_, U0) in enumerate(ens):
            b.upload(i, U0.astype(np.float32))
        b.constrainU()

    load()
    t0 = time.perf_counter()
    rows, wpb, _ = b.autotune(3)
    tune_s = time.perf_counter() - t0
    b.setLaunch(rows, wpb)
    load()
    b.update(3)                 # capture and prime the graph outside the timed region
    load()
    b.sync()
    t0 = time.perf_counter()
    b.update(steps, graph=True)
    b.sync()
    wall = time.perf_counter() - t0
    out = [b.download(i) for i in range(count)]

    # the same grids advanced one after the other by the plain solver (state resident)
    plains = []
    for kw, U0 in ens:
        s = HipFVSolver(dim=2, gridSize=kw['size'], mins=kw['mins'], maxs=kw['maxs'], usePLM=kw['plm'],
                        slopeLimiter=kw['slope_limiter'], integrator=kw['integrator'], cfl=kw['cfl'],
                        boundary=kw['boundary'], real='float', fastMath=True)
        s.resetState(U0.astype(np.float32))
        plains.append(s)
    for s in plains:
        s.sync()
    t0 = time.perf_counter()
    for s in plains:           # one solve after the other: each finishes before the next starts
        s.update(steps)
        s.sync()
    seq_wall = time.perf_counter() - t0
    same = all(np.array_equal(out[i], plains[i].UBufObj.toCPU()) for i in range(0, count, max(1, count // 16)))
    del plains

    # FP32 error against FP64 (IEEE) on a sample of the grids
    rng = np.random.default_rng(seed + 99)
    pick = sorted(rng.choice(count, size=min(accuracy_grids, count), replace=False).tolist())
    worst = 0.
    for i in pick:
        kw, U0 = ens[i]
        s = HipFVSolver(dim=2, gridSize=kw['size'], mins=kw['mins'], maxs=kw['maxs'], usePLM=kw['plm'],
                        slopeLimiter=kw['slope_limiter'], integrator=kw['integrator'], cfl=kw['cfl'],
                        boundary=kw['boundary'], real='double', fastMath=False)
        s.resetState(U0.astype(np.float32).astype(np.float64))
        s.update(steps)
        e = rel_l1(interior(out[i], 2)[..., :NQ], interior(s.UBufObj.toCPU(), 2)[..., :NQ])
        worst = max(worst, float(e.max()))
        del s
    return {
        'workload': f'{count} independent 2D Euler four-quadrant grids, sides log-uniform in [64, 512] '
                    f'(seed {seed}), freeflow, Roe + plm cons/minmod, RK2-TVD, CFL 0.4, FP32, {steps} steps',
        'grids': count, 'cells': cells, 'steps': steps,
        'solves_per_s': count / wall, 'gcell_updates_per_s': cells * steps / wall / 1e9,
        'sequential_plain_solves_per_s': count / seq_wall,
        'sequential_plain_gcell_updates_per_s': cells * steps / seq_wall / 1e9,
        'speedup_vs_sequential': seq_wall / wall,
        'launch': {'rows_per_wave': rows, 'waves_per_block': wpb, 'autotune_s': tune_s, 'graph': True},
        'same_bits_as_plain': bool(same),
        'fp32_rel_l1_vs_fp64_max': worst, 'fp32_rel_l1_grids': pick,
    }


def blast3d_leg(n, steps, seed=0):
    """3D Euler spherical blast, n^3 freeflow on [-1, 1]^3, Roe + van Leer
    flux limiter, RK3-TVD, CFL 0.3: FP32 bulk throughput, mass and energy
    drift (FP64 compensated totals), and the FP32 L1 difference from the same
    run in FP64"""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler

    def solver(real):
        s = HipFVSolver(dim=3, gridSize=(n, n, n), mins=(-1, -1, -1), maxs=(1, 1, 1), fluxLimiter='van Leer',
                        integrator='Runge-Kutta 3, TVD', cfl=.3, boundary='freeflow', real=real, fastMath=True)
        U0 = init_euler.to_ghosted(s.cfg, init_euler.sphere_blast(s.cfg, seed=seed))
        s.resetState(U0.astype(s.dtype))
        return s

    s = solver('float')
    tot0 = s.conservedTotals()
    s.sync()
    t0 = time.perf_counter()
    kt = s.timeUpdates(steps, perKernel=False)
    s.sync()
    wall = time.perf_counter() - t0
    tot1 = s.conservedTotals()
    U32 = interior(s.UBufObj.toCPU(), 3)[..., :NQ]
    t32, rows = s.t, s.getLaunch()[0]
    del s
    cells = n ** 3
    avg_ms = kt.total_ms / (3 * steps)
    alg = 8 * NQ * 4 / 3 * cells          # RK3-TVD: 8 q w per cell-update over 3 stage launches
    s64 = solver('double')
    s64.update(steps)
    U64 = interior(s64.UBufObj.toCPU(), 3)[..., :NQ]
    t64 = s64.t
    del s64
    e = rel_l1(U32, U64)
    return {
        'workload': f'3D Euler spherical blast {n}^3 freeflow on [-1,1]^3 (P 10 in r < .1 about a seeded centre, '
                    f'rho 1, P .1 outside), Roe + van Leer flux limiter, RK3-TVD, CFL 0.3, FP32 bulk, {steps} steps',
        'value': cells * steps / wall / 1e9, 'unit': 'Gcell-updates/s', 'rows_per_block_chunk': rows,
        'avg_launch_ms': avg_ms, 'roofline_frac': alg / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        'mass_drift': float((tot1[0] - tot0[0]) / tot0[0]), 'energy_drift': float((tot1[4] - tot0[4]) / tot0[4]),
        'rel_l1_vs_fp64': [float(x) for x in e], 't_fp32': t32, 't_fp64': t64,
        'note': 'drift includes the fluxes through the freeflow faces; each precision takes its own dt',
    }


def serving_leg(count, seed=0, steps_per_tick=20, budget=32 << 30, closed_only=False):
    from hydro import workloads as wl
    from hydro.serve import serve, serve_sequential
    reqs = wl.serving_mix(count, seed=seed)
    cu = sum(r.cell_updates() for r in reqs)
    lat_c, mk_c, _ = serve(reqs, budget, steps_per_tick, closed_loop=True)
    capacity = count / mk_c
    if closed_only:
        return {'closed_loop_solves_per_s': capacity, 'closed_loop_gcell_updates_per_s': cu / mk_c / 1e9,
                'closed_loop_makespan_s': mk_c, 'cell_updates': cu,
                'note': 'closed-loop served run only (--serve-closed-only)'}
    rate = .8 * capacity
    reqs_p = wl.serving_mix(count, seed=seed, rate=rate)
    lat, mk, _ = serve(reqs_p, budget, steps_per_tick)
    lat_s, mk_s, _ = serve_sequential(reqs, closed_loop=True)
    kinds = {k: sum(r.kind == k for r in reqs) for k in ('euler1d', 'euler2d', 'glm2d')}
    # each kind served alone, against its share of the sequential run (the
    # sequential run serves the requests in order, so a request's time is the
    # gap between its completion and the one before)
    done = np.sort(np.asarray(lat_s))
    took = np.diff(np.concatenate([[0.], done]))
    took_of = dict(zip(np.argsort(np.asarray(lat_s), kind='stable'), took))
    alone = {}
    for k in kinds:
        rk = [r for r in reqs if r.kind == k]
        if not rk:
            continue
        _, mk_k, _ = serve(rk, budget, steps_per_tick, closed_loop=True)
        seq_k = float(sum(took_of[i] for i, r in enumerate(reqs) if r.kind == k))
        cu_k = sum(r.cell_updates() for r in rk)
        alone[k] = {'served_makespan_s': mk_k, 'served_gcell_updates_per_s': cu_k / mk_k / 1e9,
                    'sequential_s': seq_k, 'speedup_vs_sequential': seq_k / mk_k}
    by_kind = {}
    for k in kinds:
        idx = [i for i, r in enumerate(reqs_p) if r.kind == k]
        if idx:
            by_kind[k] = {'p50_latency_s': float(np.percentile(lat[idx], 50)),
                          'p99_latency_s': float(np.percentile(lat[idx], 99)),
                          'cell_updates': int(sum(reqs_p[i].cell_updates() for i in idx))}
    return {
        'workload': f'{count} solve requests (seed {seed}): 1D Euler N in [256, 65536], 2D Euler side '
                    f'log-uniform [64, 1024], 2D GLM-Maxwell side log-uniform [128, 512] (sigma 0, eps = mu = 1); '
                    f'100-2000 steps each, FP32, Roe + plm cons/minmod, RK2-TVD',
        'mix': kinds, 'cell_updates': cu, 'poisson_by_kind': by_kind, 'closed_loop_alone_by_kind': alone,
        'closed_loop_solves_per_s': capacity, 'closed_loop_gcell_updates_per_s': cu / mk_c / 1e9,
        'poisson_rate_per_s': rate, 'solves_per_s': count / mk,
        'p50_latency_s': float(np.percentile(lat, 50)), 'p99_latency_s': float(np.percentile(lat, 99)),
        'sequential_plain_solves_per_s': count / mk_s,
        'sequential_plain_p50_latency_s': float(np.percentile(lat_s, 50)),
        'speedup_vs_sequential': mk_s / mk_c,
        'steps_per_tick': steps_per_tick, 'hbm_budget_bytes': budget,
        'latency': 'arrival to result on the host; sequential latencies are closed-loop (all requests at t=0)',
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--real', default='double', choices=['double', 'float'])
    ap.add_argument('--size', type=int, default=N)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-fp32', action='store_true', help='skip the FP32 KH leg')
    ap.add_argument('--no-fluxes', action='store_true', help='skip the KH legs of the other fluxes and reconstructions')
    ap.add_argument('--no-batch', action='store_true', help='skip the batched leg')
    ap.add_argument('--no-serve', action='store_true', help='skip the serving leg')
    ap.add_argument('--no-3d', action='store_true', help='skip the 3D blast leg')
    ap.add_argument('--no-hbm', action='store_true', help='skip the HBM-regime (4096^2) KH leg')
    ap.add_argument('--size-hbm', type=int, default=4096)
    ap.add_argument('--steps-hbm', type=int, default=50)
    ap.add_argument('--size3d', type=int, default=256)
    ap.add_argument('--steps3d', type=int, default=200)
    ap.add_argument('--exact-math', action='store_true',
                    help='IEEE division/sqrt (bitwise reference arithmetic) instead of the faithful fast sequences')
    ap.add_argument('--rows', type=int, default=0, help='rows per wave of the stage kernel (0: default)')
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--batch-grids', type=int, default=512)
    ap.add_argument('--batch-steps', type=int, default=300)
    ap.add_argument('--serve-requests', type=int, default=120)
    ap.add_argument('--serve-closed-only', action='store_true',
                    help='serving leg: the closed-loop served run only (for profiling)')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='first run one device copy of this many MiB (counter calibration for profiling)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py runs on one GPU only (got --gpus {args.gpus}); the grid path is not '
              'sharded across GPUs in this build', file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        # kernels of known traffic for the FETCH_SIZE / WRITE_SIZE counters: a
        # device copy of MIB MiB, and an elementwise negation of MIB MiB of
        # FP64 and of FP32 whose operands start one element past a 16-byte
        # boundary, so that each lane moves one 8- or 4-byte value per access,
        # as the stage kernels do
        import torch
        a = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device='cuda')
        b = torch.empty_like(a)
        b.copy_(a)
        for dt, w in ((torch.float64, 8), (torch.float32, 4)):
            n = (args.copy_calibration << 20) // w
            x = torch.ones(n + 1, dtype=dt, device='cuda')
            y = torch.empty_like(x)
            torch.neg(x[1:], out=y[1:])
        torch.cuda.synchronize()
        del a, b, x, y

    n = args.size
    fast = not args.exact_math
    s = kh_solver(n, args.real, fast, args.rows)
    wall, ev_ms = timed_steps(s, args.warmup, args.steps)
    cells = n * n
    value = cells * args.steps / wall / 1e9

    # stage kernel duration: the timed region holds only stage launches (two
    # per step, dt fused into the last), so the HIP-event time of the region on
    # the solver stream over the launch count is the average launch duration
    U_main = s.UBufObj.toCPU()
    t_main = s.t
    avg_ms = ev_ms / (2 * args.steps)
    w = 8 if args.real == 'double' else 4
    bytes_per_update = algorithmic_bytes_per_cell_update(w, 5)
    bytes_per_launch = bytes_per_update / 2 * cells      # two stage launches per step
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9

    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes of this same command (counters cannot be read inside the run)
    traffic, traffic_src = committed_traffic(f'kh{n}_{args.real}')

    out = {
        'metric': 'Gcell-updates/s (2D Euler Roe+PLM) + batched solves/s & p99 latency vs plain',
        'value': value,
        'unit': 'Gcell-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if args.real == 'double' else 'f32',
        'data': 'synthetic (seeded Kelvin-Helmholtz shear layer; seeded quadrant / Riemann / pulse states)',
        'config': {
            'workload': f'2D Euler Kelvin-Helmholtz {n}x{n} periodic, Roe + plm cons/minmod, '
                        f'RK2-TVD, CFL 0.4, {"FP64" if args.real == "double" else "FP32 bulk"}',
            'cells': cells, 'stage_launches_per_step': 2,
            'rows_per_wave': s.getLaunch()[0],
            'math': 'ieee' if args.exact_math else 'fast (faithfully rounded div/sqrt)',
        },
        'roofline': {
            'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
            'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic, 'traffic_source': traffic_src,
            'kernel': f'hydro::stage_kernel<{args.real}, Euler, 2, plm cons, minmod>',
            'avg_launch_ms': avg_ms,
            'algorithmic_bytes_per_launch': bytes_per_launch,
        },
        't_sim': t_main,
    }
    del s

    if not args.no_hbm:
        out['hbm_regime'] = hbm_leg(args.size_hbm, args.steps_hbm, args.warmup, fast, args.no_fp32)
    if not args.no_fp32 and args.real == 'double':
        s32 = kh_solver(n, 'float', True)
        wall32, ev32 = timed_steps(s32, args.warmup, args.steps)
        e = rel_l1(interior(s32.UBufObj.toCPU(), 2)[..., :NQ], interior(U_main, 2)[..., :NQ])
        out['fp32'] = {'value': cells * args.steps / wall32 / 1e9, 'unit': 'Gcell-updates/s',
                       'avg_launch_ms': ev32 / (2 * args.steps),
                       'roofline_frac': 5 * NQ * 4 / 2 * cells / (ev32 / (2 * args.steps) * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       'rows_per_wave': s32.getLaunch()[0], 't_sim': s32.t,
                       'rel_l1_vs_fp64': [float(x) for x in e],
                       'note': f'after {args.warmup + args.steps} steps; each form takes its own dt'}
        del s32
    if not args.no_fluxes and args.real == 'double':
        # the same workload with the other flux plugins (plm cons/minmod, no flux limiter)
        # and the Roe flux with the other reconstructions ('plm prim', 'plm eig')
        out['fluxes'], out['reconstructions'] = {}, {}
        variants = [('fluxes', f, f, 'plm cons') for f in ('hll', 'rusanov', 'euler-hllc', 'marquina')]
        variants += [('reconstructions', p, 'roe', p)
                     for p in ('plm prim', 'plm eig', 'plm cons with flux', 'plm athena', 'ppm', 'plm eig prim',
                               'plm eig prim ref')]
        for leg, key, flux, plm in variants:
            sf = kh_solver(n, args.real, fast, args.rows, flux=flux, plm=plm)
            steps_f = max(20, args.steps // 5)
            wallf, evf = timed_steps(sf, args.warmup, steps_f)
            avg_f = evf / (2 * steps_f)
            out[leg][key] = {'value': cells * steps_f / wallf / 1e9, 'unit': 'Gcell-updates/s',
                             'steps': steps_f, 'avg_launch_ms': avg_f,
                             'roofline_frac': 5 * NQ * 8 / 2 * cells / (avg_f * 1e-3) / 1e9 / HBM_PEAK_GBS}
            del sf
        out['equations'] = equations_leg(n, max(20, args.steps // 5), args.warmup, fast)
    if not args.no_batch:
        out['batched'] = batched_leg(args.batch_grids, args.batch_steps)
    if not args.no_3d:
        out['blast3d'] = blast3d_leg(args.size3d, args.steps3d)
    if not args.no_serve:
        out['serving'] = serving_leg(args.serve_requests, closed_only=args.serve_closed_only)
    if not args.no_cpu:
        out['cpu_baseline'] = cpu_baseline(args.cpu_seconds)
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out))


if __name__ == '__main__':
    main()
