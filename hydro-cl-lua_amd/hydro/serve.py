"""Solve server over the C ABI (hydro_server_*): continuous batching with
HBM-budget admission, plus the wall-clock serving loop that feeds it an
arrival stream and records per-request latency."""
import ctypes
import time

import numpy as np

from . import _abi
from ._cfg import make_cfg, cfg_ghost_shape


class Server:
    def __init__(self, budget_bytes, steps_per_tick=20):
        h = ctypes.c_void_p()
        _abi.check(_abi.lib().hydro_server_create(int(budget_bytes), int(steps_per_tick), ctypes.byref(h)))
        self.h = h
        self.shapes = {}
        self.dtypes = {}

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            _abi.lib().hydro_server_destroy(h)
            self.h = None

    def submit(self, kw, U, steps):
        cfg = make_cfg(**kw)
        dtype = np.float64 if cfg.precision == 0 else np.float32
        U = np.ascontiguousarray(np.asarray(U, dtype=dtype))
        rid = ctypes.c_longlong()
        _abi.check(_abi.lib().hydro_server_submit(self.h, ctypes.byref(cfg), U.ctypes.data, U.nbytes, int(steps),
                                                  ctypes.byref(rid)))
        self.shapes[rid.value] = cfg_ghost_shape(cfg)
        self.dtypes[rid.value] = dtype
        return rid.value

    def tick(self):
        n = ctypes.c_int()
        _abi.check(_abi.lib().hydro_server_tick(self.h, ctypes.byref(n)))
        return n.value

    def poll(self, max_ids=4096):
        ids = (ctypes.c_longlong * max_ids)()
        n = ctypes.c_int()
        _abi.check(_abi.lib().hydro_server_poll(self.h, ids, max_ids, ctypes.byref(n)))
        return list(ids[:n.value])

    def result(self, rid):
        if rid not in self.shapes:    # not submitted here: let the library report it
            _abi.check(_abi.lib().hydro_server_result(self.h, int(rid), None, 0, None, None))
        out = np.empty(self.shapes[rid], dtype=self.dtypes[rid])
        t, st = ctypes.c_double(), ctypes.c_int()
        _abi.check(_abi.lib().hydro_server_result(self.h, int(rid), out.ctypes.data, out.nbytes, ctypes.byref(t),
                                                  ctypes.byref(st)))
        del self.shapes[rid], self.dtypes[rid]
        return out, t.value, st.value

    def info(self):
        i = _abi.HydroServerInfo()
        _abi.check(_abi.lib().hydro_server_info_get(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _abi.HydroServerInfo._fields_}

    def setLaunch(self, rows=0, wavesPerBlock=0):
        _abi.check(_abi.lib().hydro_server_set_launch(self.h, int(rows), int(wavesPerBlock)))


def serve(requests, budget_bytes, steps_per_tick=20, keep_results=False, closed_loop=False, launch=None):
    """Runs the arrival stream against a Server in wall-clock time. Each
    request is submitted when its arrival time comes (all at once when
    closed_loop); returns per-request latency (arrival -> result on the host),
    the makespan, and optionally the results by request index. launch =
    (rows per wave, waves per block) sets the pools' launch shape."""
    srv = Server(budget_bytes, steps_per_tick)
    if launch:
        srv.setLaunch(*launch)
    states = [r.initial_state() for r in requests]
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    lat = [None] * len(requests)
    results = {}
    ids = {}
    nxt = 0
    t0 = time.perf_counter()
    done = 0
    while done < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(order) and (closed_loop or requests[order[nxt]].arrival <= now):
            i = order[nxt]
            ids[srv.submit(requests[i].kw, states[i], requests[i].steps)] = i
            nxt += 1
        info = srv.info()
        if info['queued'] + info['resident'] == 0:
            if nxt < len(order):
                time.sleep(max(0., requests[order[nxt]].arrival - (time.perf_counter() - t0)))
            continue
        srv.tick()
        for rid in srv.poll():
            U, t, status = srv.result(rid)
            i = ids.pop(rid)
            arrival = 0. if closed_loop else requests[i].arrival
            lat[i] = time.perf_counter() - t0 - arrival
            if keep_results:
                results[i] = (U, t, status)
            done += 1
    makespan = time.perf_counter() - t0
    return np.array(lat), makespan, results


def serve_sequential(requests, keep_results=False, closed_loop=False):
    """The same stream served one request at a time by the plain solver: each
    request is created, uploaded, advanced with plain launches and read back
    before the next starts."""
    from .solver.hipfvsolver import HipFVSolver
    states = [r.initial_state() for r in requests]
    order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
    lat = [None] * len(requests)
    results = {}
    t0 = time.perf_counter()
    for i in order:
        r = requests[i]
        if not closed_loop:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        kw = r.kw
        s = HipFVSolver(dim=kw['dim'], gridSize=kw['size'], mins=kw.get('mins'), maxs=kw.get('maxs'),
                        eqn=kw.get('eqn', 'euler'), usePLM=kw.get('plm'), slopeLimiter=kw.get('slope_limiter', 'minmod'),
                        fluxLimiter=kw.get('flux_limiter', 'donor cell'), integrator=kw['integrator'], cfl=kw['cfl'],
                        boundary=kw['boundary'], real=('double' if kw.get('precision', 'double') == 'double' else 'float'),
                        fastMath=kw.get('fast_math', False))
        s.UBufObj.fromCPU(states[i])
        s.update(r.steps)
        U = s.UBufObj.toCPU()
        t = s.t
        del s
        lat[i] = time.perf_counter() - t0 - (0. if closed_loop else r.arrival)
        if keep_results:
            results[i] = (U, t, 0)
    makespan = time.perf_counter() - t0
    return np.array(lat), makespan, results
