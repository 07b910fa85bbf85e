"""GPU: the solve server (continuous batching with HBM-budget admission).
A request's result must be the same bits whether it is served in a mix of
other requests, under a budget that makes it wait, or alone by the plain
solver; oversized and forward-Euler requests are refused at submit; a state
that turns non-finite is flagged in the result status."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_mix(count, seed, real='float', fast=True):
    from hydro import workloads as wl
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = ('euler1d', 'euler2d', 'glm2d')[i % 3]
        steps = int(rng.integers(3, 40))
        if kind == 'euler1d':
            kw = wl.euler_1d_kw(int(rng.integers(100, 3000)), real, fast)
        elif kind == 'euler2d':
            kw = wl.quadrant_2d_kw(int(rng.integers(20, 90)), int(rng.integers(20, 90)), real, fast)
        else:
            s = int(rng.integers(24, 70))
            kw = wl.glm_2d_kw(s, real, fast)
        out.append(wl.Request(kind, kw, steps, int(rng.integers(0, 2 ** 31))))
    return out


@pytest.mark.parametrize('real,fast', [('float', True), ('double', False), ('double', True)])
def test_mix_equals_alone(real, fast):
    from hydro.serve import serve, serve_sequential
    reqs = _small_mix(12, seed=7, real=real, fast=fast)
    _, _, mixed = serve(reqs, budget_bytes=256 << 20, steps_per_tick=7, keep_results=True, closed_loop=True)
    _, _, alone = serve_sequential(reqs, keep_results=True, closed_loop=True)
    assert sorted(mixed) == list(range(len(reqs)))
    for i in range(len(reqs)):
        Um, tm, st = mixed[i]
        Ua, ta, _ = alone[i]
        assert st == 0, (i, reqs[i].kind)
        assert tm == ta, (i, reqs[i].kind, tm, ta)
        assert np.array_equal(Um, Ua), (i, reqs[i].kind, np.abs(Um - Ua).max())


def test_budget_admission_queues_and_matches():
    """a budget that holds only a few requests: the rest wait in the queue,
    memory is reused after retirement, and results do not change"""
    from hydro.serve import Server, serve_sequential
    from hydro import workloads as wl
    rng = np.random.default_rng(11)
    reqs = [wl.Request('euler2d', wl.quadrant_2d_kw(int(rng.integers(80, 111)), int(rng.integers(80, 111))),
                       int(rng.integers(3, 30)), int(rng.integers(0, 2 ** 31))) for _ in range(9)]
    states = [r.initial_state() for r in reqs]
    # device bytes of the largest possible request, read from a probe server
    big = wl.Request('euler2d', wl.quadrant_2d_kw(110, 110), 1, 0)
    probe = Server(budget_bytes=64 << 20, steps_per_tick=1)
    probe.submit(big.kw, big.initial_state(), 2)
    probe.tick()
    per_req = probe.info()['bytes_in_use']
    assert per_req > 0
    del probe
    srv = Server(budget_bytes=max(1 << 20, int(2.5 * per_req)), steps_per_tick=5)
    ids = {srv.submit(r.kw, U, r.steps): i for i, (r, U) in enumerate(zip(reqs, states))}
    seen_queue = False
    results = {}
    for _ in range(10000):
        info = srv.info()
        assert info['bytes_in_use'] <= info['budget']
        if info['queued'] + info['resident'] == 0:
            break
        srv.tick()
        info = srv.info()
        seen_queue = seen_queue or (info['queued'] > 0 and info['resident'] > 0)
        for rid in srv.poll():
            results[ids[rid]] = srv.result(rid)
    assert len(results) == len(reqs)
    assert seen_queue, 'the budget was expected to hold back some requests'
    assert srv.info()['bytes_in_use'] == 0
    _, _, alone = serve_sequential(reqs, keep_results=True, closed_loop=True)
    for i in range(len(reqs)):
        assert results[i][1] == alone[i][1]
        assert np.array_equal(results[i][0], alone[i][0]), i


def test_submit_rejections():
    from hydro.serve import Server
    from hydro import workloads as wl
    from hydro._abi import HydroError
    srv = Server(budget_bytes=1 << 20, steps_per_tick=4)
    big = wl.Request('euler2d', wl.quadrant_2d_kw(300, 300), 10, 1)
    with pytest.raises(HydroError, match='budget'):
        srv.submit(big.kw, big.initial_state(), 10)
    kw = dict(wl.euler_1d_kw(64), integrator='forward Euler')
    r = wl.Request('euler1d', kw, 5, 1)
    with pytest.raises(HydroError, match='forward Euler'):
        srv.submit(kw, r.initial_state(), 5)
    ok = wl.Request('euler1d', wl.euler_1d_kw(64), 5, 1)
    with pytest.raises(HydroError, match='state size'):
        srv.submit(ok.kw, ok.initial_state()[:-1], 5)
    with pytest.raises(HydroError):
        srv.result(12345)
    assert srv.info()['queued'] == 0


def test_non_finite_flagged():
    from hydro.serve import Server
    from hydro import workloads as wl
    srv = Server(budget_bytes=16 << 20, steps_per_tick=4)
    good = wl.Request('euler1d', wl.euler_1d_kw(128), 6, 3)
    U = good.initial_state()
    bad = U.copy()
    bad[0, 0, 40, 4] = np.nan
    ig = srv.submit(good.kw, U, 6)
    ib = srv.submit(good.kw, bad, 6)
    out = {}
    while len(out) < 2:
        srv.tick()
        for rid in srv.poll():
            out[rid] = srv.result(rid)
    assert out[ig][2] == 0 and np.isfinite(out[ig][0]).all()
    assert out[ib][2] == 1


def test_served_fp32_results_against_fp64_oracle():
    """every served FP32 result (1D Euler, 2D Euler and 2D GLM-Maxwell, as in
    the serving mix) against the FP64 oracle run from the same float inputs:
    <= 1e-5 relative L1 per conserved variable, and the same time"""
    import pyoracle
    from conftest import rel_l1
    from gpu_util import interior
    from hydro._cfg import make_cfg
    from hydro.serve import serve
    reqs = _small_mix(9, seed=21, real='float', fast=True)
    for r in reqs:
        r.steps = max(r.steps, 25)
    _, _, served = serve(reqs, budget_bytes=256 << 20, steps_per_tick=6, keep_results=True, closed_loop=True)
    for i, r in enumerate(reqs):
        U, t, status = served[i]
        assert status == 0
        o = pyoracle.Oracle(make_cfg(**dict(r.kw, precision='double', fast_math=False)))
        o.set_U(r.initial_state().astype(np.float64))
        o.boundary()
        o.update(r.steps)
        q = 5 if r.kind != 'glm2d' else 8
        dim = r.kw['dim']
        err = rel_l1(interior(dim, U)[..., :q].astype(np.float64), interior(dim, o.U)[..., :q])
        print(r.kind, r.kw['size'], r.steps, err.max())
        assert np.all(err <= 1e-5), (i, r.kind, err)
        assert abs(t - o.t) <= 1e-5 * o.t


def test_pinned_staging_limits_admission_and_copy_out():
    """requests whose packed states fill the server's pinned staging (256 MiB
    under a larger budget): admission waits for staging as it waits for
    memory, finished requests wait for staging to copy out, and results do
    not change; a request whose packed state alone exceeds the staging is
    refused at submit"""
    from hydro.serve import Server, serve_sequential
    from hydro import workloads as wl
    from hydro._abi import HydroError
    rng = np.random.default_rng(5)
    reqs = [wl.Request('euler2d', wl.quadrant_2d_kw(1300, 1280 + 10 * i, real='double', fast=True),
                       int(rng.integers(2, 6)), int(rng.integers(0, 2 ** 31))) for i in range(6)]
    srv = Server(budget_bytes=1 << 30, steps_per_tick=2)
    big = wl.Request('euler2d', wl.quadrant_2d_kw(2900, 2900, real='double', fast=True), 1, 0)
    with pytest.raises(HydroError, match='staging'):
        srv.submit(big.kw, np.zeros(big_shape(big.kw)), 1)
    ids = {srv.submit(r.kw, r.initial_state(), r.steps): i for i, r in enumerate(reqs)}
    results = {}
    queued_seen = False
    for _ in range(10000):
        info = srv.info()
        if info['queued'] + info['resident'] == 0:
            break
        srv.tick()
        queued_seen = queued_seen or srv.info()['queued'] > 0
        for rid in srv.poll():
            results[ids[rid]] = srv.result(rid)
    assert len(results) == len(reqs)
    assert queued_seen, 'the staging was expected to hold back some requests'
    _, _, alone = serve_sequential(reqs, keep_results=True, closed_loop=True)
    for i in range(len(reqs)):
        assert results[i][2] == 0
        assert results[i][1] == alone[i][1]
        assert np.array_equal(results[i][0], alone[i][0]), i


def big_shape(kw):
    """the ghosted cons_t shape of a request's state"""
    from hydro._cfg import make_cfg, cfg_ghost_shape
    return cfg_ghost_shape(make_cfg(**kw))
