"""GPU: the batched form. Each grid of a batch must give the same bits as the
same grid solved alone by the plain solver (same kernels, no cross-grid
arithmetic), and match the oracle to the FP64 / FP32 bars."""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

SIZES = [(64, 64), (100, 37), (130, 200), (61, 300), (250, 90), (33, 33)]


def _cfg(n, real='double', fast=False, integ='Runge-Kutta 2, TVD', bc='freeflow', cfl=.4):
    return dict(dim=2, size=n, mins=(-1, -1), maxs=(1, 1), plm='plm cons', slope_limiter='minmod',
                integrator=integ, cfl=cfl, boundary=bc, precision=real, fast_math=fast)


def _init(cfg, seed):
    from hydro.init import euler as ie
    rng = np.random.default_rng(seed)
    return ie.to_ghosted(cfg, ie.quadrant(cfg, ie.random_quadrant_states(rng)))


def _plain(kw, U0, steps):
    from hydro.solver.hipfvsolver import HipFVSolver
    s = HipFVSolver(dim=2, gridSize=kw['size'], mins=kw['mins'], maxs=kw['maxs'], usePLM='plm cons',
                    slopeLimiter='minmod', integrator=kw['integrator'], cfl=kw['cfl'], boundary=kw['boundary'],
                    real=kw['precision'], fastMath=kw['fast_math'])
    s.resetState(U0.astype(s.dtype))
    s.update(steps)
    return s.UBufObj.toCPU(), s.t


@pytest.mark.parametrize('real,fast,graph', [('double', False, True), ('double', True, False),
                                             ('float', True, True)])
def test_batch_equals_plain(real, fast, graph):
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [_cfg(n, real, fast) for n in SIZES]
    b = HipBatch(kws)
    U0s = [_init(make_cfg(**kw), 10 + i) for i, kw in enumerate(kws)]
    for i, U0 in enumerate(U0s):
        b.upload(i, U0.astype(b.dtype))
    b.constrainU()
    steps = 14
    b.update(steps, graph=graph)
    for i, kw in enumerate(kws):
        ref, t = _plain(kw, U0s[i], steps)
        got = b.download(i)
        assert np.array_equal(got, ref), (i, np.abs(got - ref).max())
        assert b.time(i)[0] == t


def test_batch_vs_oracle_and_inactive_grid():
    import pyoracle
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [_cfg(n) for n in SIZES[:4]]
    b = HipBatch(kws)
    U0s = [_init(make_cfg(**kw), 20 + i) for i, kw in enumerate(kws)]
    for i, U0 in enumerate(U0s):
        b.upload(i, U0)
    b.constrainU()
    b.setActive(2, False)       # grid 2 stays put
    b.update(9)
    for i, kw in enumerate(kws):
        o = pyoracle.Oracle(make_cfg(**kw))
        o.set_U(U0s[i])
        o.boundary()
        o.constrain_U()
        if i != 2:
            o.update(9)
        got = o.interior(b.download(i))[..., :5]
        err = rel_l1(got, o.interior()[..., :5])
        assert np.all(err <= 1e-12), (i, err)


def test_batch_fp32_vs_fp64_oracle():
    import pyoracle
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [_cfg(n, 'float', True) for n in SIZES[:3]]
    b = HipBatch(kws)
    U0s = [_init(make_cfg(**kw), 30 + i) for i, kw in enumerate(kws)]
    for i, U0 in enumerate(U0s):
        b.upload(i, U0.astype(np.float32))
    b.constrainU()
    b.update(30)
    for i, kw in enumerate(kws):
        kw64 = dict(kw, precision='double', fast_math=False)
        o = pyoracle.Oracle(make_cfg(**kw64))
        o.set_U(U0s[i].astype(np.float32).astype(np.float64))
        o.boundary()
        o.constrain_U()
        o.update(30)
        err = rel_l1(o.interior(b.download(i))[..., :5], o.interior()[..., :5])
        assert np.all(err <= 1e-5), (i, err)


def test_batch_autotune_keeps_results():
    """the launch shape changes the decomposition, never the numbers"""
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [_cfg(n, 'float', True) for n in SIZES]
    U0s = [_init(make_cfg(**kw), 40 + i) for i, kw in enumerate(kws)]
    outs = []
    for shape in [(2, 1), (8, 4), None]:
        b = HipBatch(kws)
        for i, U0 in enumerate(U0s):
            b.upload(i, U0.astype(np.float32))
        b.constrainU()
        if shape:
            b.setLaunch(*shape)
        else:
            # autotune on a scratch copy, then run the real state
            b2 = HipBatch(kws)
            for i, U0 in enumerate(U0s):
                b2.upload(i, U0.astype(np.float32))
            b2.constrainU()
            r, w, ms = b2.autotune(2)
            assert r > 0 and w in (1, 2, 4) and ms > 0
            b.setLaunch(r, w)
        b.update(7)
        outs.append([b.download(i) for i in range(len(kws))])
    for o in outs[1:]:
        for a, c in zip(outs[0], o):
            assert np.array_equal(a, c)


@pytest.mark.parametrize('graph', [False, True])
def test_batch_reactivated_grid_matches_oracle(graph):
    """grids switched off for one, two and three steps (every offset of the
    three-slot dt rotation) and back on continue from their own state with
    their own dt: each grid's result equals the oracle's after the steps it
    took part in"""
    import pyoracle
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [_cfg(n) for n in SIZES[:4]]
    b = HipBatch(kws)
    U0s = [_init(make_cfg(**kw), 50 + i) for i, kw in enumerate(kws)]
    for i, U0 in enumerate(U0s):
        b.upload(i, U0)
    b.constrainU()
    taken = [0] * len(kws)
    plan = [(2, ()), (1, (1, 2, 3)), (1, (2, 3)), (1, (3,)), (4, ()), (3, (0,)), (5, ())]
    for steps, off in plan:
        for i in range(len(kws)):
            b.setActive(i, i not in off)
        b.update(steps, graph=graph)
        for i in range(len(kws)):
            if i not in off:
                taken[i] += steps
    for i, kw in enumerate(kws):
        o = pyoracle.Oracle(make_cfg(**kw))
        o.set_U(U0s[i])
        o.boundary()
        o.constrain_U()
        o.update(taken[i])
        got = o.interior(b.download(i))[..., :5]
        err = rel_l1(got, o.interior()[..., :5])
        assert np.all(err <= 1e-12), (i, taken[i], err)
        assert b.time(i)[0] == pytest.approx(o.t, rel=1e-14)


@pytest.mark.parametrize('real,fast,tol', [('double', True, 1e-12), ('float', True, 1e-5)])
def test_batch_glm_ppm_vs_oracle(real, fast, tol):
    """GLM-Maxwell with 'ppm' in the batched kernels"""
    import pyoracle
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    from hydro.init import glm_maxwell
    from hydro.init import euler as ie
    kws = [dict(dim=2, size=n, mins=(-1, -1), maxs=(1, 1), eqn='glm_maxwell', plm='ppm', slope_limiter='minmod',
                integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', precision=real, fast_math=fast)
           for n in [(48, 48), (70, 33), (33, 80)]]
    b = HipBatch(kws)
    U0s = []
    for i, kw in enumerate(kws):
        cfg = make_cfg(**kw)
        U0 = ie.to_ghosted(cfg, glm_maxwell.gaussian_pulses(cfg, seed=60 + i, conductivity=.5)).astype(b.dtype)
        U0s.append(U0)
        b.upload(i, U0)
    b.update(20)
    for i, kw in enumerate(kws):
        o = pyoracle.Oracle(make_cfg(**dict(kw, precision='double', fast_math=False)))
        o.set_U(U0s[i].astype(np.float64))
        o.boundary()
        o.update(20)
        err = rel_l1(o.interior(b.download(i))[..., :8].astype(np.float64), o.interior()[..., :8])
        assert np.all(err <= tol), (i, err)
