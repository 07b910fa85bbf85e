"""GPU parity of the shallow water equations (hydro/eqn/shallow-water.cl,
depthInCell = false) through the C ABI against the oracle: a 1D dam break
and 2D humps on a sloped, bumpy depth (the depth source active), every
boundary kind, with and without 'plm cons', with a flux limiter, several
integrators; EXACT FP64 <= 1e-12 relative L1 per integrated field (most
bitwise), FAST FP64 <= 1e-12, FP32 (inputs rounded to float on both sides)
<= 1e-5; the batched form equals the plain solver bit for bit; the total
height is conserved on a periodic grid."""
import numpy as np
import pytest

from conftest import rel_l1
from gpu_util import interior

pytestmark = pytest.mark.gpu

G = .8
MODES = {'exact': dict(real='double', fastMath=False), 'fast': dict(real='double', fastMath=True),
         'fp32': dict(real='float', fastMath=True)}
TOL = {'exact': 1e-12, 'fast': 1e-12, 'fp32': 1e-5}


def _pair(init, **kw):
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(eqn='shallow_water', eqnArgs={'gravity': G}, **kw)
    U0 = init_euler.to_ghosted(s.cfg, init(s.cfg)).astype(s.dtype)
    s.resetState(U0)
    okw = {k: v for k, v in kw.items() if k not in ('real', 'fastMath')}
    ocfg = pyoracle.make_cfg(dim=okw['dim'], size=okw['gridSize'], mins=okw.get('mins'), maxs=okw.get('maxs'),
                             eqn='shallow_water', plm=okw.get('usePLM'), slope_limiter=okw.get('slopeLimiter', 'minmod'),
                             flux_limiter=okw.get('fluxLimiter', 'donor cell'), integrator=okw['integrator'],
                             cfl=okw['cfl'], boundary=okw['boundary'], gravity=G)
    o = pyoracle.Oracle(ocfg)
    o.set_U(U0.astype(np.float64))
    o.boundary()
    return s, o


def _dam(cfg):
    from hydro.init import shallow_water as isw
    return isw.dam_break(cfg, 1., .4)


def _dam_y(cfg):
    from hydro.init import shallow_water as isw
    return isw.dam_break(cfg, 1., .4, axis=1)


def _humps(seed):
    from hydro.init import shallow_water as isw
    return lambda cfg: isw.humps(cfg, seed=seed)


CASES = [
    ('dam-1d-plm-rk2', _dam, dict(dim=1, gridSize=300, usePLM='plm cons', slopeLimiter='minmod',
                                   integrator='Runge-Kutta 2, TVD', cfl=.5, boundary='freeflow'), 40),
    ('dam-1d-superbee-fe', _dam, dict(dim=1, gridSize=257, fluxLimiter='superbee', integrator='forward Euler',
                                       cfl=.5, boundary='mirror'), 40),
    ('humps-2d-plm-rk2-periodic', _humps(1), dict(dim=2, gridSize=(96, 80), usePLM='plm cons', slopeLimiter='minmod',
                                                  integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic'), 30),
    ('humps-2d-none-rk3-mirror', _humps(2), dict(dim=2, gridSize=(70, 90), integrator='Runge-Kutta 3, TVD', cfl=.4,
                                                 boundary='mirror'), 25),
    ('humps-2d-vanleer-rk2-freeflow', _humps(3), dict(dim=2, gridSize=(65, 47), fluxLimiter='van Leer',
                                                      integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow'), 25),
    ('dam-2d-y-plm-rk4', _dam_y,
     dict(dim=2, gridSize=(40, 130), mins=(-1, -1), maxs=(1, 1), usePLM='plm cons', slopeLimiter='superbee',
          integrator='Runge-Kutta 4', cfl=.4, boundary='freeflow'), 30),
]


@pytest.mark.parametrize('mode', list(MODES))
@pytest.mark.parametrize('name,init,kw,steps', CASES, ids=[c[0] for c in CASES])
def test_shallow_water_vs_oracle(name, init, kw, steps, mode):
    s, o = _pair(init, **kw, **MODES[mode])
    s.update(steps)
    o.update(steps)
    dim = kw['dim']
    got = interior(dim, s.UBufObj.toCPU())[..., :5].astype(np.float64)
    ref = interior(dim, o.U)[..., :5]
    assert np.all(np.isfinite(got))
    err = rel_l1(got[..., :4], ref[..., :4])
    print(f'{name} {mode}: rel L1 {err}, bitwise {np.array_equal(got, ref)}')
    assert np.all(err <= TOL[mode]), err
    assert np.array_equal(got[..., 4], ref[..., 4])          # the depth is carried, never integrated
    assert abs(s.t - o.t) <= max(1e-14, TOL[mode]) * o.t


def test_shallow_water_height_conserved_periodic():
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(dim=2, gridSize=(128, 96), eqn='shallow_water', eqnArgs={'gravity': G}, usePLM='plm cons',
                    slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic')
    s.resetState(init_euler.to_ghosted(s.cfg, _humps(5)(s.cfg)))
    t0 = s.conservedTotals()
    s.update(100)
    t1 = s.conservedTotals()
    assert abs(t1[0] - t0[0]) <= 1e-13 * abs(t0[0])


def test_shallow_water_batch_equals_plain():
    from hydro.solver.batch import HipBatch
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro._cfg import make_cfg
    from hydro.init import euler as init_euler
    sizes = [(64, 50), (90, 33), (40, 120)]
    kws = [dict(dim=2, size=n, mins=(-1, -1), maxs=(1, 1), eqn='shallow_water', plm='plm cons',
                slope_limiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow',
                precision='double', fast_math=True, gravity=G) for n in sizes]
    b = HipBatch(kws)
    U0s = [init_euler.to_ghosted(make_cfg(**kw), _humps(20 + i)(make_cfg(**kw))) for i, kw in enumerate(kws)]
    for i, U0 in enumerate(U0s):
        b.upload(i, U0)
    b.update(12, graph=True)
    for i, kw in enumerate(kws):
        s = HipFVSolver(dim=2, gridSize=kw['size'], mins=kw['mins'], maxs=kw['maxs'], eqn='shallow_water',
                        eqnArgs={'gravity': G}, usePLM='plm cons', slopeLimiter='minmod',
                        integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow', fastMath=True)
        s.resetState(U0s[i])
        s.update(12)
        assert np.array_equal(b.download(i), s.UBufObj.toCPU()), i
        assert b.time(i)[0] == s.t


def test_shallow_water_rejections():
    from hydro._abi import HydroError
    from hydro.solver.hipfvsolver import HipFVSolver
    with pytest.raises(HydroError, match='plm cons'):
        HipFVSolver(dim=2, gridSize=(16, 16), eqn='shallow_water', usePLM='plm prim')
    with pytest.raises(HydroError, match='Euler'):
        HipFVSolver(dim=2, gridSize=(16, 16), eqn='shallow_water', flux='hll')
    with pytest.raises(HydroError, match='3D'):
        HipFVSolver(dim=3, gridSize=(8, 8, 8), eqn='shallow_water')
