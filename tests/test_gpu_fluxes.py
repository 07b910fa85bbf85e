"""GPU: the HLL, Rusanov, HLLC and Marquina fluxes (hydro/flux/hll.cl,
rusanov.cl, euler-hllc.cl, marquina.cl) through the C ABI against the CPU
oracle, to the same bars as the Roe flux: FP64 within
1e-12 relative L1 per conserved variable (IEEE and fast math), FP32 bulk
within 1e-5 of the FP64 oracle; 1D, 2D (several strips, every boundary kind)
and 3D; the batched form equal to the plain one; the flux limiter ignored as
the reference does (only Roe uses one, fvsolver.lua:63)."""
import numpy as np
import pytest

from conftest import rel_l1
from test_gpu_parity import _pair, _check, _quad, _sod, TOL64, TOL32

pytestmark = pytest.mark.gpu

FLUXES = ['hll', 'rusanov', 'euler-hllc', 'marquina']


@pytest.mark.parametrize('fast', [False, True])
@pytest.mark.parametrize('flux', FLUXES)
def test_sod_1d(flux, fast):
    left = [1., 0, 0, 0, 1. / .4, 0]
    right = [.125, 0, 0, 0, .1 / .4, 0]
    kw = dict(dim=1, gridSize=256, flux=flux, integrator='Runge-Kutta 2, TVD', cfl=.5,
              boundary={'xmin': 'fixed', 'xmax': 'fixed'}, boundaryFixed={'xmin': left, 'xmax': right})
    s, o = _pair(_sod, 60, fastMath=fast, **kw)
    _check(s, o, 1, TOL64, 60, f'sod-{flux}-{fast}')


@pytest.mark.parametrize('fast', [False, True])
@pytest.mark.parametrize('plm', [False, True])
@pytest.mark.parametrize('flux', FLUXES)
def test_quadrant_2d(flux, plm, fast):
    rec = dict(usePLM='plm cons', slopeLimiter='van Leer') if plm else {}
    kw = dict(dim=2, gridSize=(150, 45), flux=flux, integrator='Runge-Kutta 3, TVD', cfl=.3,
              boundary={'xmin': 'fixed', 'xmax': 'mirror', 'ymin': 'periodic', 'ymax': 'periodic'},
              boundaryFixed={'xmin': [1., .1, 0, 0, 2.5, 0]}, **rec)
    s, o = _pair(_quad(7), 8, fastMath=fast, **kw)
    _check(s, o, 2, TOL64, 8, f'quad-{flux}-{plm}-{fast}')


@pytest.mark.parametrize('flux', FLUXES)
def test_fp32_vs_fp64_oracle(flux):
    from test_gpu_parity import _kh
    kw = dict(dim=2, gridSize=(128, 128), mins=(-.5, -.5), maxs=(.5, .5), flux=flux, usePLM='plm cons',
              slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic')
    s, o = _pair(_kh(3), 20, real='float', fastMath=True, **kw)
    _check(s, o, 2, TOL32, 20, f'kh32-{flux}')


@pytest.mark.parametrize('plm', [False, True])
@pytest.mark.parametrize('flux', FLUXES)
def test_3d(flux, plm):
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    from test_gpu_3d import _octants
    box = dict(mins=(-1, -1, -1), maxs=(1, 1, 1))
    rec = dict(usePLM='plm cons', slopeLimiter='minmod') if plm else {}
    s = HipFVSolver(dim=3, gridSize=(70, 13, 9), flux=flux, integrator='Runge-Kutta 2, TVD', cfl=.3,
                    boundary='freeflow', **rec, **box)
    U0 = ie.to_ghosted(s.cfg, _octants(4)(s.cfg))
    s.resetState(U0)
    o = pyoracle.Oracle(pyoracle.make_cfg(dim=3, size=(70, 13, 9), flux=flux, integrator='Runge-Kutta 2, TVD',
                                          cfl=.3, boundary='freeflow', plm='plm cons' if plm else None,
                                          slope_limiter='minmod', precision='double', **box))
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    s.update(5)
    o.update(5)
    err = rel_l1(o.interior(s.UBufObj.toCPU())[..., :5], o.interior()[..., :5])
    print(f'3d-{flux}-{plm}: rel L1 {err}')
    assert np.all(err <= TOL64), err


@pytest.mark.parametrize('flux', FLUXES)
def test_flux_limiter_ignored(flux):
    """a flux limiter setting changes nothing for these fluxes"""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    outs = []
    for lim in ('donor cell', 'superbee'):
        s = HipFVSolver(dim=2, gridSize=(70, 40), flux=flux, fluxLimiter=lim, integrator='Runge-Kutta 2, TVD',
                        cfl=.3, boundary='freeflow')
        rng = np.random.default_rng(8)
        s.resetState(ie.to_ghosted(s.cfg, ie.quadrant(s.cfg, ie.random_quadrant_states(rng))))
        s.update(6)
        outs.append(s.UBufObj.toCPU())
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize('flux', FLUXES)
def test_batch_equals_plain(flux):
    from hydro.solver.batch import HipBatch
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro._cfg import make_cfg
    from hydro.init import euler as ie
    sizes = [(64, 64), (100, 37), (130, 70)]
    kws = [dict(dim=2, size=n, mins=(-1, -1), maxs=(1, 1), flux=flux, plm='plm cons', slope_limiter='minmod',
                integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow', precision='double', fast_math=True)
           for n in sizes]
    b = HipBatch(kws)
    U0s = []
    for i, kw in enumerate(kws):
        rng = np.random.default_rng(20 + i)
        cfg = make_cfg(**kw)
        U0s.append(ie.to_ghosted(cfg, ie.quadrant(cfg, ie.random_quadrant_states(rng))))
        b.upload(i, U0s[i])
    b.constrainU()
    b.update(9)
    for i, kw in enumerate(kws):
        s = HipFVSolver(dim=2, gridSize=kw['size'], mins=kw['mins'], maxs=kw['maxs'], flux=flux,
                        usePLM='plm cons', slopeLimiter='minmod', integrator=kw['integrator'], cfl=kw['cfl'],
                        boundary='freeflow', fastMath=True)
        s.resetState(U0s[i])
        s.update(9)
        assert np.array_equal(b.download(i), s.UBufObj.toCPU()), i


def test_glm_rejects_hll():
    from hydro._abi import HydroError
    from hydro.solver.hipfvsolver import HipFVSolver
    with pytest.raises(HydroError):
        HipFVSolver(dim=2, gridSize=(32, 32), eqn='glm_maxwell', flux='hll', integrator='Runge-Kutta 2, TVD',
                    cfl=.4, boundary='periodic')


@pytest.mark.parametrize('fast', [False, True])
@pytest.mark.parametrize('rec', ['plm', 'flux limiter'])
def test_roe_entropy_fix(rec, fast):
    """Roe with Harten's entropy fix (roe.cl:83-106, args.useEntropyFluxFix) on
    transonic quadrant states: the oracle's bar, and results that differ from
    the unfixed flux"""
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    kw = dict(usePLM='plm cons', slopeLimiter='minmod') if rec == 'plm' else dict(fluxLimiter='superbee')
    okw = dict(plm='plm cons', slope_limiter='minmod') if rec == 'plm' else dict(flux_limiter='superbee')
    outs = []
    for fix in (True, False):
        s = HipFVSolver(dim=2, gridSize=(130, 40), integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='freeflow',
                        fastMath=fast, fluxArgs={'useEntropyFluxFix': fix, 'entropyFluxFixLimitingFactor': .3}, **kw)
        rng = np.random.default_rng(11)
        states = ie.random_quadrant_states(rng)
        U0 = ie.to_ghosted(s.cfg, ie.quadrant(s.cfg, states))
        s.resetState(U0)
        s.update(6)
        outs.append(s.UBufObj.toCPU())
        if fix:
            o = pyoracle.Oracle(pyoracle.make_cfg(dim=2, size=(130, 40), integrator='Runge-Kutta 2, TVD', cfl=.3,
                                                  boundary='freeflow', precision='double', use_entropy_flux_fix=True,
                                                  entropy_flux_fix_limiting_factor=.3, **okw))
            o.set_U(U0)
            o.boundary()
            o.constrain_U()
            o.update(6)
            err = rel_l1(o.interior(outs[0])[..., :5], o.interior()[..., :5])
            print(f'roe-fix-{rec}-{fast}: rel L1 {err}')
            assert np.all(err <= TOL64), err
    assert not np.array_equal(outs[0], outs[1])


PLM_KINDS = ['plm prim', 'plm eig', 'plm cons with flux', 'plm athena', 'ppm', 'plm eig prim', 'plm eig prim ref']


@pytest.mark.parametrize('real,fast', [('double', False), ('double', True), ('float', True)])
@pytest.mark.parametrize('plm', PLM_KINDS)
def test_plm_kinds_2d(plm, real, fast):
    """'plm prim' (plm.cl:191-254): slopes of the primitives; 'plm eig'
    (plm.cl:256-432): limited characteristic slopes with half a step of the
    edge fluxes; 'plm cons with flux' (plm.cl:95-188): cons slopes with the
    same half step; 'plm athena' (plm.cl:782-879): Athena's characteristic
    slope of the primitives with clamped edges; 'ppm' (plm.cl:882-944): the
    kappa = 1/3 edges; 'plm eig prim' / 'ref' (plm.cl:536-780): characteristic
    slopes of the primitives; all with the Roe flux"""
    kw = dict(dim=2, gridSize=(150, 45), usePLM=plm, slopeLimiter='van Leer', integrator='Runge-Kutta 2, TVD',
              cfl=.3, boundary={'xmin': 'mirror', 'xmax': 'freeflow', 'ymin': 'periodic', 'ymax': 'periodic'})
    s, o = _pair(_quad(12), 8, real=real, fastMath=fast, **kw)
    _check(s, o, 2, TOL64 if real == 'double' else TOL32, 8, f'{plm}-{real}-{fast}')


@pytest.mark.parametrize('real', ['double', 'float'])
@pytest.mark.parametrize('plm', PLM_KINDS)
def test_plm_kinds_2d_fast_minmod(plm, real):
    """the same reconstructions with minmod in the FAST kernels, which take its
    compile-time form (phi(r) dU without forming r) instead of the runtime
    limiter switch (seed 12 as above: on some seeded quadrant states 'ppm''s
    kappa = 1/3 edges reach negative pressures and the reference arithmetic
    itself turns non-finite)"""
    kw = dict(dim=2, gridSize=(150, 45), usePLM=plm, slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD',
              cfl=.3, boundary={'xmin': 'mirror', 'xmax': 'freeflow', 'ymin': 'periodic', 'ymax': 'periodic'})
    s, o = _pair(_quad(12), 8, real=real, fastMath=True, **kw)
    _check(s, o, 2, TOL64 if real == 'double' else TOL32, 8, f'{plm}-{real}-fast-minmod')


@pytest.mark.parametrize('plm', PLM_KINDS)
def test_plm_kinds_1d_3d_batch(plm):
    kw = dict(dim=1, gridSize=256, usePLM=plm, slopeLimiter='minmod', integrator='Runge-Kutta 3, TVD',
              cfl=.5, boundary='freeflow')
    s, o = _pair(_sod, 30, **kw)
    _check(s, o, 1, TOL64, 30, f'{plm}-1d')
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    from test_gpu_3d import _octants
    box = dict(mins=(-1, -1, -1), maxs=(1, 1, 1))
    s = HipFVSolver(dim=3, gridSize=(70, 13, 9), usePLM=plm, slopeLimiter='minmod',
                    integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='mirror', **box)
    U0 = ie.to_ghosted(s.cfg, _octants(6)(s.cfg))
    s.resetState(U0)
    o = pyoracle.Oracle(pyoracle.make_cfg(dim=3, size=(70, 13, 9), plm=plm, slope_limiter='minmod',
                                          integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='mirror',
                                          precision='double', **box))
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    s.update(4)
    o.update(4)
    err = rel_l1(o.interior(s.UBufObj.toCPU())[..., :5], o.interior()[..., :5])
    print(plm, '3d', err)
    assert np.all(err <= TOL64), err
    from hydro.solver.batch import HipBatch
    from hydro._cfg import make_cfg
    kws = [dict(dim=2, size=n, mins=(-1, -1), maxs=(1, 1), plm=plm, slope_limiter='minmod',
                integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow', precision='double', fast_math=True)
           for n in [(64, 64), (100, 37)]]
    b = HipBatch(kws)
    U0s = []
    for i, kwi in enumerate(kws):
        cfg = make_cfg(**kwi)
        U0s.append(ie.to_ghosted(cfg, ie.quadrant(cfg, ie.random_quadrant_states(np.random.default_rng(30 + i)))))
        b.upload(i, U0s[i])
    b.constrainU()
    b.update(5)
    for i, kwi in enumerate(kws):
        sp = HipFVSolver(dim=2, gridSize=kwi['size'], mins=kwi['mins'], maxs=kwi['maxs'], usePLM=plm,
                         slopeLimiter='minmod', integrator=kwi['integrator'], cfl=kwi['cfl'], boundary='freeflow',
                         fastMath=True)
        sp.resetState(U0s[i])
        sp.update(5)
        assert np.array_equal(b.download(i), sp.UBufObj.toCPU()), i


def test_piecewise_constant_equals_no_reconstruction():
    """'piecewise constant' (plm.cl:10-24) gives the edge states U itself:
    the same bits as no reconstruction"""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    outs = []
    for plm in ('piecewise constant', None):
        s = HipFVSolver(dim=2, gridSize=(64, 40), usePLM=plm, integrator='Runge-Kutta 2, TVD', cfl=.4,
                        boundary='periodic')
        s.resetState(ie.to_ghosted(s.cfg, ie.quadrant(s.cfg, ie.random_quadrant_states(np.random.default_rng(6)))))
        s.update(10)
        outs.append(s.UBufObj.toCPU())
    assert np.array_equal(outs[0], outs[1])
