"""GPU parity: the HIP path through the C ABI against the CPU oracle.

FP64 results must agree with the oracle to <= 1e-12 relative L1 per conserved
variable; FP32-bulk results to <= 1e-5 against the FP64 oracle.
"""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _pair(init, nsteps, **kw):
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(**kw)
    Uint = init(s.cfg)
    U0 = init_euler.to_ghosted(s.cfg, Uint)
    s.resetState(U0.astype(s.dtype))
    o = pyoracle.Oracle(pyoracle.make_cfg(**_oracle_kw(kw)))
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    return s, o


def _oracle_kw(kw):
    m = {'dim': 'dim', 'gridSize': 'size', 'mins': 'mins', 'maxs': 'maxs', 'usePLM': 'plm', 'eqn': 'eqn', 'flux': 'flux',
         'slopeLimiter': 'slope_limiter', 'fluxLimiter': 'flux_limiter', 'integrator': 'integrator',
         'cfl': 'cfl', 'boundary': 'boundary', 'fixedDT': 'fixed_dt', 'boundaryFixed': 'boundary_fixed'}
    out = {m[k]: v for k, v in kw.items() if k in m}
    out['precision'] = 'double'
    return out


def _interior(dim, U):
    sl = [slice(None)] * 4
    for i in range(dim):
        sl[2 - i] = slice(2, -2)
    return U[tuple(sl)]


def _kh(seed=0):
    from hydro.init import euler as init_euler
    return lambda cfg: init_euler.kelvin_helmholtz(cfg, seed=seed)


def _sod(cfg):
    from hydro.init import euler as init_euler
    return init_euler.sod(cfg)


def _quad(seed):
    from hydro.init import euler as init_euler
    rng = np.random.default_rng(seed)
    states = init_euler.random_quadrant_states(rng)
    return lambda cfg: init_euler.quadrant(cfg, states)


def _check(s, o, dim, tol, nsteps, label):
    s.update(nsteps)
    o.update(nsteps)
    got = _interior(dim, s.UBufObj.toCPU())[..., :5]
    ref = _interior(dim, o.U)[..., :5]
    err = rel_l1(got, ref)
    print(f'{label}: rel L1 per field {err}')
    assert np.all(np.isfinite(got))
    assert np.all(err <= tol), f'{label}: {err}'
    assert abs(s.t - o.t) <= max(1e-12, tol) * abs(o.t)
    return err


def test_sod_1d_superbee_fe():
    """1D Sod, Roe + superbee flux limiter, forward Euler, cfl .5, fixed ends, 100 steps"""
    left = [1., 0, 0, 0, 1. / .4, 0]
    right = [.125, 0, 0, 0, .1 / .4, 0]
    kw = dict(dim=1, gridSize=256, fluxLimiter='superbee', integrator='forward Euler', cfl=.5,
              boundary={'xmin': 'fixed', 'xmax': 'fixed'}, boundaryFixed={'xmin': left, 'xmax': right})
    s, o = _pair(_sod, 100, **kw)
    _check(s, o, 1, TOL64, 100, 'sod1d')


@pytest.mark.parametrize('n', [(64, 64), (70, 37), (130, 9)])
def test_kh_2d_plm_rk2_periodic(n):
    kw = dict(dim=2, gridSize=n, mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic')
    s, o = _pair(_kh(1), 20, **kw)
    _check(s, o, 2, TOL64, 20, f'kh{n}')


@pytest.mark.parametrize('bc', ['freeflow', 'mirror'])
def test_quadrant_2d_plm_bcs(bc):
    kw = dict(dim=2, gridSize=(61, 66), usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary=bc)
    s, o = _pair(_quad(3), 15, **kw)
    _check(s, o, 2, TOL64, 15, f'quad-{bc}')


@pytest.mark.parametrize('lim', ['superbee', 'van Leer', 'minmod', 'Koren'])
def test_quadrant_2d_flux_limiter(lim):
    kw = dict(dim=2, gridSize=(48, 40), fluxLimiter=lim, integrator='Runge-Kutta 3, TVD', cfl=.3,
              boundary='freeflow')
    s, o = _pair(_quad(4), 10, **kw)
    _check(s, o, 2, TOL64, 10, f'fluxlim-{lim}')


@pytest.mark.parametrize('rows', [1, 3, 32])
@pytest.mark.parametrize('mode', ['plm', 'flux limiter', 'donor cell'])
def test_strip_edges_2d(rows, mode):
    """64-column strips take their two edge interfaces from a per-row-block pass:
    three strips (the last one partial), each reconstruction, rows per wave
    1..32, mirror faces in x"""
    from hydro._abi import HydroError
    rec = {'plm': dict(usePLM='plm cons', slopeLimiter='van Leer'), 'flux limiter': dict(fluxLimiter='superbee'),
           'donor cell': {}}[mode]
    kw = dict(dim=2, gridSize=(150, 45), integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='mirror', **rec)
    s, o = _pair(_quad(5), 8, **kw)
    with pytest.raises(HydroError):
        s.setLaunch(33, 0)
    s.setLaunch(rows, 0)
    _check(s, o, 2, TOL64, 8, f'edges-{mode}-{rows}')


@pytest.mark.parametrize('fast', [False, True])
@pytest.mark.parametrize('mode', ['plm', 'flux limiter'])
def test_strip_edges_2d_mixed_faces(mode, fast):
    """edge interfaces next to fixed, mirror and freeflow faces: three strips
    (the last partial), fixed values entering both the edge pass and the rows"""
    rec = {'plm': dict(usePLM='plm cons', slopeLimiter='minmod'), 'flux limiter': dict(fluxLimiter='van Leer')}[mode]
    left = [1., .1, 0, 0, 2.5, 0]
    top = [.5, 0, -.2, 0, 1.2, 0]
    kw = dict(dim=2, gridSize=(150, 45), integrator='Runge-Kutta 3, TVD', cfl=.3,
              boundary={'xmin': 'fixed', 'xmax': 'mirror', 'ymin': 'freeflow', 'ymax': 'fixed'},
              boundaryFixed={'xmin': left, 'ymax': top}, **rec)
    s, o = _pair(_quad(6), 6, fastMath=fast, **kw)
    s.setLaunch(7, 0)
    _check(s, o, 2, TOL64, 6, f'mixed-{mode}-{fast}')


@pytest.mark.parametrize('integ', ['forward Euler', 'Runge-Kutta 2', 'Runge-Kutta 2 Heun', 'Runge-Kutta 2 Ralston',
                                   'Runge-Kutta 3', 'Runge-Kutta 4', 'Runge-Kutta 4, 3/8ths rule',
                                   'Runge-Kutta 2, non-TVD', 'Runge-Kutta 4, TVD', 'Runge-Kutta 4, non-TVD'])
def test_integrators(integ):
    kw = dict(dim=2, gridSize=(40, 40), usePLM='plm cons', slopeLimiter='monotized central',
              integrator=integ, cfl=.3, boundary='periodic', mins=(-.5, -.5), maxs=(.5, .5))
    s, o = _pair(_kh(2), 6, **kw)
    _check(s, o, 2, TOL64, 6, f'int-{integ}')


def test_roe_no_limiter_1d():
    kw = dict(dim=1, gridSize=200, integrator='Runge-Kutta 2, TVD', cfl=.5, boundary='mirror')
    s, o = _pair(_sod, 40, **kw)
    _check(s, o, 1, TOL64, 40, 'roe1d')


def test_calc_dt_and_host_step():
    kw = dict(dim=2, gridSize=(64, 32), usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='freeflow')
    s, o = _pair(_quad(5), 0, **kw)
    dt_s = s.calcDT()
    dt_o = o.calc_dt()
    assert dt_s == dt_o
    s.step(dt_s)
    o.step(dt_o)
    got = _interior(2, s.UBufObj.toCPU())[..., :5]
    assert np.all(rel_l1(got, _interior(2, o.U)[..., :5]) <= TOL64)


def test_download_ghosts_match_boundary():
    """ghost cells of the download equal the oracle's after boundary()"""
    for bc in ['periodic', 'mirror', 'freeflow']:
        kw = dict(dim=2, gridSize=(13, 11), usePLM='plm cons', slopeLimiter='minmod',
                  integrator='Runge-Kutta 2, TVD', cfl=.4, boundary=bc)
        s, o = _pair(_quad(6), 0, **kw)
        s.update(2)
        o.update(2)
        got = s.UBufObj.toCPU()[..., :5]
        ref = o.U[..., :5]
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), bc


def test_conserved_totals_periodic():
    kw = dict(dim=2, gridSize=(64, 64), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons',
              slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic')
    s, o = _pair(_kh(0), 0, **kw)
    t0 = s.conservedTotals()
    ref0 = _interior(2, o.U)[..., :5].reshape(-1, 5).sum(axis=0)
    assert np.allclose(t0, ref0, rtol=1e-14, atol=1e-12)
    s.update(10)
    t1 = s.conservedTotals()
    # periodic fluxes telescope: mass, momentum, energy are conserved to roundoff
    assert np.allclose(t1[[0, 1, 4]], t0[[0, 1, 4]], rtol=1e-12, atol=1e-11)


def test_kh_fp32_vs_fp64_oracle():
    kw = dict(dim=2, gridSize=(128, 128), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons',
              slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic',
              real='float')
    s, o = _pair(_kh(0), 50, **kw)
    _check(s, o, 2, TOL32, 50, 'kh-fp32')


def _glm(seed, sigma=0.):
    from hydro.init import glm_maxwell
    return lambda cfg: glm_maxwell.gaussian_pulses(cfg, seed=seed, conductivity=sigma)


@pytest.mark.parametrize('n,plm,bc,sigma', [((64, 48), 'plm cons', 'periodic', 0.), ((50, 40), None, 'freeflow', .5),
                                             ((40, 33), 'plm cons', 'mirror', 1.), ((44, 37), 'ppm', 'mirror', .5)])
def test_glm_maxwell_2d(n, plm, bc, sigma):
    kw = dict(dim=2, gridSize=n, eqn='glm_maxwell', usePLM=plm, slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary=bc)
    s, o = _pair(_glm(7, sigma), 20, **kw)
    s.update(20)
    o.update(20)
    got = _interior(2, s.UBufObj.toCPU())
    ref = _interior(2, o.U)
    err = rel_l1(got[..., :8], ref[..., :8])
    print('glm', n, err)
    assert np.all(err <= TOL64), err
    assert np.array_equal(got[..., 8:], ref[..., 8:])   # static fields carried unchanged


def test_glm_maxwell_1d_ppm():
    """'ppm' (plm.cl:882-944) with the GLM-Maxwell equations: the kappa = 1/3
    edges of the integrated fields, static fields copied"""
    kw = dict(dim=1, gridSize=300, eqn='glm_maxwell', usePLM='ppm', slopeLimiter='van Leer',
              integrator='Runge-Kutta 3, TVD', cfl=.4, boundary='freeflow')
    s, o = _pair(_glm(12), 30, **kw)
    s.update(30)
    o.update(30)
    got = _interior(1, s.UBufObj.toCPU())
    ref = _interior(1, o.U)
    err = rel_l1(got[..., :8], ref[..., :8])
    print('glm ppm 1d', err)
    assert np.all(err <= TOL64), err


def test_glm_maxwell_1d_flux_limiter():
    kw = dict(dim=1, gridSize=300, eqn='glm_maxwell', fluxLimiter='superbee',
              integrator='Runge-Kutta 3, TVD', cfl=.4, boundary='freeflow')
    s, o = _pair(_glm(8), 30, **kw)
    s.update(30)
    o.update(30)
    err = rel_l1(_interior(1, s.UBufObj.toCPU())[..., :8], _interior(1, o.U)[..., :8])
    assert np.all(err <= TOL64), err


# ---- FAST math policy: faithfully rounded division / sqrt, still <= 1e-12
@pytest.mark.parametrize('n,steps', [((128, 128), 100), ((96, 70), 60)])
def test_kh_fast_math_fp64(n, steps):
    kw = dict(dim=2, gridSize=n, mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', fastMath=True)
    s, o = _pair(_kh(4), steps, **kw)
    _check(s, o, 2, TOL64, steps, f'kh-fast{n}')


@pytest.mark.parametrize('lim,integ', [('superbee', 'forward Euler'), ('van Leer', 'Runge-Kutta 3, TVD'),
                                       ('Koren', 'Runge-Kutta 2, TVD')])
def test_flux_limiter_fast_math(lim, integ):
    kw = dict(dim=2, gridSize=(48, 52), fluxLimiter=lim, integrator=integ, cfl=.3, boundary='freeflow',
              fastMath=True)
    s, o = _pair(_quad(9), 20, **kw)
    _check(s, o, 2, TOL64, 20, f'fast-{lim}')


@pytest.mark.parametrize('slim', ['superbee', 'van Leer', 'monotized central'])
def test_plm_slope_limiters_fast_math(slim):
    """the fast policy forms phi(r) * dU without r for these limiters"""
    kw = dict(dim=2, gridSize=(64, 40), usePLM='plm cons', slopeLimiter=slim, integrator='Runge-Kutta 2, TVD',
              cfl=.4, boundary='freeflow', fastMath=True)
    s, o = _pair(_quad(13), 25, **kw)
    _check(s, o, 2, TOL64, 25, f'fast-plm-{slim}')


def test_sod_fast_math():
    left = [1., 0, 0, 0, 2.5, 0]
    right = [.125, 0, 0, 0, .25, 0]
    kw = dict(dim=1, gridSize=256, fluxLimiter='superbee', integrator='forward Euler', cfl=.5,
              boundary={'xmin': 'fixed', 'xmax': 'fixed'}, boundaryFixed={'xmin': left, 'xmax': right},
              fastMath=True)
    s, o = _pair(_sod, 100, **kw)
    _check(s, o, 1, TOL64, 100, 'sod-fast')


def test_glm_fast_math():
    kw = dict(dim=2, gridSize=(64, 64), eqn='glm_maxwell', usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', fastMath=True)
    s, o = _pair(_glm(11), 30, **kw)
    s.update(30)
    o.update(30)
    err = rel_l1(_interior(2, s.UBufObj.toCPU())[..., :8], _interior(2, o.U)[..., :8])
    assert np.all(err <= TOL64), err


def test_nonzero_out_of_plane_momentum_2d():
    """a 2D state with m.z != 0 takes the three-component kernels"""
    def init(cfg):
        U = _kh(5)(cfg)
        U[..., 3] = .1 * U[..., 0]
        U[..., 4] += .5 * .01 * U[..., 0]
        return U
    kw = dict(dim=2, gridSize=(64, 40), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic')
    for fast in (False, True):
        s, o = _pair(init, 20, fastMath=fast, **kw)
        s.update(20)
        o.update(20)
        got = _interior(2, s.UBufObj.toCPU())[..., :5]
        err = rel_l1(got, _interior(2, o.U)[..., :5])
        assert np.all(err <= TOL64), (fast, err)
        assert np.abs(got[..., 3]).max() > 0


def test_kh_fp32_fast_vs_fp64_oracle():
    kw = dict(dim=2, gridSize=(128, 128), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons',
              slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic',
              real='float', fastMath=True)
    s, o = _pair(_kh(0), 50, **kw)
    _check(s, o, 2, TOL32, 50, 'kh-fp32-fast')


@pytest.mark.parametrize('plm,n', [('plm cons', (128, 96)), ('ppm', (90, 70)), (None, (80, 64))])
def test_glm_maxwell_2d_fp32_vs_fp64_oracle(plm, n):
    """FP32 GLM-Maxwell (the serving mix's third kind: periodic, RK2-TVD,
    sigma = 0) against the FP64 oracle from the same float inputs: <= 1e-5.
    'ppm' (not in the serving mix) is held to ten times the oracle's own
    response to one float rounding of its input instead: its phi(dUL / dUR)
    takes phi(0) where dUR is exactly 0 and phi(+-inf) beside it, so on the
    pulses' flat background a rounding of the input moves the oracle's FP64
    result by 6-7e-5 after 40 steps ('plm cons': 2.5e-7), and a float run
    differs from it by as much."""
    from gpu_util import init_euler_ghosted, oracle_sensitivity, pair
    kw = dict(dim=2, gridSize=n, eqn='glm_maxwell', usePLM=plm, slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', real='float', fastMath=True)
    s, o = pair(_glm(14, 0.), **kw)
    s.update(40)
    o.update(40)
    got = _interior(2, s.UBufObj.toCPU()).astype(np.float64)
    ref = _interior(2, o.U)
    err = rel_l1(got[..., :8], ref[..., :8])
    sens = np.zeros(8)
    if plm == 'ppm':
        import pyoracle
        U0 = init_euler_ghosted(_glm(14, 0.), s.cfg, np.float32)
        sens = oracle_sensitivity(pyoracle.make_cfg(**_oracle_kw(kw)), U0, 2, 40, 2. ** -24, 8)
    bar = np.maximum(TOL32, 10 * sens)
    print('glm fp32', plm, err, 'oracle sensitivity', sens, 'bar', bar)
    assert np.all(err <= bar), (err, bar)
    assert abs(s.t - o.t) <= TOL32 * o.t


@pytest.mark.parametrize('dim,n', [(2, (52, 45)), (1, 260)])
def test_glm_maxwell_ppm_fast_fp64(dim, n):
    """'ppm' with GLM-Maxwell in the FAST FP64 kernels (<= 1e-12)"""
    kw = dict(dim=dim, gridSize=n, eqn='glm_maxwell', usePLM='ppm', slopeLimiter='van Leer',
              integrator='Runge-Kutta 3, TVD', cfl=.4, boundary='mirror' if dim == 2 else 'freeflow', fastMath=True)
    s, o = _pair(_glm(15, .5), 25, **kw)
    s.update(25)
    o.update(25)
    err = rel_l1(_interior(dim, s.UBufObj.toCPU())[..., :8], _interior(dim, o.U)[..., :8])
    print('glm ppm fast', dim, err)
    assert np.all(err <= TOL64), err


@pytest.mark.parametrize('eqn', ['glm_maxwell', 'shallow_water'])
def test_halo_strips_rows_same_bits(eqn):
    """the equations with a source march halo strips (60 output columns of
    64 lanes); a grid advances to the same bits at any rows per wave,
    including the long strips the host picks for GLM-Maxwell FP64 (10-16):
    three strips per row (the last partial), FAST FP64, 12 RK2-TVD steps"""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler, glm_maxwell as iglm, shallow_water as isw
    if eqn == 'glm_maxwell':
        init, bc = (lambda cfg: iglm.gaussian_pulses(cfg, seed=3, conductivity=0.)), 'periodic'
    else:
        init, bc = (lambda cfg: isw.humps(cfg, seed=3)), 'freeflow'
    out = {}
    for rows in (3, 7, 10, 16):
        s = HipFVSolver(dim=2, gridSize=(130, 70), mins=(-.5, -.5), maxs=(.5, .5), eqn=eqn, usePLM='plm cons',
                        slopeLimiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=.4, boundary=bc,
                        real='double', fastMath=True)
        s.setLaunch(rows, 0)
        s.resetState(init_euler.to_ghosted(s.cfg, init(s.cfg)).astype(s.dtype))
        s.update(12)
        out[rows] = (s.UBufObj.toCPU(), s.t)
        del s
    U3, t3 = out[3]
    assert np.isfinite(U3).all()
    for rows, (U, t) in out.items():
        assert t == t3, rows
        assert np.array_equal(U, U3), rows
