"""GPU: the 3D stage kernel (Euler) through the C ABI against the CPU oracle.

FP64 (IEEE and fast) within 1e-12 relative L1 per conserved variable, FP32
bulk within 1e-5 of the FP64 oracle. The grids are small and uneven so that
strips, row blocks (3 rows and the halo wave) and plane chunks end inside the
grid, and every boundary kind meets every face; grids of 1 and 2 rows put
row waves outside the grid and the halo wave's rows through the y faces.
"""
import numpy as np
import pytest

from conftest import rel_l1

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _octants(seed):
    """eight constant states, one per octant about the box centre, with
    rho, P ~ U[.2, 2] and |v| <= .5"""
    from hydro.init import euler as ie

    def init(cfg):
        rng = np.random.default_rng(seed)
        x, y, z = ie.mesh(cfg)
        shape = ie.interior_shape(cfg)
        mid = [.5 * (cfg.mins[i] + cfg.maxs[i]) for i in range(3)]
        idx = (np.broadcast_to(x > mid[0], shape).astype(int) + 2 * np.broadcast_to(y > mid[1], shape)
               + 4 * np.broadcast_to(z > mid[2], shape))
        rho = rng.uniform(.2, 2, 8)[idx]
        P = rng.uniform(.2, 2, 8)[idx]
        v = [rng.uniform(-.3, .3, 8)[idx] for _ in range(3)]
        return ie.cons_from_prim(rho, v, P, 1.4)
    return init


def _blast(seed, radius=.45):
    from hydro.init import euler as ie
    return lambda cfg: ie.sphere_blast(cfg, seed=seed, radius=radius)


def _run(init, steps, real='double', fast=False, rows=0, **kw):
    """the HIP solver and the oracle on the same initial state; kw are
    make_cfg names (size, boundary, flux_limiter, plm, ...)"""
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    from hydro._cfg import make_cfg
    box = dict(mins=(-1, -1, -1), maxs=(1, 1, 1))
    s = HipFVSolver(dim=3, gridSize=kw['size'], integrator=kw['integrator'], cfl=kw['cfl'],
                    boundary=kw['boundary'], boundaryFixed=kw.get('boundary_fixed'),
                    fluxLimiter=kw.get('flux_limiter', 'donor cell'), usePLM=kw.get('plm'),
                    slopeLimiter=kw.get('slope_limiter', 'minmod'), real=real, fastMath=fast, **box)
    if rows:
        s.setLaunch(rows, 0)
    U0 = ie.to_ghosted(s.cfg, init(s.cfg))
    s.resetState(U0.astype(s.dtype))
    o = pyoracle.Oracle(make_cfg(dim=3, precision='double', **box, **kw))
    o.set_U(U0.astype(s.dtype).astype(np.float64))
    o.boundary()
    o.constrain_U()
    s.update(steps)
    o.update(steps)
    got = o.interior(s.UBufObj.toCPU())[..., :5]
    ref = o.interior()[..., :5]
    err = rel_l1(got, ref)
    print(f'3D {kw.get("size")} {real} fast={fast}: rel L1 {err}')
    return err, s, o


@pytest.mark.parametrize('size,bc', [
    ((20, 12, 10), 'freeflow'),
    ((33, 9, 17), 'periodic'),
    ((14, 19, 11), 'mirror'),
    ((64, 8, 6), {'xmin': 'mirror', 'xmax': 'freeflow', 'ymin': 'periodic', 'ymax': 'periodic',
                  'zmin': 'freeflow', 'zmax': 'mirror'}),
])
def test_3d_flux_limiter_exact(size, bc):
    err, s, o = _run(_octants(1), 6, size=size, boundary=bc, flux_limiter='van Leer',
                     integrator='Runge-Kutta 3, TVD', cfl=.3)
    assert np.all(err <= TOL64), err
    assert s.t == o.t


@pytest.mark.parametrize('mode', ['plm', 'none', 'superbee'])
def test_3d_reconstructions_exact(mode):
    kw = dict(size=(23, 13, 9), boundary='freeflow', integrator='Runge-Kutta 2, TVD', cfl=.3)
    if mode == 'plm':
        kw.update(plm='plm cons', slope_limiter='monotized central')
    elif mode == 'superbee':
        kw.update(flux_limiter='superbee')
    err, s, o = _run(_octants(2), 5, **kw)
    assert np.all(err <= TOL64), (mode, err)


def test_3d_fixed_faces_and_forward_euler():
    left = [1., 0, 0, 0, 2.5, 0]
    right = [.5, 0, 0, 0, 1., 0]
    bc = {f: 'fixed' for f in ('xmin', 'xmax', 'ymin', 'ymax', 'zmin', 'zmax')}
    fixed = {'xmin': left, 'xmax': right, 'ymin': left, 'ymax': right, 'zmin': right, 'zmax': left}
    err, s, o = _run(_octants(3), 8, size=(18, 10, 12), boundary=bc, boundary_fixed=fixed,
                     flux_limiter='minmod', integrator='forward Euler', cfl=.3)
    assert np.all(err <= TOL64), err


@pytest.mark.parametrize('size,mode,bc,fast', [
    ((20, 1, 10), 'superbee', 'periodic', False),
    ((17, 2, 9), 'plm', 'mirror', False),
    ((19, 2, 7), 'none', 'freeflow', False),
    ((13, 1, 8), 'van Leer', 'mirror', True),
    ((22, 4, 9), 'van Leer', {'xmin': 'periodic', 'xmax': 'periodic', 'ymin': 'mirror', 'ymax': 'freeflow',
                              'zmin': 'freeflow', 'zmax': 'freeflow'}, True),
])
def test_3d_thin_in_y(size, mode, bc, fast):
    """fewer rows than a block (or one more): the rows outside the grid are
    computed, never stored, and the halo rows come through the y faces"""
    kw = dict(size=size, boundary=bc, integrator='Runge-Kutta 2, TVD', cfl=.3)
    if mode == 'plm':
        kw.update(plm='plm cons', slope_limiter='minmod')
    elif mode != 'none':
        kw.update(flux_limiter=mode)
    err, s, o = _run(_octants(7), 6, fast=fast, **kw)
    assert np.all(err <= TOL64), (size, mode, err)
    assert s.t == o.t or fast


@pytest.mark.parametrize("rows", [4, 7, 16])
def test_3d_plane_chunks_do_not_change_bits(rows):
    """the plane chunk per block changes the decomposition, not the numbers"""
    err, s, o = _run(_octants(4), 4, size=(21, 11, 30), boundary='periodic', flux_limiter='van Leer',
                     integrator='Runge-Kutta 2, TVD', cfl=.3, rows=rows)
    assert np.all(err <= TOL64), err


def test_3d_plane_chunk_limit():
    from hydro._abi import HydroError
    from hydro.solver.hipfvsolver import HipFVSolver
    s = HipFVSolver(dim=3, gridSize=(8, 8, 8), integrator='forward Euler', cfl=.3, boundary='periodic')
    s.setLaunch(16, 0)
    with pytest.raises(HydroError):
        s.setLaunch(17, 0)


def test_3d_fast_math():
    err, s, o = _run(_blast(5), 10, fast=True, size=(24, 20, 18), boundary='freeflow',
                     flux_limiter='van Leer', integrator='Runge-Kutta 3, TVD', cfl=.3)
    assert np.all(err <= TOL64), err


def test_3d_fp32_vs_fp64_oracle():
    err, s, o = _run(_blast(6), 10, real='float', fast=True, size=(32, 24, 20), boundary='freeflow',
                     flux_limiter='van Leer', integrator='Runge-Kutta 3, TVD', cfl=.3)
    assert np.all(err <= TOL32), err


def test_3d_conserved_totals_periodic():
    """periodic 3D: mass and energy totals stay put to roundoff (FP64)"""
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    s = HipFVSolver(dim=3, gridSize=(24, 16, 20), mins=(-1, -1, -1), maxs=(1, 1, 1), fluxLimiter='van Leer',
                    integrator='Runge-Kutta 3, TVD', cfl=.3, boundary='periodic', fastMath=True)
    s.resetState(ie.to_ghosted(s.cfg, _octants(7)(s.cfg)))
    t0 = s.conservedTotals()
    s.update(20)
    t1 = s.conservedTotals()
    assert abs(t1[0] - t0[0]) <= 1e-12 * abs(t0[0])
    assert abs(t1[4] - t0[4]) <= 1e-12 * abs(t0[4])


def test_3d_rejected_where_not_built():
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.solver.batch import HipBatch
    from hydro._abi import HydroError
    with pytest.raises(HydroError, match='Euler'):
        HipFVSolver(dim=3, gridSize=(8, 8, 8), eqn='glm_maxwell')
    with pytest.raises(HydroError, match='1D and 2D'):
        HipBatch([dict(dim=3, size=(8, 8, 8), integrator='Runge-Kutta 2, TVD')])
