"""GPU parity at the full size of the Kelvin-Helmholtz configuration: 1024^2
cells, periodic, Roe + 'plm cons' / minmod, RK2-TVD, CFL .4, 500 steps, FP64,
in both arithmetic policies (EXACT: IEEE division and square root, no FP
contraction; FAST: the faithfully rounded sequences with FP contraction,
which bench.py times). The oracle runs the same steps on the host's cores as
the checker. Bar: <= 1e-12 relative L1 per conserved variable, same t. And at
the full size of the 3D blast configuration (256^3), a few steps of FAST FP64
(<= 1e-12) and of the FP32 bulk form (<= 1e-5 against the FP64 oracle)."""
import os

import numpy as np
import pytest

from conftest import rel_l1
from gpu_util import interior

pytestmark = pytest.mark.gpu

N, STEPS = 1024, 500


@pytest.fixture(scope='module')
def kh_oracle_run():
    import pyoracle
    from hydro._cfg import make_cfg
    from hydro.init import euler as init_euler
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    cfg = make_cfg(dim=2, size=(N, N), mins=(-.5, -.5), maxs=(.5, .5), plm='plm cons', slope_limiter='minmod',
                   integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', precision='double')
    U0 = init_euler.to_ghosted(cfg, init_euler.kelvin_helmholtz(cfg, seed=0))
    o = pyoracle.Oracle(cfg)
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    print(f'oracle: KH {N}^2, {STEPS} steps on {pyoracle.get_threads()} threads', flush=True)
    o.update(STEPS)
    return U0, interior(2, o.U)[..., :5].copy(), o.t


@pytest.mark.parametrize('fast', [False, True], ids=['exact', 'fast'])
def test_kh_1024_500_steps_fp64(kh_oracle_run, fast):
    from hydro.solver.hipfvsolver import HipFVSolver
    U0, ref, t_ref = kh_oracle_run
    s = HipFVSolver(dim=2, gridSize=(N, N), mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons', slopeLimiter='minmod',
                    integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', fastMath=fast)
    s.resetState(U0)
    s.update(STEPS)
    got = interior(2, s.UBufObj.toCPU())[..., :5]
    err = rel_l1(got, ref)
    print(f'KH {N}^2 x {STEPS} {"FAST" if fast else "EXACT"}: rel L1 {err}, t {s.t!r} vs {t_ref!r}, '
          f'bitwise {np.array_equal(got, ref)}')
    assert np.all(err <= 1e-12), err
    assert abs(s.t - t_ref) <= 1e-14 * t_ref


# The 3D configuration at its full size: the spherical blast, 256^3 freeflow on
# [-1, 1]^3, Roe + van Leer flux limiter, RK3-TVD, CFL .3 (as bench.py's 3D
# leg). The oracle takes a few steps at this size in seconds; the bars are
# those of the 2D case for FAST FP64 and 1e-5 for the FP32 bulk form.
N3, STEPS3 = 256, 3


@pytest.fixture(scope='module')
def blast_oracle_run():
    import pyoracle
    from hydro._cfg import make_cfg
    from hydro.init import euler as init_euler
    pyoracle.set_threads(min(16, os.cpu_count() or 1))
    cfg = make_cfg(dim=3, size=(N3, N3, N3), mins=(-1, -1, -1), maxs=(1, 1, 1), flux_limiter='van Leer',
                   integrator='Runge-Kutta 3, TVD', cfl=.3, boundary='freeflow', precision='double')
    U0 = init_euler.to_ghosted(cfg, init_euler.sphere_blast(cfg, seed=0))
    o = pyoracle.Oracle(cfg)
    o.set_U(U0)
    o.boundary()
    o.constrain_U()
    print(f'oracle: blast {N3}^3, {STEPS3} steps on {pyoracle.get_threads()} threads', flush=True)
    o.update(STEPS3)
    return U0, interior(3, o.U)[..., :5].copy(), o.t


@pytest.mark.parametrize('real', ['double', 'float'])
def test_blast_256_cubed(blast_oracle_run, real):
    from hydro.solver.hipfvsolver import HipFVSolver
    U0, ref, t_ref = blast_oracle_run
    s = HipFVSolver(dim=3, gridSize=(N3, N3, N3), mins=(-1, -1, -1), maxs=(1, 1, 1), fluxLimiter='van Leer',
                    integrator='Runge-Kutta 3, TVD', cfl=.3, boundary='freeflow', real=real, fastMath=True)
    s.resetState(U0.astype(s.dtype))
    s.update(STEPS3)
    got = interior(3, s.UBufObj.toCPU())[..., :5].astype(np.float64)
    err = rel_l1(got, ref)
    print(f'blast {N3}^3 x {STEPS3} {real}: rel L1 {err}, t {s.t!r} vs {t_ref!r}')
    if real == 'double':
        assert np.all(err <= 1e-12), err
        assert abs(s.t - t_ref) <= 1e-14 * t_ref
    else:
        assert np.all(err <= 1e-5), err
        assert abs(s.t - t_ref) <= 1e-5 * t_ref
