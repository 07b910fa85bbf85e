"""GPU: the display statistics of a stored field (hydro_var_stats,
calcDisplayVarRangeAndAvgAndStdDev of solverbase.lua:2871-2994) against the
oracle's: min and max over the ghosted grid exactly (ghosts as every boundary
kind fills them), average and standard deviation over the interior within
1e-12 relative (FP64; the sums run in another order) and 1e-5 for FP32."""
import numpy as np
import pytest

from test_gpu_parity import _pair, _quad, _sod

pytestmark = pytest.mark.gpu


def _stats_oracle(o, f):
    import ctypes
    import pyoracle
    out = np.zeros(4)
    pyoracle.lib().oracle_var_stats(o.h, f, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _compare(s, o, fields, tol):
    for f in fields:
        got = np.array(s.calcDisplayVarRangeAndAvgAndStdDev(f))
        ref = _stats_oracle(o, f)
        print(f, got, ref)
        if tol == 0:
            assert got[0] == ref[0] and got[1] == ref[1], (f, got, ref)
        else:
            assert np.allclose(got[:2], ref[:2], rtol=tol, atol=tol * np.abs(ref).max()), (f, got, ref)
        scale = max(abs(ref[2]), ref[3], 1e-300)
        assert abs(got[2] - ref[2]) <= max(tol, 1e-12) * scale, (f, got, ref)
        assert abs(got[3] - ref[3]) <= max(tol, 1e-12) * max(ref[3], 1e-300), (f, got, ref)


@pytest.mark.parametrize('bc', ['mirror', 'periodic', 'mixed'])
def test_stats_2d(bc):
    boundary = bc if bc != 'mixed' else {'xmin': 'fixed', 'xmax': 'mirror', 'ymin': 'freeflow', 'ymax': 'fixed'}
    kw = dict(dim=2, gridSize=(70, 33), integrator='Runge-Kutta 2, TVD', cfl=.3, boundary=boundary,
              usePLM='plm cons', slopeLimiter='minmod')
    if bc == 'mixed':
        kw['boundaryFixed'] = {'xmin': [2.5, .6, .1, 0, 6., 0], 'ymax': [.05, 0, -.02, 0, .3, 0]}
    s, o = _pair(_quad(9), 5, **kw)
    s.update(5)
    o.update(5)
    _compare(s, o, range(5), 0)


def test_stats_1d_fixed_and_3d():
    left = [1., 0, 0, 0, 2.5, 0]
    right = [.125, 0, 0, 0, .25, 0]
    kw = dict(dim=1, gridSize=200, fluxLimiter='superbee', integrator='forward Euler', cfl=.5,
              boundary={'xmin': 'fixed', 'xmax': 'fixed'}, boundaryFixed={'xmin': left, 'xmax': right})
    s, o = _pair(_sod, 20, **kw)
    s.update(20)
    o.update(20)
    _compare(s, o, range(5), 0)
    from test_gpu_3d import _octants
    kw = dict(dim=3, gridSize=(20, 14, 9), integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='mirror',
              mins=(-1, -1, -1), maxs=(1, 1, 1))
    s, o = _pair(_octants(5), 3, **kw)
    s.update(3)
    o.update(3)
    _compare(s, o, range(5), 0)


def test_stats_glm_static_fields():
    from test_gpu_parity import _glm
    kw = dict(dim=2, gridSize=(40, 30), eqn='glm_maxwell', usePLM='plm cons', slopeLimiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='mirror')
    s, o = _pair(_glm(4), 4, **kw)
    s.update(4)
    o.update(4)
    _compare(s, o, range(12), 0)


def test_stats_fp32():
    kw = dict(dim=2, gridSize=(64, 48), integrator='Runge-Kutta 2, TVD', cfl=.3, boundary='freeflow',
              usePLM='plm cons', slopeLimiter='minmod')
    s, o = _pair(_quad(10), 4, real='float', fastMath=True, **kw)
    s.update(4)
    o.update(4)
    _compare(s, o, range(5), 1e-5)
    from hydro._abi import HydroError
    with pytest.raises(HydroError):
        s.calcDisplayVarRangeAndAvgAndStdDev(5)


def test_save_fits_matches_download(tmp_path):
    """GridSolver:save (gridsolver.lua:1465-1470) of the state: the FITS
    file holds the downloaded ghosted buffer in the reference's plane layout,
    2D and 3D (z interleaved along x), FP64 and FP32"""
    from hydro import fits
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as ie
    for kw in (dict(dim=2, gridSize=(40, 24)), dict(dim=3, gridSize=(20, 12, 8), real='float')):
        s = HipFVSolver(integrator='Runge-Kutta 2, TVD', boundary='freeflow', **kw)
        s.resetState(ie.to_ghosted(s.cfg, ie.quadrant(s.cfg, ie.random_quadrant_states(np.random.default_rng(4)))).astype(s.dtype))
        s.update(3)
        names = s.save(str(tmp_path / f'results{kw["dim"]}'))
        assert names == [str(tmp_path / f'results{kw["dim"]}_UBuf.fits')]
        planes, hdr = fits.read_image(names[0])
        U = s.UBufObj.toCPU()
        d, h, w, c = U.shape
        assert (hdr['NAXIS1'], hdr['NAXIS2'], hdr['NAXIS3']) == (w * d, h, c)
        assert planes.dtype == U.dtype
        assert np.array_equal(planes, fits.buffer_planes(U))
        assert np.array_equal(planes[0, :, :w], U[0, :, :, 0])


@pytest.mark.parametrize('real', ['double', 'float'])
def test_stats_display_vars(real):
    """the derived Euler display variables (euler.lua:240-260) on the GPU
    against the oracle: FP64 within 1e-13 (pow of 'S' may differ in the last
    place), the FP32 state within 1e-5 of the FP64 oracle"""
    from hydro._cfg import EULER_DISPLAY_VARS
    kw = dict(dim=2, gridSize=(70, 33), integrator='Runge-Kutta 2, TVD', cfl=.3,
              boundary={'xmin': 'fixed', 'xmax': 'mirror', 'ymin': 'freeflow', 'ymax': 'periodic'},
              boundaryFixed={'xmin': [2.5, .6, .1, 0, 6., 0]}, usePLM='plm cons', slopeLimiter='minmod')
    s, o = _pair(_quad(19), 5, real=real, **kw)
    s.update(5)
    o.update(5)
    tol = 1e-13 if real == 'double' else 1e-5
    for name, vid in EULER_DISPLAY_VARS.items():
        got = np.array(s.calcDisplayVarRangeAndAvgAndStdDev(name))
        ref = _stats_oracle(o, vid)
        print(name, got, ref)
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.allclose(got, ref, rtol=tol, atol=tol * scale), (name, got, ref)
    from hydro import _abi
    with pytest.raises(_abi.HydroError):
        s.calcDisplayVarRangeAndAvgAndStdDev(200)


def test_display_vars_refused_for_glm():
    """the HYDRO_VAR_* ids are the Euler equations' display variables"""
    from hydro import _abi
    from hydro.solver.hipfvsolver import HipFVSolver
    s = HipFVSolver(dim=1, gridSize=32, eqn='glm_maxwell')
    with pytest.raises(_abi.HydroError):
        s.calcDisplayVarRangeAndAvgAndStdDev('P')
    with pytest.raises(ValueError):
        s.calcDisplayVarRangeAndAvgAndStdDev('no such var')
