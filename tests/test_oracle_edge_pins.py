"""Pins the oracle's edge-case branches: constrainU's clamps (euler.cl:697-717),
the vacuum and low-enthalpy branches of eigen_forInterface (euler.cl:347-430),
the identity transforms below rhoMin (euler.cl:434-542), and the per-axis
area / volume scale of the flux divergence (fvsolver.cl:64-121).

Every expected value is computed here from the textbook definitions (ideal gas
prims, Roe averages, the Fourier symbol of first-order upwinding), not by
restating the oracle's code path; each case is chosen so that a plausible
transcription mistake (a clamp applied after primFromCons, `<=` in the rhoMin
test of calc_P, the averaged Roe state on a vacuum interface, hTotal kept where
h < 1e-5, 1/dx in place of 1/dy) fails it, and several tests assert that the
mistake's value differs from the expected one by far more than the tolerance.
"""
import ctypes

import numpy as np
import pytest

import pyoracle
from pyoracle import lib, params, _ptr
from hydro._cfg import make_cfg
from hydro.init import euler as init_euler

GAMMA = 1.4
RHO_MIN = 1e-7
P_MIN = 1e-7
P = params(gamma=GAMMA, rho_min=RHO_MIN, P_min=P_MIN)


def _cons(rho, v, p):
    v = np.asarray(v, dtype=float)
    return np.array([rho, *(rho * v), rho * .5 * (v @ v) + p / (GAMMA - 1), 0.])


def _constrain(U):
    U = np.array(U, dtype=float)
    lib().oracle_euler_constrain(ctypes.byref(P), _ptr(U))
    return U


def _pressure(U):
    return (GAMMA - 1) * (U[4] - .5 * (U[1:4] @ U[1:4]) / U[0])


# ---------------------------------------------------------------- constrainU
def test_constrain_admissible_state_unchanged():
    for rho, v, p in [(1., (.3, -.2, .1), .7), (RHO_MIN, (0, 0, 0), 1.), (2., (1, 0, 0), P_MIN)]:
        U = _cons(rho, v, p)
        out = _constrain(U)
        assert np.allclose(out[:5], U[:5], rtol=1e-15, atol=0)


def test_constrain_density_clamp_before_prims():
    """rho < rhoMin: rho becomes rhoMin first, the velocity is m / rhoMin, so
    the momentum comes back unchanged and P is taken with the clamped density"""
    U = np.array([1e-9, 1e-8, -2e-8, 0., 1e-6, 0.])
    out = _constrain(U)
    assert out[0] == RHO_MIN
    assert np.allclose(out[1:4], U[1:4], rtol=1e-15, atol=0)
    v = U[1:4] / RHO_MIN
    p = (GAMMA - 1) * (U[4] - .5 * RHO_MIN * (v @ v))
    assert p > P_MIN
    want_E = RHO_MIN * .5 * (v @ v) + p / (GAMMA - 1)
    assert out[4] == pytest.approx(want_E, rel=1e-13)
    # clamping after primFromCons would keep v = m / 1e-9: 100x the momentum
    wrong_m = U[1:4] / U[0] * RHO_MIN
    assert np.abs(wrong_m - out[1:4]).max() > 1e-7


def test_constrain_rho_at_min_keeps_pressure():
    """after the clamp rho == rhoMin, which calc_P's `rho < rhoMin` test does
    not catch (euler.cl:63-83): P = (gamma - 1) ETotal for a state at rest;
    a `<=` there would give P = 0 -> PMin and ETotal = PMin / (gamma - 1)"""
    out = _constrain([1e-9, 0., 0., 0., 1., 0.])
    assert out[0] == RHO_MIN
    assert out[4] == pytest.approx(1., rel=1e-15)
    assert abs(out[4] - P_MIN / (GAMMA - 1)) > .9


@pytest.mark.parametrize('rho,v,E_excess', [(1., (.5, 0, 0), 1e-9), (2., (.5, .5, 0), -.2), (.3, (0, 0, -1.), -5.)])
def test_constrain_pressure_clamp(rho, v, E_excess):
    """P < PMin (including negative P): rho and v kept,
    ETotal = rho v^2 / 2 + PMin / (gamma - 1) exactly as consFromPrim forms it"""
    v = np.asarray(v, dtype=float)
    U = np.array([rho, *(rho * v), .5 * rho * (v @ v) + E_excess, 0.])
    assert _pressure(U) < P_MIN
    out = _constrain(U)
    assert out[0] == rho
    assert np.allclose(out[1:4], U[1:4], rtol=1e-15, atol=0)
    want = rho * (.5 * (v @ v)) + P_MIN / (GAMMA - 1)
    assert out[4] == pytest.approx(want, rel=1e-15)
    assert _pressure(out) == pytest.approx(P_MIN, rel=1e-6)


def test_constrain_both_clamps_and_idempotence():
    U = np.array([1e-9, 1e-8, 0., 0., 0., 0.])
    out = _constrain(U)
    v = 1e-8 / RHO_MIN
    assert out[0] == RHO_MIN
    assert out[4] == pytest.approx(RHO_MIN * .5 * v * v + P_MIN / (GAMMA - 1), rel=1e-15)
    rng = np.random.default_rng(3)
    for _ in range(50):
        U = np.array([10 ** rng.uniform(-10, 0.5), *rng.uniform(-1, 1, 3), rng.uniform(-1, 2), 0.])
        a = _constrain(U)
        b = _constrain(a)
        # P of the result is PMin up to the rounding of ETotal - EKin
        assert a[0] >= RHO_MIN and _pressure(a) >= P_MIN - 1e-15 * abs(a[4])
        assert np.allclose(b[:5], a[:5], rtol=1e-12, atol=1e-15 * abs(a[4]))


# ---------------------------------------------------- eigen_forInterface
def _eig(UL, UR):
    e = np.zeros(10)
    lib().oracle_euler_eigen_for_interface(ctypes.byref(P), _ptr(np.asarray(UL, float)), _ptr(np.asarray(UR, float)),
                                           _ptr(e))
    return e


def _textbook_state(rho, v, p):
    """the eigen state of one cell: rho, v, hTotal = gamma P / ((gamma - 1) rho) + v^2 / 2,
    Cs = sqrt(gamma P / rho), v^2"""
    v = np.asarray(v, dtype=float)
    return rho, v, GAMMA * p / ((GAMMA - 1) * rho) + .5 * (v @ v), np.sqrt(GAMMA * p / rho), v @ v


@pytest.mark.parametrize('vac_rho', [1e-6, 5e-6, 1e-9])
@pytest.mark.parametrize('vac_side', ['left', 'right'])
def test_eigen_one_sided_vacuum_takes_the_other_state(vac_rho, vac_side):
    """one side below rhoEpsilon = 1e-5: the Roe state is the other side's own
    state (euler.cl:360-385), whatever the vacuum side holds"""
    good = _cons(.7, (.3, -.2, .1), .9)
    vac = np.array([vac_rho, 3., -7., 2., 50., 0.])
    e = _eig(vac, good) if vac_side == 'left' else _eig(good, vac)
    rho, v, H, Cs, vSq = _textbook_state(.7, (.3, -.2, .1), .9)
    assert e[0] == rho
    assert np.allclose(e[1:4], v, rtol=1e-15, atol=1e-16) and np.allclose(e[7:10], v, rtol=1e-15, atol=1e-16)
    assert e[4] == pytest.approx(H, rel=1e-14)
    assert e[5] == pytest.approx(Cs, rel=1e-14)
    assert e[6] == pytest.approx(vSq, rel=1e-14)
    # the averaged Roe state would differ: sqrt(rho_L rho_R) != .7
    assert abs(np.sqrt(vac_rho * .7) - .7) > .1


def test_eigen_both_sides_vacuum_is_zero():
    e = _eig([1e-6, .1, .2, 0., 1., 0.], [9e-6, -1., 0., 3., 2., 0.])
    assert np.array_equal(e, np.zeros(10))


def test_eigen_low_enthalpy_branch():
    """h = hTotal - v^2 / 2 < 1e-5 (a cold gas): Cs = 0 and hTotal = v^2 / 2
    (euler.cl:410-420); just above, Cs = sqrt((gamma - 1) h)"""
    U = _cons(1., (.5, .2, 0.), 1e-8)           # h = gamma P / ((gamma - 1) rho) = 3.5e-8
    e = _eig(U, U)
    assert e[5] == 0.
    assert e[4] == pytest.approx(.5 * (.25 + .04), rel=1e-15)
    # keeping hTotal there would add h = 3.5e-8
    assert abs(e[4] - (.5 * .29 + 3.5e-8)) > 1e-8
    U = _cons(1., (.5, .2, 0.), 1e-5)           # h = 3.5e-5
    e = _eig(U, U)
    assert e[5] == pytest.approx(np.sqrt((GAMMA - 1) * 3.5e-5), rel=1e-12)
    assert e[4] == pytest.approx(.5 * .29 + 3.5e-5, rel=1e-14)


def _apply(fn, e, X, side):
    Y = np.zeros(6)
    fn(ctypes.byref(P), _ptr(np.asarray(e, float)), _ptr(np.asarray(X, float)), side, _ptr(Y))
    return Y[:5]


@pytest.mark.parametrize('side', [0, 1, 2])
def test_transforms_identity_below_rho_min(side):
    """eig.rho < rhoMin: left and right transforms are the identity; at
    eig.rho = rhoMin they are the eigenvector maps (L R = I, not X)"""
    X = np.array([.3, -1.2, .7, 2., -.4, 0.])
    Cs = np.sqrt((GAMMA - 1) * (2. - .5 * .14))   # consistent with hTotal = 2, v^2 = .14
    e = np.array([5e-8, .3, .1, -.2, 2., Cs, .14, .3, .1, -.2])
    assert np.array_equal(_apply(lib().oracle_euler_left, e, X, side), X[:5])
    assert np.array_equal(_apply(lib().oracle_euler_right, e, X, side), X[:5])
    e[0] = RHO_MIN
    L = _apply(lib().oracle_euler_left, e, X, side)
    assert np.abs(L - X[:5]).max() > .1
    assert np.allclose(_apply(lib().oracle_euler_right, e, np.r_[L, 0.], side), X[:5], rtol=1e-13, atol=1e-13)


def _flux(U, side):
    F = np.zeros(6)
    lib().oracle_euler_flux_from_cons(ctypes.byref(P), _ptr(np.asarray(U, float)), side, _ptr(F))
    return F[:5]


def _roe(UL, UR, side):
    F = np.zeros(6)
    lib().oracle_roe_flux(ctypes.byref(P), 0, 0, 0., _ptr(np.asarray(UL, float)), _ptr(np.asarray(UR, float)), None,
                          side, _ptr(F))
    return F[:5]


def _jacobian(rho, v, p, side):
    """textbook dF/dU of the Euler flux along axis `side` at one state"""
    _, v, H, _, vsq = _textbook_state(rho, v, p)
    n = np.zeros(3)
    n[side] = 1
    vn = v[side]
    g = GAMMA
    A = np.zeros((5, 5))
    A[0, 1:4] = n
    for i in range(3):
        A[1 + i, 0] = -v[i] * vn + .5 * (g - 1) * vsq * n[i]
        A[1 + i, 1:4] = v[i] * n - (g - 1) * n[i] * v
        A[1 + i, 1 + i] += vn
        A[1 + i, 4] = (g - 1) * n[i]
    A[4, 0] = vn * (.5 * (g - 1) * vsq - H)
    A[4, 1:4] = H * n - (g - 1) * vn * v
    A[4, 4] = g * vn
    return A


@pytest.mark.parametrize('side', [0, 1, 2])
def test_roe_flux_vacuum_interfaces(side):
    """both sides below 1e-5: eig = 0, so lambda = 0 and the transforms are the
    identity, leaving the central flux (F(UL) + F(UR)) / 2. One side below:
    the dissipation is |A| (UR - UL) with the Jacobian of the other side's own
    state (numpy's eigendecomposition of the textbook Jacobian)"""
    UL = _cons(1e-6, (.2, 0, 0), 1e-6)
    UR = _cons(2e-6, (-.1, .3, 0), 3e-6)
    assert np.allclose(_roe(UL, UR, side), .5 * (_flux(UL, side) + _flux(UR, side)), rtol=1e-15, atol=1e-22)
    good = (.8, (.4, -.3, .2), 1.3)
    Ug = _cons(*good)
    Uv = _cons(3e-6, (1., 2., 0), 1e-6)
    A = _jacobian(*good, side)
    lam, V = np.linalg.eig(A)
    absA = (V @ np.diag(np.abs(lam)) @ np.linalg.inv(V)).real
    for UL, UR in ((Uv, Ug), (Ug, Uv)):
        want = .5 * (_flux(UL, side) + _flux(UR, side)) - .5 * absA @ (UR - UL)[:5]
        assert np.allclose(_roe(UL, UR, side), want, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------ divergence scale
def _stab(integ):
    return {'forward Euler': lambda z: 1 + z, 'Runge-Kutta 2, TVD': lambda z: 1 + z + z * z / 2,
            'Runge-Kutta 3, TVD': lambda z: 1 + z + z * z / 2 + z ** 3 / 6}[integ]


def _wave_run(dim, size, maxs, k_modes, v, integ='Runge-Kutta 3, TVD', dt=.004):
    """a small density wave rho = 1 + eps cos(k . x) advected at uniform v and
    P on a periodic grid, one step; returns the measured factor of the mode and
    the cell sizes"""
    cfg = make_cfg(dim=dim, size=size, mins=(0., 0., 0.)[:dim], maxs=maxs, integrator=integ, boundary='periodic',
                   fixed_dt=dt, cfl=.5)
    o = pyoracle.Oracle(cfg)
    x, y, z = init_euler.mesh(cfg)
    coords = [x, y, z]
    phase = np.zeros(init_euler.interior_shape(cfg))
    for a in range(dim):
        phase = phase + 2 * np.pi * k_modes[a] / maxs[a] * coords[a]
    eps = 1e-6
    rho = 1 + eps * np.cos(phase)
    vel = [v[a] + 0 * rho for a in range(3)]
    o.set_U(init_euler.to_ghosted(cfg, init_euler.cons_from_prim(rho, vel, 1. + 0 * rho, GAMMA)))
    o.boundary()
    idx = tuple(k_modes[a] if a < dim else 0 for a in (2, 1, 0))   # (z, y, x) FFT indices
    fft = lambda: np.fft.fftn(o.interior()[..., 0] - 1)[idx]
    a0 = fft()
    o.step(dt)
    dx = [maxs[a] / size[a] for a in range(dim)]
    return fft() / a0, dx


def _symbol(v, k_modes, maxs, dx, scale=None):
    """Fourier symbol of first-order upwinding summed over the axes:
    -sum_a (v_a / s_a)(1 - exp(-i k_a dx_a)) for v_a >= 0, where s_a is the
    width the divergence divides by (dx_a when right)"""
    scale = dx if scale is None else scale
    lam = 0j
    for a in range(len(dx)):
        k = 2 * np.pi * k_modes[a] / maxs[a]
        lam += -(v[a] / scale[a]) * (1 - np.exp(-1j * k * dx[a]))
    return lam


@pytest.mark.parametrize('dim,size,maxs,k_modes,v', [
    (2, (6, 40), (1.5, .7), (0, 3), (0., 1., 0.)),            # along y, dx != dy
    (2, (40, 6), (.7, 1.5), (3, 0), (1., 0., 0.)),            # along x
    (2, (32, 24), (1.1, .6), (2, 3), (.7, .5, 0.)),           # oblique: both axes' terms
    (3, (4, 5, 36), (1., .8, 1.3), (0, 0, 4), (0., 0., 1.)),  # along z, dx != dy != dz
    (3, (4, 30, 5), (1., .9, 1.3), (0, 2, 0), (0., 1., 0.)),  # along y in 3D
    (3, (20, 16, 12), (1., .72, .6), (1, 2, 1), (.6, .5, .4)), # oblique in 3D
])
def test_divergence_scale_per_axis(dim, size, maxs, k_modes, v):
    """calcDerivFromFlux divides each axis's flux difference by that axis's
    cell width (areaL / volume = 1 / dx_a for a Cartesian grid): one RK3-TVD
    step multiplies the mode by R(lambda dt) with the symbol of the axes'
    own widths. A swap of two axes' widths changes the factor by far more than
    the tolerance."""
    dt = .004
    got, dx = _wave_run(dim, size, maxs, k_modes, v, dt=dt)
    R = _stab('Runge-Kutta 3, TVD')
    want = R(_symbol(v, k_modes, maxs, dx) * dt)
    assert got == pytest.approx(want, rel=1e-6, abs=1e-9)
    moving = [a for a in range(dim) if v[a] != 0]
    for a in moving:
        for b in range(dim):
            if b != a and abs(dx[b] - dx[a]) > 1e-3 * dx[a]:
                scale = list(dx)
                scale[a] = dx[b]
                wrong = R(_symbol(v, k_modes, maxs, dx, scale) * dt)
                assert abs(wrong - want) > 1e3 * 1e-6 * abs(want)
