"""Pins the oracle's shallow water equations (hydro/eqn/shallow-water.cl with
depthInCell = false, extraTermInFlux = false) by the mathematics:

- L R = I and R L = I on random Roe states, every axis;
- R Lambda L = the analytic flux Jacobian at the Roe state
  (A = dF/dU with v~ and c~^2 = g h~);
- the Roe property R Lambda L (UR - UL) = F(UR) - F(UL), exact for the
  arithmetic-mean height and sqrt(h)-weighted velocity that
  eigen_forInterface forms (shallow-water.cl:228-265);
- wave speeds v_n -+ sqrt(g h~) and the Roe flux's consistency F(U, U) = F(U);
- calcDT's closed form;
- the depth source g h d(depth)/dx on a linear depth, where the central
  difference is exact, and its absence for a flat depth;
- a dam break against the exact solution of the shallow water Riemann
  problem (rarefaction + shock, Stoker), converging under refinement.

A planted mistake (a tangential term of the right transform, the sign of the
source, the mean height replaced by the geometric one) fails one of them.
"""
import ctypes

import numpy as np
import pytest

import pyoracle
from pyoracle import lib, params, _ptr
from hydro._cfg import make_cfg
from hydro.init import shallow_water as init_sw
from hydro.init import euler as init_euler

G = .7
P = params(gravity=G)


def _state(rng, dim=3):
    h = rng.uniform(.2, 2.)
    v = rng.uniform(-1, 1, 3)
    v[dim:] = 0
    return np.array([h, *(v * h), rng.uniform(-.5, .5)])


def _eig(UL, UR):
    e = np.zeros(6)
    lib().oracle_sw_eigen_for_interface(P, _ptr(UL), _ptr(UR), _ptr(e))
    return e


def _apply(f, e, X, side):
    Y = np.zeros(5)
    f(P, _ptr(e), _ptr(np.asarray(X, dtype=np.float64)), side, _ptr(Y))
    return Y[:4]


def _mat(f, e, side):
    M = np.zeros((4, 4))
    for j in range(4):
        X = np.zeros(5)
        X[j] = 1
        M[:, j] = _apply(f, e, X, side)
    return M


def _flux(U, side):
    F = np.zeros(5)
    lib().oracle_sw_flux_from_cons(P, _ptr(np.asarray(U, dtype=np.float64)), side, _ptr(F))
    return F[:4]


def _jacobian(h, v, side):
    """dF/dU of the shallow water flux along `side` at height h, velocity v:
    F_h = m_s, F_mi = m_i m_s / h + delta_is g h^2 / 2"""
    A = np.zeros((4, 4))
    A[0, 1 + side] = 1
    for i in range(3):
        A[1 + i, 0] = -v[i] * v[side] + (G * h if i == side else 0)
        A[1 + i, 1 + i] += v[side]
        A[1 + i, 1 + side] += v[i]
    return A


@pytest.mark.parametrize('side', [0, 1, 2])
def test_left_right_inverse(side):
    rng = np.random.default_rng(10 + side)
    for _ in range(20):
        e = _eig(_state(rng), _state(rng))
        L = _mat(lib().oracle_sw_left, e, side)
        R = _mat(lib().oracle_sw_right, e, side)
        assert np.allclose(L @ R, np.eye(4), atol=1e-13)
        assert np.allclose(R @ L, np.eye(4), atol=1e-13)


@pytest.mark.parametrize('side', [0, 1, 2])
def test_eigensystem_is_flux_jacobian(side):
    rng = np.random.default_rng(20 + side)
    for _ in range(20):
        e = _eig(_state(rng), _state(rng))
        lam = np.zeros(4)
        lib().oracle_sw_waves(P, _ptr(e), side, _ptr(lam))
        c = np.sqrt(G * e[0])
        assert np.allclose(lam, [e[1 + side] - c, e[1 + side], e[1 + side], e[1 + side] + c], rtol=0, atol=1e-15)
        L = _mat(lib().oracle_sw_left, e, side)
        R = _mat(lib().oracle_sw_right, e, side)
        assert np.allclose(R @ np.diag(lam) @ L, _jacobian(e[0], e[1:4], side), atol=1e-13)


@pytest.mark.parametrize('side', [0, 1, 2])
def test_roe_property(side):
    rng = np.random.default_rng(30 + side)
    for _ in range(30):
        UL, UR = _state(rng), _state(rng)
        e = _eig(UL, UR)
        lam = np.zeros(4)
        lib().oracle_sw_waves(P, _ptr(e), side, _ptr(lam))
        dW = _apply(lib().oracle_sw_left, e, UR - UL, side)
        X = np.zeros(5)
        X[:4] = lam * dW
        dF = _apply(lib().oracle_sw_right, e, X, side)
        assert np.allclose(dF, _flux(UR, side) - _flux(UL, side), rtol=0, atol=1e-13)


def test_roe_average_closed_form():
    UL = np.array([.5, .3, -.2, .1, .0])
    UR = np.array([2., -.4, .6, .2, .0])
    e = _eig(UL, UR)
    sl, sr = np.sqrt(.5), np.sqrt(2.)
    vL, vR = UL[1:4] / UL[0], UR[1:4] / UR[0]
    assert e[0] == pytest.approx(1.25, abs=0)               # arithmetic mean height
    assert np.allclose(e[1:4], (sl * vL + sr * vR) / (sl + sr), rtol=0, atol=1e-15)
    assert e[4] == pytest.approx(np.sqrt(G * 1.25), rel=1e-15)


@pytest.mark.parametrize('side', [0, 1, 2])
def test_roe_flux_consistency(side):
    rng = np.random.default_rng(40 + side)
    for _ in range(10):
        U = _state(rng)
        F = np.zeros(5)
        lib().oracle_roe_flux(P, 2, 0, 0., _ptr(U), _ptr(U), None, side, _ptr(F))
        assert np.allclose(F[:4], _flux(U, side), rtol=0, atol=1e-15)


def test_flux_closed_form():
    U = np.array([1.5, .3, -.6, .9, .2])
    v = U[1:4] / U[0]
    for side in range(3):
        want = np.array([U[1 + side], *(U[1:4] * v[side])])
        want[1 + side] += .5 * G * U[0] ** 2
        assert np.allclose(_flux(U, side), want, rtol=1e-15, atol=0)


def test_dt_cell_closed_form():
    U = np.array([1.5, .3, -.6, .9, .2])
    dx = np.array([.1, .05, .2])
    c = np.sqrt(G * 1.5)
    v = U[1:4] / U[0]
    want = min(dx[s] / (abs(v[s]) + c) for s in range(3))
    got = lib().oracle_sw_dt_cell(P, _ptr(U), 3, _ptr(dx))
    assert got == pytest.approx(want, rel=1e-15)


def _oracle(dim, n, init, **kw):
    cfg = make_cfg(dim=dim, size=n, eqn='shallow_water', gravity=G, **kw)
    o = pyoracle.Oracle(cfg)
    o.set_U(init_euler.to_ghosted(cfg, init(cfg)))
    o.boundary()
    return o, cfg


@pytest.mark.parametrize('dim', [1, 2])
def test_depth_source_on_linear_depth(dim):
    """still water of uniform height on depth = a x + b y: every flux
    difference is 0, and the momentum derivative is g h (a, b) (the central
    difference of a linear function is its slope); flat depth: 0"""
    a, b, h = .3, -.7, 1.2

    def init(cfg):
        x, y, z = init_euler.mesh(cfg)
        shape = init_euler.interior_shape(cfg)
        zero = np.zeros(shape)
        return init_sw.cons_from_prim(h + zero, (zero, zero, zero), a * x + b * y + zero)
    n = 16 if dim == 1 else (16, 12)
    o, cfg = _oracle(dim, n, init, boundary='freeflow')
    # the ghosts of a linear depth: extend it linearly so every interior cell
    # sees the exact slope
    U = o.U
    gs = U.shape
    for k in range(gs[1]):
        for i in range(gs[2]):
            xx = (i - 2 + .5) / cfg.size[0] * 2 - 1
            yy = (k - 2 + .5) / cfg.size[1] * 2 - 1 if dim > 1 else 0.
            U[0, k, i, 4] = a * xx + b * yy
    d = o.calc_deriv(.01)
    inner = o.interior(d)
    assert np.allclose(inner[..., 0], 0, atol=1e-14)
    assert np.allclose(inner[..., 1], G * h * a, rtol=1e-12, atol=0)
    if dim > 1:
        assert np.allclose(inner[..., 2], G * h * b, rtol=1e-12, atol=0)
    else:
        assert np.all(inner[..., 2] == 0)
    # flat depth: still water stays still
    o2, _ = _oracle(dim, n, lambda cfg: init_sw.cons_from_prim(
        h + np.zeros(init_euler.interior_shape(cfg)), [np.zeros(init_euler.interior_shape(cfg))] * 3,
        np.zeros(init_euler.interior_shape(cfg))), boundary='freeflow')
    assert np.all(np.abs(o2.interior(o2.calc_deriv(.01))[..., :4]) < 1e-15)


def _stoker(x, t, hl, hr, g):
    """exact solution of the dam break (still water hl | hr, hl > hr > 0): a
    rarefaction into the left column and a shock into the right one"""
    cl, cr = np.sqrt(g * hl), np.sqrt(g * hr)

    def f(hs):   # u* from the left (rarefaction) minus u* from the right (shock)
        ul = 2 * (cl - np.sqrt(g * hs))
        ur = (hs - hr) * np.sqrt(.5 * g * (hs + hr) / (hs * hr))
        return ul - ur
    lo, hi = hr, hl
    for _ in range(200):
        m = .5 * (lo + hi)
        if f(m) > 0:
            lo = m
        else:
            hi = m
    hs = .5 * (lo + hi)
    us = 2 * (cl - np.sqrt(g * hs))
    cs = np.sqrt(g * hs)
    shock = hs * us / (hs - hr)
    out = np.empty_like(x)
    for i, xi in enumerate(x):
        s = xi / t
        if s < -cl:
            out[i] = hl
        elif s < us - cs:
            out[i] = ((2 * cl - s) / 3) ** 2 / g
        elif s < shock:
            out[i] = hs
        else:
            out[i] = hr
    return out


def test_dam_break_against_exact_solution():
    """1D dam break h = 1 | .5, Roe + plm cons / minmod, RK2-TVD: L1(h) vs the
    Stoker solution is small and halves (at least) from 200 to 800 cells"""
    errs = []
    for n in (200, 800):
        o, cfg = _oracle(1, n, lambda cfg: init_sw.dam_break(cfg, 1., .5), plm='plm cons', slope_limiter='minmod',
                         integrator='Runge-Kutta 2, TVD', cfl=.5, boundary='freeflow')
        while o.t < .3:
            o.update(1)
        x = (np.arange(n) + .5) / n * 2 - 1
        h = o.interior()[0, 0, :, 0]
        errs.append(np.abs(h - _stoker(x, o.t, 1., .5, G)).mean() * 2)
    assert errs[0] < 5e-3
    assert errs[1] < .5 * errs[0]


def test_mirror_reflects_normal_momentum():
    cfg = make_cfg(dim=2, size=(4, 3), eqn='shallow_water', boundary='mirror')
    o = pyoracle.Oracle(cfg)
    rng = np.random.default_rng(3)
    U = init_euler.to_ghosted(cfg, rng.uniform(.5, 1.5, (1, 3, 4, 5)))
    o.set_U(U)
    o.boundary()
    g = o.U[0]
    assert np.array_equal(g[2, 1, :], np.array([1, -1, 1, 1, 1]) * g[2, 2, :])   # x face: m.x flips
    assert np.array_equal(g[1, 3, :], np.array([1, 1, -1, 1, 1]) * g[2, 3, :])   # y face: m.y flips
