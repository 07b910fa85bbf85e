"""The plugin surfaces of the reference kept by the drop-in (hydro/eqn,
hydro/flux, hydro/int): names, state counts, and the integrator tables the
library runs, pinned by their stability polynomials on u' = lambda u."""
import numpy as np
import pytest

from hydro import eqn, flux, int as integ
from hydro._cfg import INTEGRATORS, EQN_NUM_STATES, EQN_NUM_INT_STATES
from hydro import _abi


def test_equations_match_library():
    for name in ('euler', 'glm_maxwell'):
        e = eqn.get(name)
        lib_index = ['euler', 'glm_maxwell'].index(name)
        assert e.numStates == EQN_NUM_STATES[name] == _abi.lib().hydro_eqn_num_states(lib_index)
        assert e.numIntStates == EQN_NUM_INT_STATES[name] == _abi.lib().hydro_eqn_num_int_states(lib_index)
        assert sum(3 if t == 'real3' else 1 for _, t in e.consVars) == e.numStates


def test_euler_prim_cons_roundtrip():
    e = eqn.get('euler')
    rng = np.random.default_rng(3)
    rho, P = rng.uniform(.1, 2, 50), rng.uniform(.1, 2, 50)
    v = rng.uniform(-1, 1, (3, 50))
    U = e.consFromPrim(rho, v, P)
    r2, v2, P2 = e.primFromCons(U)
    assert np.allclose(r2, rho, rtol=1e-15) and np.allclose(P2, P, rtol=1e-13)
    assert np.allclose(np.array(v2), v, rtol=1e-14)


def test_flux_plugin():
    f = flux.get('roe')
    assert f.name == 'roe' and f.usesFluxLimiter and not f.useEntropyFluxFix
    h, r = flux.get('hll'), flux.get('rusanov')
    assert h.name == 'hll' and not h.usesFluxLimiter and h.hllCalcWaveMethod == 'Davis direct bounded'
    assert r.name == 'rusanov' and not r.usesFluxLimiter
    with pytest.raises(ValueError):
        flux.get('hllc')


def _R(alphas, betas, z):
    u = [1. + 0j]
    for a, b in zip(alphas, betas):
        u.append(sum((a[k] + z * b[k]) * u[k] for k in range(len(u))))
    return u[-1]


# textbook stability polynomials; the reference's 'Runge-Kutta 3' and
# 'Runge-Kutta 4, 3/8ths rule' tables are not the textbook methods
# (int/all.lua:50-54, 79-84) and are checked entry by entry below
POLY = {1: lambda z: 1 + z, 2: lambda z: 1 + z + z * z / 2, 3: lambda z: 1 + z + z * z / 2 + z ** 3 / 6,
        4: lambda z: 1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24}
ORDER = {'forward Euler': 1, 'Runge-Kutta 2': 2, 'Runge-Kutta 2 Heun': 2, 'Runge-Kutta 2 Ralston': 2,
         'Runge-Kutta 4': 4, 'Runge-Kutta 2, TVD': 2, 'Runge-Kutta 2, non-TVD': 2, 'Runge-Kutta 3, TVD': 3,
         'Runge-Kutta 4, TVD': 4, 'Runge-Kutta 4, non-TVD': 4}


@pytest.mark.parametrize('name', INTEGRATORS)
def test_integrator_tables(name):
    it = integ.get(name)
    assert it.name == name
    a, b = it.table()
    if name in ORDER:
        for z in (.3, -.7 + .2j, 1.1j):
            assert abs(_R(a, b, z) - POLY[ORDER[name]](z)) < 1e-13
    elif name == 'Runge-Kutta 3':
        assert b == [[.5, 0, 0], [-1, 2, 0], [1 / 6, 2 / 6, 1 / 6]] and a == [[1, 0, 0]] * 3
    else:
        assert b == [[1 / 3, 0, 0, 0], [-1 / 3, 0, 0, 0], [1, -1, 1, 0], [1 / 8, 3 / 8, 3 / 8, 1 / 8]]
        assert a == [[1, 0, 0, 0]] * 4
    # explicit: stage i uses only earlier stages
    for i in range(len(a)):
        assert all(a[i][k] == 0 and b[i][k] == 0 for k in range(i + 1, len(a)))


def test_unknown_integrator_rejected():
    with pytest.raises(_abi.HydroError):
        _abi.check(_abi.lib().hydro_integrator_table(99, None, None, None))
