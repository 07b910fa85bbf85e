"""CPU-side checks of the C ABI: the library loads and exports every symbol
include/hydro_hip.h declares. No compute call is made."""
import ctypes
import os
import re

from hydro import _abi
from hydro._cfg import HydroCfg, make_cfg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(REPO, 'include', 'hydro_hip.h')).read()
    return sorted(set(re.findall(r'\b(hydro_[A-Za-z0-9_]+)\s*\(', txt)))


def test_library_exports_header_symbols():
    L = _abi.lib()
    syms = _header_symbols()
    assert syms, 'no declarations found'
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_abi.EXPORTS) == syms


def test_cfg_struct_matches_default():
    L = _abi.lib()
    c = HydroCfg()
    L.hydro_cfg_default(ctypes.byref(c))
    assert c.dim == 1 and c.eqn == 0 and c.cfl == .5 and abs(c.gamma - 1.4) < 1e-15
    assert list(c.boundary) == [2] * 6
    assert c.rho_min == 1e-7 and c.P_min == 1e-7


def test_eqn_sizes():
    L = _abi.lib()
    assert L.hydro_eqn_num_states(0) == 6 and L.hydro_eqn_num_int_states(0) == 5
    assert L.hydro_eqn_num_states(1) == 12 and L.hydro_eqn_num_int_states(1) == 8


def test_create_rejects_plm_with_flux_limiter():
    """gridsolver.lua:119: PLM and a flux limiter together are an error"""
    L = _abi.lib()
    c = make_cfg(dim=2, size=(8, 8), plm='plm cons', flux_limiter='superbee')
    h = ctypes.c_void_p()
    rc = L.hydro_create(ctypes.byref(c), ctypes.byref(h))
    assert rc != 0
    assert b'flux and slope limiters' in L.hydro_last_error()


def test_create_rejects_plm_kinds_outside_euler_roe():
    """'plm prim' and 'plm eig' are built for the Euler equations with the Roe
    flux; other pairings are refused before any device work"""
    L = _abi.lib()
    for kw in (dict(eqn='glm_maxwell', plm='plm eig'), dict(flux='hll', plm='plm eig'),
               dict(eqn='glm_maxwell', plm='plm cons with flux'), dict(flux='marquina', plm='plm cons with flux'),
               dict(flux='hll', plm='ppm'), dict(flux='rusanov', plm='plm athena'),
               dict(eqn='glm_maxwell', plm='plm eig prim'), dict(flux='euler-hllc', plm='plm eig prim ref'),
               dict(plm='plm eig', use_entropy_flux_fix=True), dict(eqn='glm_maxwell', plm='plm prim')):
        c = make_cfg(dim=2, size=(8, 8), **kw)
        h = ctypes.c_void_p()
        assert L.hydro_create(ctypes.byref(c), ctypes.byref(h)) != 0, kw
        assert b'built for the Euler equations' in L.hydro_last_error() or b'entropy fix' in L.hydro_last_error()


def test_piecewise_constant_is_no_reconstruction():
    """plm.cl:10-24: 'piecewise constant' edge states are the cell's own, so
    the configuration is plm None; a flux limiter with it is refused as
    gridsolver.lua:119 does"""
    import pytest
    c = make_cfg(dim=2, size=(8, 8), plm='piecewise constant')
    assert c.plm == 0 and c.flux_limiter == 0
    with pytest.raises(ValueError, match='flux and slope limiters'):
        make_cfg(dim=2, size=(8, 8), plm='piecewise constant', flux_limiter='superbee')
