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
