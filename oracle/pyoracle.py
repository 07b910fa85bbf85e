"""ctypes wrapper of the CPU oracle — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REPO = os.path.dirname(_HERE)
_PKG = os.path.join(_REPO, 'hydro-cl-lua_amd')
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from hydro._cfg import HydroCfg, make_cfg, cfg_ghost_shape  # noqa: E402

LIB_PATH = os.path.join(_HERE, 'build', 'libhydro_oracle.so')
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class OracleParams(ctypes.Structure):
    _fields_ = [('gamma', ctypes.c_double), ('rho_min', ctypes.c_double),
                ('P_min', ctypes.c_double), ('div_phi', ctypes.c_double),
                ('div_psi', ctypes.c_double), ('roe_fix', ctypes.c_int),
                ('roe_fix_factor', ctypes.c_double), ('gravity', ctypes.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'oracle not built: {LIB_PATH} (run __graft_entry__.build())')
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_create.restype = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.POINTER(HydroCfg)]
        L.oracle_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_num_cells.restype = ctypes.c_size_t
        L.oracle_num_cells.argtypes = [ctypes.c_void_p]
        L.oracle_num_states.argtypes = [ctypes.c_void_p]
        L.oracle_U.restype = _dp
        L.oracle_U.argtypes = [ctypes.c_void_p]
        for f in ('oracle_boundary', 'oracle_constrain_U'):
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.oracle_calc_dt.restype = ctypes.c_double
        L.oracle_calc_dt.argtypes = [ctypes.c_void_p]
        L.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oracle_update.restype = ctypes.c_double
        L.oracle_update.argtypes = [ctypes.c_void_p]
        L.oracle_time.restype = ctypes.c_double
        L.oracle_time.argtypes = [ctypes.c_void_p]
        L.oracle_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oracle_calc_deriv.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        L.oracle_limiter.restype = ctypes.c_double
        L.oracle_limiter.argtypes = [ctypes.c_int, ctypes.c_double]
        pp = ctypes.POINTER(OracleParams)
        for f in ('oracle_euler_prim_from_cons', 'oracle_euler_cons_from_prim'):
            getattr(L, f).argtypes = [pp, _dp, _dp]
        for f in ('oracle_euler_eigen_for_interface', 'oracle_glm_eigen_for_interface'):
            getattr(L, f).argtypes = [pp, _dp, _dp, _dp]
        for f in ('oracle_euler_left', 'oracle_euler_right', 'oracle_glm_left', 'oracle_glm_right'):
            getattr(L, f).argtypes = [pp, _dp, _dp, ctypes.c_int, _dp]
        for f in ('oracle_euler_flux_from_cons', 'oracle_glm_flux_from_cons',
                  'oracle_euler_waves', 'oracle_glm_waves'):
            getattr(L, f).argtypes = [pp, _dp, ctypes.c_int, _dp]
        L.oracle_euler_dt_cell.restype = ctypes.c_double
        L.oracle_euler_dt_cell.argtypes = [pp, _dp, ctypes.c_int, _dp]
        L.oracle_euler_constrain.argtypes = [pp, _dp]
        L.oracle_roe_flux.argtypes = [pp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                      _dp, _dp, ctypes.POINTER(_dp), ctypes.c_int, _dp]
        L.oracle_plm_cons.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_var_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.oracle_euler_display.restype = ctypes.c_double
        L.oracle_euler_display.argtypes = [pp, ctypes.c_int, _dp]
        L.oracle_plm_prim.argtypes = [pp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.oracle_plm_eig.argtypes = [pp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_plm_cons_flux.argtypes = L.oracle_plm_eig.argtypes
        L.oracle_plm_athena.argtypes = [pp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.oracle_plm_ppm.argtypes = L.oracle_plm_cons.argtypes
        L.oracle_plm_eig_prim.argtypes = [pp, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp]
        for f in ('oracle_hll_flux', 'oracle_rusanov_flux', 'oracle_hllc_flux', 'oracle_marquina_flux'):
            getattr(L, f).argtypes = [pp, _dp, _dp, ctypes.c_int, _dp]
        L.oracle_sw_prim_from_cons.argtypes = [pp, _dp, _dp]
        L.oracle_sw_eigen_for_interface.argtypes = [pp, _dp, _dp, _dp]
        for f in ('oracle_sw_left', 'oracle_sw_right'):
            getattr(L, f).argtypes = [pp, _dp, _dp, ctypes.c_int, _dp]
        for f in ('oracle_sw_flux_from_cons', 'oracle_sw_waves'):
            getattr(L, f).argtypes = [pp, _dp, ctypes.c_int, _dp]
        L.oracle_sw_dt_cell.restype = ctypes.c_double
        L.oracle_sw_dt_cell.argtypes = [pp, _dp, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def params(gamma=1.4, rho_min=1e-7, P_min=1e-7, div_phi=1.0, div_psi=1.0, roe_fix=0, roe_fix_factor=.2,
           gravity=1.0):
    return OracleParams(gamma, rho_min, P_min, div_phi, div_psi, roe_fix, roe_fix_factor, gravity)


class Oracle:
    """FP64 CPU restatement of the reference solver, ghosted AoS state."""

    def __init__(self, cfg=None, **kw):
        self.cfg = cfg if cfg is not None else make_cfg(**kw)
        L = lib()
        self.h = L.oracle_create(ctypes.byref(self.cfg))
        if not self.h:
            raise ValueError(L.oracle_last_error().decode())
        self.shape = cfg_ghost_shape(self.cfg)
        n = int(np.prod(self.shape))
        self.U = np.ctypeslib.as_array(L.oracle_U(self.h), shape=(n,)).reshape(self.shape)

    def __del__(self):
        if getattr(self, 'h', None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_U(self, U):
        self.U[...] = np.asarray(U, dtype=np.float64).reshape(self.shape)

    def boundary(self):
        lib().oracle_boundary(self.h)

    def constrain_U(self):
        lib().oracle_constrain_U(self.h)

    def calc_dt(self):
        return lib().oracle_calc_dt(self.h)

    def step(self, dt):
        lib().oracle_step(self.h, dt)

    def update(self, n=1):
        dt = None
        for _ in range(n):
            dt = lib().oracle_update(self.h)
        return dt

    @property
    def t(self):
        return lib().oracle_time(self.h)

    @t.setter
    def t(self, v):
        lib().oracle_set_time(self.h, v)

    def calc_deriv(self, dt):
        d = np.zeros(self.shape)
        lib().oracle_calc_deriv(self.h, dt, _ptr(d))
        return d

    def interior(self, U=None):
        U = self.U if U is None else U
        d = self.cfg.dim
        sl = [slice(None)] * 4
        for i in range(d):
            sl[2 - i] = slice(2, -2)
        return U[tuple(sl)]


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return lib().oracle_get_threads()
