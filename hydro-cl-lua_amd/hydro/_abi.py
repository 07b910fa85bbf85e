"""ctypes binding of libhydro_hip.so (include/hydro_hip.h).

The product path: every call goes to the HIP library. Loading fails loudly when
the library is missing; there is no CPU fallback.
"""
import ctypes
import os

from ._cfg import HydroCfg

LIB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'lib')
LIB_PATH = os.path.join(LIB_DIR, 'libhydro_hip.so')

# every symbol include/hydro_hip.h declares
EXPORTS = [
    'hydro_cfg_default', 'hydro_last_error', 'hydro_eqn_num_states', 'hydro_eqn_num_int_states',
    'hydro_integrator_table',
    'hydro_device_count', 'hydro_create', 'hydro_destroy', 'hydro_num_cells', 'hydro_upload_U',
    'hydro_download_U', 'hydro_constrain_U', 'hydro_boundary', 'hydro_calc_dt', 'hydro_step',
    'hydro_update', 'hydro_get_time', 'hydro_set_time', 'hydro_conserved_totals', 'hydro_conserved_totals_device',
    'hydro_var_stats', 'hydro_sync',
    'hydro_time_updates', 'hydro_device_bytes', 'hydro_set_launch', 'hydro_get_launch',
    'hydro_batch_create', 'hydro_batch_destroy', 'hydro_batch_size', 'hydro_batch_device_bytes',
    'hydro_batch_upload_U', 'hydro_batch_download_U', 'hydro_batch_constrain_U', 'hydro_batch_update',
    'hydro_batch_get_time', 'hydro_batch_set_active', 'hydro_batch_sync', 'hydro_batch_time_updates',
    'hydro_batch_set_launch', 'hydro_batch_get_launch', 'hydro_batch_autotune',
    'hydro_server_create', 'hydro_server_destroy', 'hydro_server_submit', 'hydro_server_tick',
    'hydro_server_poll', 'hydro_server_result', 'hydro_server_info_get', 'hydro_server_set_launch',
]


class HydroServerInfo(ctypes.Structure):
    _fields_ = [('queued', ctypes.c_int), ('resident', ctypes.c_int), ('finished', ctypes.c_int),
                ('pools', ctypes.c_int), ('bytes_in_use', ctypes.c_size_t), ('budget', ctypes.c_size_t),
                ('ticks', ctypes.c_longlong)]


class HydroTiming(ctypes.Structure):
    _fields_ = [('total_ms', ctypes.c_double), ('stage_ms', ctypes.c_double),
                ('n_stage_launches', ctypes.c_int), ('dt_ms', ctypes.c_double),
                ('n_dt_launches', ctypes.c_int)]


class HydroError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HydroError(f'{LIB_PATH} is missing: build it with __graft_entry__.build() '
                         '(the HIP path has no fallback)')
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, ip, dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    cfgp = ctypes.POINTER(HydroCfg)
    L.hydro_cfg_default.argtypes = [cfgp]
    L.hydro_cfg_default.restype = None
    L.hydro_last_error.restype = ctypes.c_char_p
    L.hydro_eqn_num_states.argtypes = [ip]
    L.hydro_eqn_num_int_states.argtypes = [ip]
    L.hydro_integrator_table.argtypes = [ip, ctypes.POINTER(ip), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_double)]
    L.hydro_device_count.restype = ip
    L.hydro_create.argtypes = [cfgp, ctypes.POINTER(vp)]
    L.hydro_destroy.argtypes = [vp]
    L.hydro_num_cells.argtypes = [vp]
    L.hydro_num_cells.restype = sz
    L.hydro_device_bytes.argtypes = [vp]
    L.hydro_device_bytes.restype = sz
    L.hydro_upload_U.argtypes = [vp, vp, sz]
    L.hydro_download_U.argtypes = [vp, vp, sz]
    for f in ('hydro_constrain_U', 'hydro_boundary', 'hydro_sync'):
        getattr(L, f).argtypes = [vp]
    L.hydro_calc_dt.argtypes = [vp, dp]
    L.hydro_step.argtypes = [vp, ctypes.c_double]
    L.hydro_update.argtypes = [vp, ip]
    L.hydro_get_time.argtypes = [vp, dp, dp]
    L.hydro_set_time.argtypes = [vp, ctypes.c_double]
    L.hydro_conserved_totals.argtypes = [vp, dp]
    ll = ctypes.c_longlong
    L.hydro_conserved_totals_device.argtypes = [vp, ip, ip, ll, ll, vp, vp]
    L.hydro_var_stats.argtypes = [vp, ip, dp]
    L.hydro_time_updates.argtypes = [vp, ip, ip, ctypes.POINTER(HydroTiming)]
    L.hydro_set_launch.argtypes = [vp, ip, ip]
    ipp = ctypes.POINTER(ip)
    L.hydro_batch_create.argtypes = [cfgp, ip, ctypes.POINTER(vp)]
    L.hydro_batch_destroy.argtypes = [vp]
    L.hydro_batch_size.argtypes = [vp]
    L.hydro_batch_device_bytes.argtypes = [vp]
    L.hydro_batch_device_bytes.restype = sz
    L.hydro_batch_upload_U.argtypes = [vp, ip, vp, sz]
    L.hydro_batch_download_U.argtypes = [vp, ip, vp, sz]
    L.hydro_batch_constrain_U.argtypes = [vp]
    L.hydro_batch_update.argtypes = [vp, ip, ip]
    L.hydro_batch_get_time.argtypes = [vp, ip, dp, dp]
    L.hydro_batch_set_active.argtypes = [vp, ip, ip]
    L.hydro_batch_sync.argtypes = [vp]
    L.hydro_batch_time_updates.argtypes = [vp, ip, ip, dp]
    L.hydro_batch_set_launch.argtypes = [vp, ip, ip]
    L.hydro_batch_get_launch.argtypes = [vp, ipp, ipp]
    L.hydro_batch_autotune.argtypes = [vp, ip, ipp, ipp, dp]
    llp = ctypes.POINTER(ctypes.c_longlong)
    L.hydro_server_create.argtypes = [sz, ip, ctypes.POINTER(vp)]
    L.hydro_server_destroy.argtypes = [vp]
    L.hydro_server_submit.argtypes = [vp, cfgp, vp, sz, ip, llp]
    L.hydro_server_tick.argtypes = [vp, ipp]
    L.hydro_server_poll.argtypes = [vp, llp, ip, ipp]
    L.hydro_server_result.argtypes = [vp, ctypes.c_longlong, vp, sz, dp, ipp]
    L.hydro_server_info_get.argtypes = [vp, ctypes.POINTER(HydroServerInfo)]
    L.hydro_server_set_launch.argtypes = [vp, ip, ip]
    L.hydro_get_launch.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise HydroError(lib().hydro_last_error().decode())
    return rc
