"""HipBatch: many independent grids advanced together (the reference runs
several solvers from one app loop, hydro/app.lua:1573-1737). One persistent
launch per stage covers every active grid; three steps (one turn of the dt
slot rotation) are replayed from a captured hipGraph; the launch shape comes
from the autotuner. Each grid keeps its own dt and time."""
import ctypes

import numpy as np

from .. import _abi
from .._cfg import HydroCfg, make_cfg, cfg_ghost_shape


class HipBatch:
    def __init__(self, cfgs):
        """cfgs: list of HydroCfg or dicts of make_cfg keyword arguments"""
        cfgs = [c if isinstance(c, HydroCfg) else make_cfg(**c) for c in cfgs]
        self.cfgs = cfgs
        arr = (HydroCfg * len(cfgs))(*cfgs)
        h = ctypes.c_void_p()
        _abi.check(_abi.lib().hydro_batch_create(arr, len(cfgs), ctypes.byref(h)))
        self.h = h
        self.dtype = np.float64 if cfgs[0].precision == 0 else np.float32
        self.shapes = [cfg_ghost_shape(c) for c in cfgs]

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            _abi.lib().hydro_batch_destroy(h)
            self.h = None

    def __len__(self):
        return len(self.cfgs)

    @property
    def deviceBytes(self):
        return int(_abi.lib().hydro_batch_device_bytes(self.h))

    def upload(self, i, U):
        U = np.ascontiguousarray(np.asarray(U, dtype=self.dtype).reshape(self.shapes[i]))
        _abi.check(_abi.lib().hydro_batch_upload_U(self.h, int(i), U.ctypes.data, U.nbytes))

    def download(self, i):
        out = np.empty(self.shapes[i], dtype=self.dtype)
        _abi.check(_abi.lib().hydro_batch_download_U(self.h, int(i), out.ctypes.data, out.nbytes))
        return out

    def constrainU(self):
        _abi.check(_abi.lib().hydro_batch_constrain_U(self.h))

    def update(self, n=1, graph=True):
        _abi.check(_abi.lib().hydro_batch_update(self.h, int(n), 1 if graph else 0))

    def sync(self):
        _abi.check(_abi.lib().hydro_batch_sync(self.h))

    def time(self, i):
        t, dt = ctypes.c_double(), ctypes.c_double()
        _abi.check(_abi.lib().hydro_batch_get_time(self.h, int(i), ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def setActive(self, i, active=True):
        _abi.check(_abi.lib().hydro_batch_set_active(self.h, int(i), 1 if active else 0))

    def setLaunch(self, rows=0, wavesPerBlock=0):
        _abi.check(_abi.lib().hydro_batch_set_launch(self.h, int(rows), int(wavesPerBlock)))

    def getLaunch(self):
        r, w = ctypes.c_int(), ctypes.c_int()
        _abi.check(_abi.lib().hydro_batch_get_launch(self.h, ctypes.byref(r), ctypes.byref(w)))
        return r.value, w.value

    def timeUpdates(self, n, graph=True):
        ms = ctypes.c_double()
        _abi.check(_abi.lib().hydro_batch_time_updates(self.h, int(n), 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def autotune(self, trialSteps=3):
        r, w, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _abi.check(_abi.lib().hydro_batch_autotune(self.h, int(trialSteps), ctypes.byref(r), ctypes.byref(w),
                                                   ctypes.byref(ms)))
        return r.value, w.value, ms.value
