"""Integrator plugins (hydro/int/all.lua): the explicit Shu-Osher methods by
their reference names. Each class carries its table, read from the library
(the tables the kernels run), as the reference's `alphas` / `betas`."""
import ctypes

from .. import _abi
from .._cfg import INTEGRATORS


class Integrator:
    name = None

    @property
    def index(self):
        return INTEGRATORS.index(self.name)

    def table(self):
        n = ctypes.c_int()
        a = (ctypes.c_double * 16)()
        b = (ctypes.c_double * 16)()
        _abi.check(_abi.lib().hydro_integrator_table(self.index, ctypes.byref(n), a, b))
        k = n.value
        return ([[a[4 * i + j] for j in range(k)] for i in range(k)],
                [[b[4 * i + j] for j in range(k)] for i in range(k)])

    @property
    def alphas(self):
        return self.table()[0]

    @property
    def betas(self):
        return self.table()[1]


def _make(name):
    cls_name = ''.join(c for c in name.title() if c.isalnum())
    return type(cls_name, (Integrator,), {'name': name})


INTEGRATOR_CLASSES = {name: _make(name) for name in INTEGRATORS}


def get(name_or_obj):
    if isinstance(name_or_obj, str):
        if name_or_obj not in INTEGRATOR_CLASSES:
            raise ValueError(f'unknown integrator {name_or_obj!r} (known: {INTEGRATORS})')
        return INTEGRATOR_CLASSES[name_or_obj]()
    return name_or_obj
