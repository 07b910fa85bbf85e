"""hydro_conserved_totals_device, the entry point that allocates nothing: the
totals of a state the caller holds on the device, partial sums in the
library's static device array. Checked against hydro_conserved_totals (the
same partition and order: the same bits) on two calls in a row, the second
with a smaller input, in FP64 and FP32, and its size check."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(n, real, seed):
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(dim=2, gridSize=n, mins=(-.5, -.5), maxs=(.5, .5), usePLM='plm cons', slopeLimiter='minmod',
                    integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', real=real)
    U0 = init_euler.to_ghosted(s.cfg, init_euler.kelvin_helmholtz(s.cfg, seed=seed))
    s.resetState(U0.astype(s.dtype))
    s.update(3)
    return s


class _Dev:
    """device buffers through the HIP runtime the library itself is linked to
    (libamdhip64.so.7, already loaded with it): the test allocates and copies;
    the entry point under test does neither"""

    def __init__(self):
        self.hip = ctypes.CDLL('libamdhip64.so.7')
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # hipMemcpyHostToDevice
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _planes(s):
    """the solver's interior state as q planes (SoA, x fastest)"""
    U = s.interior(s.UBufObj.toCPU())[..., :5]           # (z, y, x, field)
    return np.ascontiguousarray(np.moveaxis(U, -1, 0).reshape(5, -1))


@pytest.mark.parametrize('real', ['double', 'float'])
def test_totals_device_matches_solver_totals(real):
    from hydro import _abi
    L = _abi.lib()
    prec = 0 if real == 'double' else 1
    dev = _Dev()
    out = dev.alloc(5 * 8)
    try:
        for n, seed in (((96, 80), 3), ((40, 28), 4)):          # the second call's input is smaller
            s = _solver(n, real, seed)
            ref = np.asarray(s.conservedTotals())
            P = _planes(s)
            d = dev.put(P)
            _abi.check(L.hydro_conserved_totals_device(d, prec, 5, P.shape[1], P.shape[1], out, None))
            got = dev.get(out, 5, np.float64)
            print(real, n, got, ref)
            assert np.array_equal(got, ref), (got, ref)
    finally:
        dev.free()


def test_totals_device_rejects_too_many_fields():
    from hydro import _abi
    L = _abi.lib()
    dev = _Dev()
    try:
        P = dev.put(np.zeros((16, 8)))
        out = dev.alloc(16 * 8)
        rc = L.hydro_conserved_totals_device(P, 0, 13, 8, 8, out, None)
        assert rc != 0 and b'num_fields 13' in L.hydro_last_error()
    finally:
        dev.free()
