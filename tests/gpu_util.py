"""Shared set-up of the GPU parity tests: the HIP solver (through the C ABI)
and the CPU oracle started from the same ghosted cons_t buffer."""
import numpy as np

_ORACLE_KEYS = {'dim': 'dim', 'gridSize': 'size', 'mins': 'mins', 'maxs': 'maxs', 'usePLM': 'plm', 'eqn': 'eqn',
                'flux': 'flux', 'slopeLimiter': 'slope_limiter', 'fluxLimiter': 'flux_limiter',
                'integrator': 'integrator', 'cfl': 'cfl', 'boundary': 'boundary', 'fixedDT': 'fixed_dt',
                'boundaryFixed': 'boundary_fixed'}


def oracle_kw(kw):
    out = {_ORACLE_KEYS[k]: v for k, v in kw.items() if k in _ORACLE_KEYS}
    out['precision'] = 'double'
    return out


def interior(dim, U):
    sl = [slice(None)] * 4
    for i in range(dim):
        sl[2 - i] = slice(2, -2)
    return U[tuple(sl)]


def pair(init, **kw):
    """(HipFVSolver, Oracle) holding init(cfg) after boundary + constrainU;
    an FP32 solver gets the state rounded to float, the oracle the same
    values in double"""
    import pyoracle
    from hydro.solver.hipfvsolver import HipFVSolver
    from hydro.init import euler as init_euler
    s = HipFVSolver(**kw)
    U0 = init_euler.to_ghosted(s.cfg, init(s.cfg)).astype(s.dtype)
    s.resetState(U0)
    o = pyoracle.Oracle(pyoracle.make_cfg(**oracle_kw(kw)))
    o.set_U(U0.astype(np.float64))
    o.boundary()
    o.constrain_U()
    return s, o
