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


def oracle_sensitivity(o_cfg, U0, dim, steps, rel, nq, trials=3, seed=1, with_t=False):
    """How far the FP64 oracle's own result after `steps` steps moves when
    its input U0 (ghosted, float64) is perturbed once by uniform relative
    noise of size `rel` (one rounding of the stored precision): per
    conserved variable, the largest relative L1 (absolute where the
    variable's reference total is 0, as conftest.rel_l1 takes it) over
    `trials` perturbations. Where the
    reference's arithmetic has jumps (calc_Cs is 0 at P <= PMin and
    sqrt(gamma P / rho) above it; 'ppm' takes phi(0) where a difference is
    exactly 0), a one-rounding change of the input moves the result by more
    than a rounding, and a result computed in another arithmetic can differ
    from the oracle by as much. with_t: also the largest relative change of
    the simulated time (dt follows the wave speeds, calc_Cs among them)."""
    import pyoracle
    from conftest import rel_l1

    def run(U):
        o = pyoracle.Oracle(o_cfg)
        o.set_U(U)
        o.boundary()
        o.constrain_U()
        o.update(steps)
        return interior(dim, o.U)[..., :nq].copy(), o.t
    ref, t_ref = run(U0)
    rng = np.random.default_rng(seed)
    worst = np.zeros(nq)
    worst_t = 0.
    for _ in range(trials):
        out, t = run(U0 * (1 + rel * rng.uniform(-1, 1, U0.shape)))
        assert np.all(np.isfinite(out))
        worst = np.maximum(worst, rel_l1(out, ref))
        worst_t = max(worst_t, abs(t - t_ref) / abs(t_ref))
    return (worst, worst_t) if with_t else worst


def init_euler_ghosted(init, cfg, dtype):
    """init(cfg) as the ghosted cons_t buffer, rounded to dtype, in float64"""
    from hydro.init import euler as init_euler
    return init_euler.to_ghosted(cfg, init(cfg)).astype(dtype).astype(np.float64)
