"""Synthetic workloads of the benchmark configurations (seeded), shared by
bench.py, the tests and the smoke check.

kh_2d           2D Euler Kelvin-Helmholtz, Roe + plm cons/minmod, RK2-TVD
sod_1d          1D Euler Sod, Roe + superbee flux limiter, forward Euler
ensemble_2d     512 independent 2D Euler four-quadrant problems, sides
                log-uniform in [64, 512]
serving_mix     Poisson stream of 1D Euler, 2D Euler and 2D GLM-Maxwell solves
"""
import numpy as np

from ._cfg import make_cfg
from .init import euler as init_euler
from .init import glm_maxwell as init_glm


def kh_2d(n=1024, real='double', fast=True, seed=0):
    kw = dict(dim=2, size=(n, n), mins=(-.5, -.5), maxs=(.5, .5), plm='plm cons', slope_limiter='minmod',
              integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', precision=real, fast_math=fast)
    cfg = make_cfg(**kw)
    return kw, init_euler.to_ghosted(cfg, init_euler.kelvin_helmholtz(cfg, seed=seed))


def sod_1d(n=256, fast=False):
    left = [1., 0, 0, 0, 1. / .4, 0]
    right = [.125, 0, 0, 0, .1 / .4, 0]
    kw = dict(dim=1, size=n, flux_limiter='superbee', integrator='forward Euler', cfl=.5,
              boundary={'xmin': 'fixed', 'xmax': 'fixed'}, boundary_fixed={'xmin': left, 'xmax': right},
              fast_math=fast)
    cfg = make_cfg(**kw)
    return kw, init_euler.to_ghosted(cfg, init_euler.sod(cfg))


def quadrant_2d_kw(side_x, side_y, real='float', fast=True, cfl=.4):
    return dict(dim=2, size=(side_x, side_y), mins=(-1, -1), maxs=(1, 1), plm='plm cons',
                slope_limiter='minmod', integrator='Runge-Kutta 2, TVD', cfl=cfl, boundary='freeflow',
                precision=real, fast_math=fast)


def ensemble_sides(count=512, seed=0, lo=64, hi=512):
    rng = np.random.default_rng(seed)
    return [int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))))) for _ in range(count)]


def ensemble_2d(count=512, seed=0, real='float', fast=True, lo=64, hi=512):
    """list of (make_cfg kwargs, ghosted initial state): square grids of side
    round(exp(U(ln lo, ln hi))), four quadrants with rho, P ~ U[.1, 2] and
    |v| <= .5, freeflow, Roe + plm cons/minmod, RK2-TVD, cfl .4"""
    sides = ensemble_sides(count, seed, lo, hi)
    rng = np.random.default_rng(seed + 1)
    out = []
    for s in sides:
        kw = quadrant_2d_kw(s, s, real, fast)
        cfg = make_cfg(**kw)
        states = init_euler.random_quadrant_states(rng)
        out.append((kw, init_euler.to_ghosted(cfg, init_euler.quadrant(cfg, states))))
    return out


def euler_1d_kw(n, real='float', fast=True):
    return dict(dim=1, size=n, plm='plm cons', slope_limiter='minmod', integrator='Runge-Kutta 2, TVD',
                cfl=.4, boundary='freeflow', precision=real, fast_math=fast)


def glm_2d_kw(side, real='float', fast=True):
    return dict(dim=2, size=(side, side), eqn='glm_maxwell', plm='plm cons', slope_limiter='minmod',
                integrator='Runge-Kutta 2, TVD', cfl=.4, boundary='periodic', precision=real, fast_math=fast)


class Request:
    """one solve request: kind, make_cfg kwargs, step count, seed of its state"""

    def __init__(self, kind, kw, steps, seed, arrival=0.):
        self.kind, self.kw, self.steps, self.seed, self.arrival = kind, kw, steps, seed, arrival

    def initial_state(self):
        cfg = make_cfg(**self.kw)
        rng = np.random.default_rng(self.seed)
        if self.kind == 'euler1d':
            left = {'rho': rng.uniform(.1, 2), 'P': rng.uniform(.1, 2), 'v': (rng.uniform(-.5, .5), 0., 0.)}
            right = {'rho': rng.uniform(.1, 2), 'P': rng.uniform(.1, 2), 'v': (rng.uniform(-.5, .5), 0., 0.)}
            U = init_euler.riemann(cfg, left, right, cfg.gamma)
        elif self.kind == 'euler2d':
            U = init_euler.quadrant(cfg, init_euler.random_quadrant_states(rng))
        else:
            U = init_glm.gaussian_pulses(cfg, seed=self.seed, conductivity=0.)
        dtype = np.float64 if cfg.precision == 0 else np.float32
        return init_euler.to_ghosted(cfg, U).astype(dtype)

    def cell_updates(self):
        n = 1
        for i in range(self.kw['dim']):
            s = self.kw['size']
            n *= s if isinstance(s, int) else s[i]
        return n * self.steps


def serving_mix(count, seed=0, rate=None, real='float', fast=True):
    """requests of the serving mix, a third of each kind: 1D Euler with N in
    [256, 65536], 2D Euler with side log-uniform in [64, 1024], 2D GLM-Maxwell
    with side log-uniform in [128, 512]; steps uniform in [100, 2000]; arrival
    times from a Poisson process of the given rate (solves/s) when rate is set,
    drawn from their own generator so that the requests are the same at any
    rate"""
    rng = np.random.default_rng(seed)
    arrivals = np.random.default_rng(seed + 1)     # the mix does not depend on the rate
    out = []
    t = 0.
    for i in range(count):
        kind = ('euler1d', 'euler2d', 'glm2d')[rng.integers(0, 3)]
        steps = int(rng.integers(100, 2001))
        if kind == 'euler1d':
            kw = euler_1d_kw(int(rng.integers(256, 65537)), real, fast)
        elif kind == 'euler2d':
            s = int(round(np.exp(rng.uniform(np.log(64), np.log(1024)))))
            kw = quadrant_2d_kw(s, s, real, fast)
        else:
            s = int(round(np.exp(rng.uniform(np.log(128), np.log(512)))))
            kw = glm_2d_kw(s, real, fast)
        if rate:
            t += arrivals.exponential(1. / rate)
        out.append(Request(kind, kw, steps, int(rng.integers(0, 2 ** 31)), t))
    return out
