"""Initial states for the shallow water equations (hydro/eqn/shallow-water.cl:
applyInitCondCell maps the Euler initial conditions' rho to the height h and
v to the velocity, m = v h, and sets the depth; cons_t here is the
depthInCell = false form: h, m.xyz, depth). Problems: a dam break (the
Riemann problem of two still columns), seeded Gaussian humps on a seeded
sloped, bumpy depth, and a still lake."""
import numpy as np

from .euler import mesh, interior_shape


def cons_from_prim(h, v, depth):
    """consFromPrim (shallow-water.cl:17-30): m = v h, depth carried"""
    h = np.asarray(h, dtype=np.float64)
    U = np.zeros(h.shape + (5,))
    U[..., 0] = h
    for i in range(3):
        U[..., 1 + i] = np.asarray(v[i], dtype=np.float64) * h
    U[..., 4] = depth
    return U


def dam_break(cfg, hl=1., hr=.5, axis=0):
    """still water of height hl where the coordinate along `axis` is below
    the middle of the domain, hr above, flat depth 0"""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    c = np.broadcast_to([x, y, z][axis], shape)
    mid = .5 * (cfg.mins[axis] + cfg.maxs[axis])
    zero = np.zeros(shape)
    return cons_from_prim(np.where(c < mid, hl, hr), (zero, zero, zero), zero)


def humps(cfg, seed=0, nhumps=3, slope=.2):
    """a unit column with seeded Gaussian humps (amplitude up to .3) and a
    seeded swirl, on a depth with a seeded linear slope and bumps"""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    rng = np.random.default_rng(seed)
    coords = [x, y, z]
    h = np.ones(shape)
    depth = np.zeros(shape)
    vel = [np.zeros(shape) for _ in range(3)]
    L = cfg.maxs[0] - cfg.mins[0]
    for i in range(cfg.dim):
        depth = depth + slope * rng.uniform(-1, 1) * coords[i]
    for _ in range(nhumps):
        c = [rng.uniform(cfg.mins[i] + .25 * (cfg.maxs[i] - cfg.mins[i]),
                         cfg.maxs[i] - .25 * (cfg.maxs[i] - cfg.mins[i])) for i in range(3)]
        w = rng.uniform(.05, .15) * L
        r2 = np.zeros(shape)
        for i in range(cfg.dim):
            r2 = r2 + (coords[i] - c[i]) ** 2
        g = np.exp(-r2 / (w * w))
        h = h + rng.uniform(.1, .3) * g
        depth = depth + .1 * rng.uniform(-1, 1) * g
        for i in range(cfg.dim):
            vel[i] = vel[i] + .2 * rng.uniform(-1, 1) * g
    return cons_from_prim(h, vel, depth)
