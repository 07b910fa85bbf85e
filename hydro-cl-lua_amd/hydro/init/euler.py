"""Initial states for the Euler equations, as host arrays.

Each function returns primitive fields (rho, v[3], P) on the interior cell
centres; `cons_from_prim` turns them into the reference's cons_t. Cell centres
follow coord.lua:1421-1436: x_i = (i + .5) / N * (max - min) + min.
The named problems follow hydro/init/euler.lua (RiemannProblem 57-118, Sod
742-751, quadrant problems 249-283, sphere 1273-1297, Kelvin-Helmholtz
1639-1742); the KH band and noise and the blast centre are seeded here
instead of coming from LuaJIT's math.random (init.lua:211-219).
"""
import numpy as np


def cell_centers(cfg):
    """list of 1D arrays of interior cell centres per used axis"""
    out = []
    for i in range(cfg.dim):
        n = cfg.size[i]
        out.append((np.arange(n) + .5) / n * (cfg.maxs[i] - cfg.mins[i]) + cfg.mins[i])
    return out


def mesh(cfg):
    """broadcastable (z, y, x) coordinate arrays of the interior"""
    c = cell_centers(cfg)
    z = c[2][:, None, None] if cfg.dim > 2 else np.zeros((1, 1, 1))
    y = c[1][None, :, None] if cfg.dim > 1 else np.zeros((1, 1, 1))
    x = c[0][None, None, :]
    return x, y, z


def interior_shape(cfg):
    return (cfg.size[2] if cfg.dim > 2 else 1, cfg.size[1] if cfg.dim > 1 else 1, cfg.size[0])


def cons_from_prim(rho, v, P, gamma):
    """consFromPrim (euler.cl:163-173): rho, m = v rho, ETotal = rho (.5 v.v) + P / (gamma - 1)"""
    rho = np.asarray(rho, dtype=np.float64)
    vx, vy, vz = (np.asarray(c, dtype=np.float64) for c in v)
    U = np.zeros(rho.shape + (6,))
    U[..., 0] = rho
    U[..., 1] = vx * rho
    U[..., 2] = vy * rho
    U[..., 3] = vz * rho
    U[..., 4] = rho * (.5 * (vx * vx + vy * vy + vz * vz)) + P / (gamma - 1.)
    return U


def to_ghosted(cfg, Uint):
    """interior cons (z, y, x, S) -> the reference's ghosted AoS buffer (ghosts zero)"""
    S = Uint.shape[-1]
    g = [cfg.size[i] + 4 if i < cfg.dim else 1 for i in range(3)]
    out = np.zeros((g[2], g[1], g[0], S))
    sl = [slice(None)] * 3
    for i in range(cfg.dim):
        sl[2 - i] = slice(2, -2)
    out[tuple(sl)] = Uint
    return out


def riemann(cfg, left, right, gamma):
    """RiemannProblem (init/euler.lua:57-118): left state where every used
    coordinate is below the domain middle"""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    lhs = np.ones(shape, dtype=bool)
    coords = [x, y, z]
    for i in range(cfg.dim):
        mid = .5 * (cfg.mins[i] + cfg.maxs[i])
        lhs = lhs & (coords[i] < mid)
    rho = np.where(lhs, left['rho'], right['rho'])
    P = np.where(lhs, left['P'], right['P'])
    v = [np.where(lhs, left.get('v', (0, 0, 0))[i], right.get('v', (0, 0, 0))[i]) for i in range(3)]
    return cons_from_prim(rho, v, P, gamma)


def sod(cfg, gamma=1.4):
    """Sod (init/euler.lua:742-751) with left rho=1, P=1 and right rho=.125, P=.1"""
    return riemann(cfg, {'rho': 1., 'P': 1.}, {'rho': .125, 'P': .1}, gamma)


def kelvin_helmholtz(cfg, gamma=1.4, seed=0, rho_in=2., rho_out=1., v_in=.5, v_out=-.5, P=2.5,
                     amplitude=.01, noise=1e-3, frequency=2.):
    """Shear layer: |y - mid| < quarter height has rho_in, vx=v_in; elsewhere rho_out,
    vx=v_out; vy = amplitude sin(2 pi frequency (x - xmin)/(xmax - xmin)) plus seeded
    uniform noise in [-noise, noise]."""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    ymid = .5 * (cfg.mins[1] + cfg.maxs[1])
    quarter = .25 * (cfg.maxs[1] - cfg.mins[1])
    inside = np.broadcast_to(np.abs(y - ymid) < quarter, shape)
    rho = np.where(inside, rho_in, rho_out)
    vx = np.where(inside, v_in, v_out)
    rng = np.random.default_rng(seed)
    theta = 2. * np.pi * frequency * (x - cfg.mins[0]) / (cfg.maxs[0] - cfg.mins[0])
    vy = np.broadcast_to(amplitude * np.sin(theta), shape) + rng.uniform(-noise, noise, size=shape)
    vz = np.zeros(shape)
    Pf = np.full(shape, P)
    return cons_from_prim(rho, (vx, vy, vz), Pf, gamma)


def quadrant(cfg, states, gamma=1.4):
    """Four-quadrant Riemann problem (init/euler.lua:249-283): states[0..3] for
    (x>mid, y>mid), (x<mid, y>mid), (x<mid, y<mid), (x>mid, y<mid), each a dict
    with rho, P, v."""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    xp = np.broadcast_to(x > .5 * (cfg.mins[0] + cfg.maxs[0]), shape)
    yp = np.broadcast_to(y > .5 * (cfg.mins[1] + cfg.maxs[1]), shape)
    q = np.where(yp, np.where(xp, 0, 1), np.where(~xp, 2, 3))
    rho = np.choose(q, [s['rho'] for s in states])
    P = np.choose(q, [s['P'] for s in states])
    v = [np.choose(q, [s['v'][i] for s in states]) for i in range(3)]
    return cons_from_prim(rho, v, P, gamma)


def random_quadrant_states(rng):
    """rho, P ~ U[.1, 2]; v = r (cos t, sin t), r ~ U[0, .5], t ~ U[0, 2 pi)"""
    out = []
    for _ in range(4):
        r = rng.uniform(0, .5)
        t = rng.uniform(0, 2 * np.pi)
        out.append({'rho': rng.uniform(.1, 2), 'P': rng.uniform(.1, 2),
                    'v': (r * np.cos(t), r * np.sin(t), 0.)})
    return out


def sphere_blast(cfg, gamma=1.4, seed=0, radius=.1, rho=1., P_out=.1, P_in=10.):
    """Spherical blast (init/euler.lua:1273-1297 'sphere'): P_in inside a sphere
    whose centre is drawn uniformly in the middle half of the box."""
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    rng = np.random.default_rng(seed)
    c = [rng.uniform(cfg.mins[i] + .25 * (cfg.maxs[i] - cfg.mins[i]),
                     cfg.maxs[i] - .25 * (cfg.maxs[i] - cfg.mins[i])) for i in range(3)]
    coords = [x, y, z]
    r2 = np.zeros(shape)
    for i in range(cfg.dim):
        r2 = r2 + (coords[i] - c[i]) ** 2
    inside = r2 < radius * radius
    Pf = np.where(inside, P_in, P_out)
    rhof = np.full(shape, rho)
    zero = np.zeros(shape)
    return cons_from_prim(rhof, (zero, zero, zero), Pf, gamma)
