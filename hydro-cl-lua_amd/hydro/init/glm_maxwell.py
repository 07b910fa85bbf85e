"""Initial states for GLM-Maxwell (hydro/eqn/glm-maxwell.cl:52-93): the
reference's applyInitCondCell sets D, B, phi = psi = 0, sigma (conductivity,
default 1), rhoCharge (default 0), sqrt_1_eps = sqrt(1/permittivity),
sqrt_1_mu = sqrt(1/permeability). Here: seeded Gaussian pulses in D and B with
eps = mu = 1 and a given conductivity (0 for the serving mix)."""
import numpy as np

from .euler import mesh, interior_shape


def gaussian_pulses(cfg, seed=0, npulses=2, conductivity=0., permittivity=1., permeability=1.):
    x, y, z = mesh(cfg)
    shape = interior_shape(cfg)
    rng = np.random.default_rng(seed)
    coords = [x, y, z]
    U = np.zeros(shape + (12,))
    for _ in range(npulses):
        c = [rng.uniform(cfg.mins[i] + .3 * (cfg.maxs[i] - cfg.mins[i]),
                         cfg.maxs[i] - .3 * (cfg.maxs[i] - cfg.mins[i])) for i in range(3)]
        w = rng.uniform(.05, .15) * (cfg.maxs[0] - cfg.mins[0])
        r2 = np.zeros(shape)
        for i in range(cfg.dim):
            r2 = r2 + (coords[i] - c[i]) ** 2
        g = np.exp(-r2 / (w * w))
        aD = rng.uniform(-1, 1, 3)
        aB = rng.uniform(-1, 1, 3)
        for k in range(3):
            U[..., k] += aD[k] * g
            U[..., 3 + k] += aB[k] * g
    U[..., 8] = 0.                               # rhoCharge
    U[..., 9] = conductivity                     # sigma
    U[..., 10] = np.sqrt(1. / permittivity)      # sqrt_1_eps
    U[..., 11] = np.sqrt(1. / permeability)      # sqrt_1_mu
    return U
