"""ctypes mirror of `hydro_cfg` (include/hydro_hip.h) and the name tables that
map the reference's plugin names to the C enums.

The names are the reference's: limiters hydro/app.lua:614-636, integrators
hydro/int/all.lua:162-175, boundaries hydro/solver/gridsolver.lua:567-780,
equations hydro/eqn/euler.lua:6 and glm-maxwell.lua:23, flux hydro/flux/roe.lua:5.
"""
import ctypes

MAX_STATES = 12

LIMITERS = [
    'donor cell', 'Lax-Wendroff', 'Beam-Warming', 'Fromm', 'CHARM', 'HCUS',
    'HQUICK', 'Koren', 'minmod', 'Oshker', 'ospre', 'smart', 'Sweby', 'UMIST',
    'van Albada 1', 'van Albada 2', 'van Leer', 'monotized central', 'superbee',
    'Barth-Jespersen',
]

INTEGRATORS = [
    'forward Euler', 'Runge-Kutta 2', 'Runge-Kutta 2 Heun', 'Runge-Kutta 2 Ralston',
    'Runge-Kutta 3', 'Runge-Kutta 4', 'Runge-Kutta 4, 3/8ths rule',
    'Runge-Kutta 2, TVD', 'Runge-Kutta 2, non-TVD', 'Runge-Kutta 3, TVD',
    'Runge-Kutta 4, TVD', 'Runge-Kutta 4, non-TVD',
]

# stages per step of each integrator (the rows of its Shu-Osher table)
INTEGRATOR_STAGES = [1, 2, 2, 2, 3, 4, 4, 2, 2, 3, 4, 4]

BOUNDARIES = ['periodic', 'mirror', 'freeflow', 'fixed']
EQNS = ['euler', 'glm_maxwell', 'shallow_water']
FLUXES = ['roe', 'hll', 'rusanov', 'euler-hllc', 'marquina']
PLMS = [None, 'plm cons', 'plm prim', 'plm eig', 'plm cons with flux', 'plm athena', 'ppm',
        'plm eig prim', 'plm eig prim ref']
# plm.cl:10-24: the edge states are the cell's own, the flux of no
# reconstruction (which the flux limiter check of gridsolver.lua:119 keeps at
# 'donor cell'), so it runs as plm None
PLM_PIECEWISE_CONSTANT = 'piecewise constant'
PRECISIONS = ['double', 'float']
FACES = ['xmin', 'xmax', 'ymin', 'ymax', 'zmin', 'zmax']

# derived display variables of the Euler equations (euler.lua:240-260), the
# HYDRO_VAR_* ids of hydro_var_stats
EULER_DISPLAY_VARS = {name: 64 + i for i, name in enumerate([
    'P', 'eInt', 'eKin', 'eTotal', 'EInt', 'EKin', 'S', 'H', 'h', 'HTotal', 'hTotal',
    'speed of sound', 'Mach number', 'v x', 'v y', 'v z'])}

EQN_NUM_STATES = {'euler': 6, 'glm_maxwell': 12, 'shallow_water': 5}
EQN_NUM_INT_STATES = {'euler': 5, 'glm_maxwell': 8, 'shallow_water': 4}


class HydroCfg(ctypes.Structure):
    _fields_ = [
        ('dim', ctypes.c_int),
        ('size', ctypes.c_int * 3),
        ('mins', ctypes.c_double * 3),
        ('maxs', ctypes.c_double * 3),
        ('eqn', ctypes.c_int),
        ('flux', ctypes.c_int),
        ('flux_limiter', ctypes.c_int),
        ('plm', ctypes.c_int),
        ('slope_limiter', ctypes.c_int),
        ('integrator', ctypes.c_int),
        ('cfl', ctypes.c_double),
        ('use_fixed_dt', ctypes.c_int),
        ('fixed_dt', ctypes.c_double),
        ('boundary', ctypes.c_int * 6),
        ('boundary_fixed', (ctypes.c_double * MAX_STATES) * 6),
        ('precision', ctypes.c_int),
        ('gamma', ctypes.c_double),
        ('rho_min', ctypes.c_double),
        ('P_min', ctypes.c_double),
        ('div_phi_wavespeed', ctypes.c_double),
        ('div_psi_wavespeed', ctypes.c_double),
        ('fast_math', ctypes.c_int),
        ('roe_entropy_fix', ctypes.c_int),
        ('roe_entropy_fix_factor', ctypes.c_double),
        ('gravity', ctypes.c_double),
    ]


def _index(table, name, what):
    if isinstance(name, int):
        if not 0 <= name < len(table):
            raise ValueError(f'{what} index {name} out of range')
        return name
    try:
        return table.index(name)
    except ValueError:
        raise ValueError(f'unknown {what} {name!r}; expected one of {table}') from None


def make_cfg(dim, size, mins=None, maxs=None, eqn='euler', flux='roe',
             flux_limiter='donor cell', plm=None, slope_limiter='minmod',
             integrator='forward Euler', cfl=0.5, fixed_dt=None,
             boundary='freeflow', boundary_fixed=None, precision='double',
             gamma=7 / 5, rho_min=1e-7, P_min=1e-7,
             div_phi_wavespeed=1.0, div_psi_wavespeed=1.0, fast_math=False,
             use_entropy_flux_fix=False, entropy_flux_fix_limiting_factor=0.2, gravity=1.0):
    """Builds a HydroCfg from reference-style arguments (names or indices).

    `boundary` is a name for all faces or a dict face -> name, like the
    reference's `args.boundary` (gridsolver.lua:122-135). `boundary_fixed` maps a
    face to the cons_t written into its ghost cells.
    """
    c = HydroCfg()
    c.dim = int(dim)
    if isinstance(size, int):
        size = [size]
    size = list(size) + [1] * (3 - len(size))
    for i in range(3):
        c.size[i] = int(size[i]) if i < dim else 1
        c.mins[i] = float(mins[i]) if mins is not None and i < len(mins) else -1.0
        c.maxs[i] = float(maxs[i]) if maxs is not None and i < len(maxs) else 1.0
    c.eqn = _index(EQNS, eqn, 'eqn')
    c.flux = _index(FLUXES, flux, 'flux')
    c.flux_limiter = _index(LIMITERS, flux_limiter, 'flux limiter')
    if plm == PLM_PIECEWISE_CONSTANT:
        if _index(LIMITERS, flux_limiter, 'flux limiter') != 0:
            raise ValueError('are you sure you want to use flux and slope limiters at the same time?')
        plm = None
    c.plm = _index(PLMS, plm, 'plm') if not isinstance(plm, bool) else int(plm)
    c.slope_limiter = _index(LIMITERS, slope_limiter, 'slope limiter')
    c.integrator = _index(INTEGRATORS, integrator, 'integrator')
    c.cfl = float(cfl)
    c.use_fixed_dt = 1 if fixed_dt is not None else 0
    c.fixed_dt = float(fixed_dt) if fixed_dt is not None else 0.0
    for f, face in enumerate(FACES):
        name = boundary if isinstance(boundary, (str, int)) else boundary.get(face, 'freeflow')
        c.boundary[f] = _index(BOUNDARIES, name, 'boundary')
        if boundary_fixed and face in boundary_fixed:
            vals = list(boundary_fixed[face])
            for j, v in enumerate(vals):
                c.boundary_fixed[f][j] = float(v)
    c.precision = _index(PRECISIONS, precision, 'precision')
    c.gamma = float(gamma)
    c.rho_min = float(rho_min)
    c.P_min = float(P_min)
    c.div_phi_wavespeed = float(div_phi_wavespeed)
    c.div_psi_wavespeed = float(div_psi_wavespeed)
    c.fast_math = 1 if fast_math else 0
    # Roe: args.useEntropyFluxFix / entropyFluxFixLimitingFactor (roe.lua:10-16)
    c.roe_entropy_fix = 1 if use_entropy_flux_fix else 0
    c.roe_entropy_fix_factor = float(entropy_flux_fix_limiting_factor)
    # shallow water: gravity (shallow-water.lua:167)
    c.gravity = float(gravity)
    return c


def cfg_num_states(cfg):
    return {0: 6, 1: 12, 2: 5}[cfg.eqn]


def cfg_num_int_states(cfg):
    return {0: 5, 1: 8, 2: 4}[cfg.eqn]


def cfg_ghost_shape(cfg):
    """numpy shape (z, y, x, S) of the reference's ghosted AoS buffer"""
    g = [cfg.size[i] + 4 if i < cfg.dim else 1 for i in range(3)]
    return (g[2], g[1], g[0], cfg_num_states(cfg))
