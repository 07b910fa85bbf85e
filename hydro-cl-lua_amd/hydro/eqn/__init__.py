"""Equation plugins (the reference's hydro/eqn/*.lua surface for this path):
name, numStates / numIntStates / numWaves, consVars / primVars / eigenVars,
default parameters (guiVars) and the host-side cons <-> prim conversions.
The stage arithmetic itself lives in the HIP kernels (csrc/fv_euler.h,
csrc/fv_glm.h, csrc/fv_sw.h); these classes describe it and select it by
name."""
import numpy as np


class Euler:
    """Euler equations (hydro/eqn/euler.lua): ideal gas, cons_t = {rho, m.xyz,
    ETotal, ePot} (euler.lua:146-177), 5 integrated states and waves
    (euler.lua:13-14), eigen_t = {rho, v, hTotal, Cs, vSq, vL} (298-307),
    heatCapacityRatio 7/5, rhoMin = PMin = 1e-7 (190-197)."""
    name = 'euler'
    numStates = 6
    numIntStates = 5
    numWaves = 5
    roeUseFluxFromCons = True
    consVars = [('rho', 'real'), ('m', 'real3'), ('ETotal', 'real'), ('ePot', 'real')]
    primVars = [('rho', 'real'), ('v', 'real3'), ('P', 'real'), ('ePot', 'real')]
    eigenVars = [('rho', 'real'), ('v', 'real3'), ('hTotal', 'real'), ('Cs', 'real'), ('vSq', 'real'),
                 ('vL', 'real3')]
    # the initial states provided by hydro.init.euler under the reference's names
    initConds = ['Sod', 'Kelvin-Helmholtz', 'sphere', 'quadrant', 'Riemann']
    # solver arguments the equation accepts (eqnArgs), with the reference's defaults
    guiVars = {'heatCapacityRatio': 7 / 5, 'rhoMin': 1e-7, 'PMin': 1e-7}

    def __init__(self, **eqnArgs):
        self.guiVars = dict(Euler.guiVars, **eqnArgs)

    @property
    def gamma(self):
        return self.guiVars['heatCapacityRatio']

    def consFromPrim(self, rho, v, P, ePot=0.):
        """euler.cl:163-173 on host arrays; returns (..., 6)"""
        rho = np.asarray(rho, dtype=np.float64)
        v = [np.asarray(c, dtype=np.float64) for c in v]
        U = np.zeros(np.broadcast(rho, *v, P).shape + (6,))
        U[..., 0] = rho
        for i in range(3):
            U[..., 1 + i] = v[i] * rho
        U[..., 4] = rho * (.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2])) + np.asarray(P) / (self.gamma - 1.)
        U[..., 5] = ePot
        return U

    def primFromCons(self, U):
        """euler.cl:140-159 on host arrays: (rho, v[3], P)"""
        U = np.asarray(U, dtype=np.float64)
        rho = U[..., 0]
        v = [U[..., 1 + i] / rho for i in range(3)]
        EKin = .5 * (U[..., 1] ** 2 + U[..., 2] ** 2 + U[..., 3] ** 2) / rho
        P = (self.gamma - 1.) * (U[..., 4] - EKin)
        return rho, v, P


class GLMMaxwell:
    """GLM-Maxwell (hydro/eqn/glm-maxwell.lua): cons_t = {D, B, phi, psi,
    rhoCharge, sigma, sqrt_1_eps, sqrt_1_mu} (glm-maxwell.lua:48-57), 8
    integrated states and 8 waves, eigen_t = {sqrt_1_eps, sqrt_1_mu} (58-61),
    divPhiWavespeed = divPsiWavespeed = 1 in units of the speed of light."""
    name = 'glm_maxwell'
    numStates = 12
    numIntStates = 8
    numWaves = 8
    consVars = [('D', 'real3'), ('B', 'real3'), ('phi', 'real'), ('psi', 'real'), ('rhoCharge', 'real'),
                ('sigma', 'real'), ('sqrt_1_eps', 'real'), ('sqrt_1_mu', 'real')]
    eigenVars = [('sqrt_1_eps', 'real'), ('sqrt_1_mu', 'real')]
    initConds = ['Gaussian pulses']
    guiVars = {'divPhiWavespeed': 1., 'divPsiWavespeed': 1.}

    def __init__(self, **eqnArgs):
        self.guiVars = dict(GLMMaxwell.guiVars, **eqnArgs)


class ShallowWater:
    """Shallow water equations (hydro/eqn/shallow-water.lua) in the reference's
    depthInCell = false form: cons_t = {h, m.xyz, depth} (shallow-water.lua:71-86),
    4 integrated states (the depth is read, never integrated) and 4 waves
    (shallow-water.lua:55-56), eigen_t = {h, v, C, depth} (293-306), gravity 1
    (167), extraTermInFlux = false (22). The momentum source is g h times the
    central difference of the depth (shallow-water.cl:366-428)."""
    name = 'shallow_water'
    numStates = 5
    numIntStates = 4
    numWaves = 4
    roeUseFluxFromCons = True
    consVars = [('h', 'real'), ('m', 'real3'), ('depth', 'real')]
    primVars = [('h', 'real'), ('v', 'real3'), ('depth', 'real')]
    eigenVars = [('h', 'real'), ('v', 'real3'), ('C', 'real'), ('depth', 'real')]
    initConds = ['dam break', 'humps']
    guiVars = {'gravity': 1.}

    def __init__(self, **eqnArgs):
        self.guiVars = dict(ShallowWater.guiVars, **eqnArgs)

    @property
    def gravity(self):
        return self.guiVars['gravity']

    def consFromPrim(self, h, v, depth=0.):
        """shallow-water.cl:17-30 on host arrays: m = v h; returns (..., 5)"""
        h = np.asarray(h, dtype=np.float64)
        v = [np.asarray(c, dtype=np.float64) for c in v]
        U = np.zeros(np.broadcast(h, *v, depth).shape + (5,))
        U[..., 0] = h
        for i in range(3):
            U[..., 1 + i] = v[i] * h
        U[..., 4] = depth
        return U

    def primFromCons(self, U):
        """shallow-water.cl:3-15 on host arrays: (h, v[3], depth)"""
        U = np.asarray(U, dtype=np.float64)
        h = U[..., 0]
        return h, [U[..., 1 + i] * (1. / h) for i in range(3)], U[..., 4]


EQUATIONS = {'euler': Euler, 'glm_maxwell': GLMMaxwell, 'glm-maxwell': GLMMaxwell,
             'shallow_water': ShallowWater, 'shallow-water': ShallowWater}


def get(name_or_obj):
    """an equation object from its reference name, or the object itself"""
    if isinstance(name_or_obj, str):
        if name_or_obj not in EQUATIONS:
            raise ValueError(f'unknown eqn {name_or_obj!r} (known: {sorted(EQUATIONS)})')
        return EQUATIONS[name_or_obj]()
    return name_or_obj
