"""Euler equations (hydro/eqn/euler.lua): ideal gas, cons_t = {rho, m.xyz,
ETotal, ePot} (euler.lua:146-177), 5 integrated states and waves
(euler.lua:13-14), eigen_t = {rho, v, hTotal, Cs, vSq, vL} (298-307),
heatCapacityRatio 7/5, rhoMin = PMin = 1e-7 (190-197)."""
import numpy as np


class Euler:
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
