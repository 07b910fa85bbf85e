"""Shallow water equations (hydro/eqn/shallow-water.lua) in the reference's
depthInCell = false form: cons_t = {h, m.xyz, depth} (shallow-water.lua:71-86),
4 integrated states (the depth is read, never integrated) and 4 waves
(shallow-water.lua:55-56), eigen_t = {h, v, C, depth} (293-306), gravity 1
(167), extraTermInFlux = false (22). The momentum source is g h times the
central difference of the depth (shallow-water.cl:366-428)."""
import numpy as np


class ShallowWater:
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
