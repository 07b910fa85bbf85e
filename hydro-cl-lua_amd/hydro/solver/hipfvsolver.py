"""HipFVSolver: the reference's FiniteVolumeSolver interface over the HIP C ABI.

Mirrors `require 'hydro.solver.fvsolver'(args)` (hydro/solver/fvsolver.lua,
gridsolver.lua, solverbase.lua): the same construction arguments (dim,
gridSize, mins, maxs, eqn, flux, fluxLimiter, usePLM, slopeLimiter,
integrator, cfl, fixedDT, boundary, initCond) and the methods the app calls —
update, step(dt), calcDT, boundary, resetState, constrainU, save — plus
UBufObj:toCPU()/fromCPU() on the ghosted cons_t buffer, and t / dt.
Errors raise like the reference's asserts.
"""
import ctypes

import numpy as np

from .. import _abi
from .. import eqn as eqn_plugins
from .. import flux as flux_plugins
from .. import int as int_plugins
from .._cfg import make_cfg, cfg_ghost_shape, INTEGRATOR_STAGES, EULER_DISPLAY_VARS
from ..init import euler as init_euler


class UBufObj:
    """`solver.UBufObj`: the cons_t buffer with ghost cells (x fastest)."""

    def __init__(self, solver):
        self.solver = solver

    def toCPU(self):
        s = self.solver
        out = np.empty(s.shape, dtype=s.dtype)
        _abi.check(_abi.lib().hydro_download_U(s.h, out.ctypes.data, out.nbytes))
        return out

    def fromCPU(self, U):
        s = self.solver
        U = np.ascontiguousarray(np.asarray(U, dtype=s.dtype).reshape(s.shape))
        _abi.check(_abi.lib().hydro_upload_U(s.h, U.ctypes.data, U.nbytes))


_INIT_CONDS = {
    'Sod': lambda cfg, gamma, args: init_euler.sod(cfg, gamma),
    'Kelvin-Helmholtz': lambda cfg, gamma, args: init_euler.kelvin_helmholtz(cfg, gamma, **args),
    'sphere': lambda cfg, gamma, args: init_euler.sphere_blast(cfg, gamma, **args),
}


class HipFVSolver:
    name = 'hipfvsolver'
    numGhost = 2

    def __init__(self, dim, gridSize, eqn='euler', flux='roe', integrator='forward Euler',
                 fluxLimiter='donor cell', usePLM=None, slopeLimiter='minmod', cfl=.5,
                 fixedDT=None, mins=None, maxs=None, boundary='freeflow', boundaryFixed=None,
                 initCond=None, initCondArgs=None, real='double', eqnArgs=None, fastMath=False,
                 fluxArgs=None):
        # plugins by name or as objects (hydro.eqn / hydro.flux / hydro.int)
        self.eqn = eqn_plugins.get(eqn)
        eqnArgs = dict(getattr(self.eqn, 'guiVars', {}), **(eqnArgs or {}))
        self.eqn.guiVars = eqnArgs
        self.flux = flux_plugins.get(flux, **(fluxArgs or {}))
        self.integrator = int_plugins.get(integrator)
        self.cfg = make_cfg(dim=dim, size=gridSize, mins=mins, maxs=maxs, eqn=self.eqn.name.replace('-', '_'),
                            flux=self.flux.name, flux_limiter=fluxLimiter, plm=usePLM, slope_limiter=slopeLimiter,
                            integrator=self.integrator.name, cfl=cfl, fixed_dt=fixedDT, boundary=boundary,
                            boundary_fixed=boundaryFixed, precision=real,
                            gamma=eqnArgs.get('heatCapacityRatio', 7 / 5),
                            rho_min=eqnArgs.get('rhoMin', 1e-7), P_min=eqnArgs.get('PMin', 1e-7),
                            div_phi_wavespeed=eqnArgs.get('divPhiWavespeed', 1.),
                            div_psi_wavespeed=eqnArgs.get('divPsiWavespeed', 1.),
                            gravity=eqnArgs.get('gravity', 1.), fast_math=fastMath,
                            use_entropy_flux_fix=getattr(self.flux, 'useEntropyFluxFix', False),
                            entropy_flux_fix_limiting_factor=getattr(self.flux, 'entropyFluxFixLimitingFactor', .2))
        self.dim = int(dim)
        self.dtype = np.float64 if real == 'double' else np.float32
        self.shape = cfg_ghost_shape(self.cfg)
        self.initCond = initCond
        self.initCondArgs = dict(initCondArgs or {})
        self.cfl = float(cfl)
        self.stages = INTEGRATOR_STAGES[self.cfg.integrator]
        h = ctypes.c_void_p()
        _abi.check(_abi.lib().hydro_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h
        self.UBufObj = UBufObj(self)
        if initCond is not None:
            self.resetState()

    def __del__(self):
        h = getattr(self, 'h', None)
        if h:
            _abi.lib().hydro_destroy(h)
            self.h = None

    # -- sizes (gridsolver.lua:60-96)
    @property
    def numCells(self):
        return int(_abi.lib().hydro_num_cells(self.h))

    @property
    def gridSize(self):
        return tuple(self.cfg.size[i] + 4 if i < self.dim else 1 for i in range(3))

    @property
    def sizeWithoutBorder(self):
        return tuple(self.cfg.size[i] if i < self.dim else 1 for i in range(3))

    @property
    def deviceBytes(self):
        return int(_abi.lib().hydro_device_bytes(self.h))

    # -- time
    @property
    def t(self):
        t = ctypes.c_double()
        _abi.check(_abi.lib().hydro_get_time(self.h, ctypes.byref(t), None))
        return t.value

    @t.setter
    def t(self, v):
        _abi.check(_abi.lib().hydro_set_time(self.h, float(v)))

    @property
    def dt(self):
        dt = ctypes.c_double()
        _abi.check(_abi.lib().hydro_get_time(self.h, None, ctypes.byref(dt)))
        return dt.value

    # -- the reference's solver methods
    def resetState(self, U=None):
        """SolverBase:resetState (solverbase.lua:2094-2114): initial state, boundary, constrainU"""
        if U is None:
            if self.initCond not in _INIT_CONDS:
                raise ValueError(f'unknown initCond {self.initCond!r}')
            Uint = _INIT_CONDS[self.initCond](self.cfg, self.cfg.gamma, self.initCondArgs)
            U = init_euler.to_ghosted(self.cfg, Uint)
        self.UBufObj.fromCPU(U)
        self.t = 0.
        self.constrainU()

    def boundary(self):
        _abi.check(_abi.lib().hydro_boundary(self.h))

    def constrainU(self):
        _abi.check(_abi.lib().hydro_constrain_U(self.h))

    def calcDT(self):
        dt = ctypes.c_double()
        _abi.check(_abi.lib().hydro_calc_dt(self.h, ctypes.byref(dt)))
        return dt.value

    def step(self, dt):
        _abi.check(_abi.lib().hydro_step(self.h, float(dt)))

    def update(self, n=1):
        """n x SolverBase:update (calcDT, step, t += dt) with dt kept on the device"""
        _abi.check(_abi.lib().hydro_update(self.h, int(n)))

    def sync(self):
        _abi.check(_abi.lib().hydro_sync(self.h))

    def conservedTotals(self):
        q = 5 if self.cfg.eqn == 0 else 8
        out = (ctypes.c_double * q)()
        _abi.check(_abi.lib().hydro_conserved_totals(self.h, out))
        return np.array(out[:])

    def calcDisplayVarRangeAndAvgAndStdDev(self, field):
        """(min, max, avg, stddev) of the stored cons_t field `field` (index),
        as SolverBase:calcDisplayVarRangeAndAvgAndStdDev (solverbase.lua:
        2871-2994) gives them for the display var 'U <field>', or of a derived
        Euler display variable by its name ('P', 'eInt', ..., 'Mach number',
        'v x'; euler.lua:240-260): min/max over the ghosted grid, avg/stddev
        over the interior; computed on the GPU"""
        if isinstance(field, str):
            if field not in EULER_DISPLAY_VARS:
                raise ValueError(f'unknown display var {field!r}; expected one of {list(EULER_DISPLAY_VARS)}')
            field = EULER_DISPLAY_VARS[field]
        out = (ctypes.c_double * 4)()
        _abi.check(_abi.lib().hydro_var_stats(self.h, int(field), out))
        return tuple(out[:])

    def saveBuffer(self, buffer, basefn):
        """GridSolver:saveBuffer (gridsolver.lua:1410-1455): the ghosted
        buffer (z, y, x, channels) as `<basefn>.fits`, one plane per channel"""
        from .. import fits
        return fits.save_buffer(buffer.toCPU() if hasattr(buffer, 'toCPU') else buffer, basefn)

    def save(self, prefix=None):
        """GridSolver:save (gridsolver.lua:1465-1470) for the buffer this
        solver holds, UBuf: writes `<prefix>_UBuf.fits` (`UBuf.fits` without
        a prefix) and returns the file names"""
        return [self.saveBuffer(self.UBufObj, (prefix + '_' if prefix else '') + 'UBuf')]

    def timeUpdates(self, n, perKernel=True):
        t = _abi.HydroTiming()
        _abi.check(_abi.lib().hydro_time_updates(self.h, int(n), 1 if perKernel else 0, ctypes.byref(t)))
        return t

    def setLaunch(self, rowsPerWave=0, xcdRemap=0):
        """launch shape of the stage kernel (xcdRemap: 1 on, 2 off, 0 keep)"""
        _abi.check(_abi.lib().hydro_set_launch(self.h, int(rowsPerWave), int(xcdRemap)))

    def getLaunch(self):
        r, x = ctypes.c_int(), ctypes.c_int()
        _abi.check(_abi.lib().hydro_get_launch(self.h, ctypes.byref(r), ctypes.byref(x)))
        return r.value, x.value

    def interior(self, U):
        sl = [slice(None)] * 4
        for i in range(self.dim):
            sl[2 - i] = slice(2, -2)
        return U[tuple(sl)]
