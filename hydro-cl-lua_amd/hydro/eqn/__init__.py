"""Equation plugins (the reference's hydro/eqn/*.lua surface for this path):
name, numStates / numIntStates / numWaves, consVars / primVars / eigenVars,
default parameters (guiVars) and the host-side cons <-> prim conversions.
The stage arithmetic itself lives in the HIP kernels (csrc/fv_euler.h,
csrc/fv_glm.h, csrc/fv_sw.h); these classes describe it and select it by
name."""
from .euler import Euler
from .glm_maxwell import GLMMaxwell
from .shallow_water import ShallowWater

EQUATIONS = {'euler': Euler, 'glm_maxwell': GLMMaxwell, 'glm-maxwell': GLMMaxwell,
             'shallow_water': ShallowWater, 'shallow-water': ShallowWater}


def get(name_or_obj):
    """an equation object from its reference name, or the object itself"""
    if isinstance(name_or_obj, str):
        if name_or_obj not in EQUATIONS:
            raise ValueError(f'unknown eqn {name_or_obj!r} (known: {sorted(EQUATIONS)})')
        return EQUATIONS[name_or_obj]()
    return name_or_obj
