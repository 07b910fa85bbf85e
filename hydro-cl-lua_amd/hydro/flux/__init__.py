"""Flux plugins (hydro/flux/*.lua): Roe, HLL, Rusanov, HLLC and Marquina."""
from .roe import Roe
from .hll import HLL
from .rusanov import Rusanov
from .hllc import EulerHLLC
from .marquina import Marquina

FLUXES = {'roe': Roe, 'hll': HLL, 'rusanov': Rusanov, 'euler-hllc': EulerHLLC, 'marquina': Marquina}


def get(name_or_obj, **fluxArgs):
    """a flux plugin by name, built with the reference's args.fluxArgs"""
    if isinstance(name_or_obj, str):
        if name_or_obj not in FLUXES:
            raise ValueError(f'unknown flux {name_or_obj!r} (known: {sorted(FLUXES)})')
        return FLUXES[name_or_obj](**fluxArgs)
    return name_or_obj
