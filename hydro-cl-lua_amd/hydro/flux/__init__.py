"""Flux plugins (hydro/flux/*.lua): the Roe flux of this path."""
from .roe import Roe

FLUXES = {'roe': Roe}


def get(name_or_obj):
    if isinstance(name_or_obj, str):
        if name_or_obj not in FLUXES:
            raise ValueError(f'unknown flux {name_or_obj!r} (known: {sorted(FLUXES)})')
        return FLUXES[name_or_obj]()
    return name_or_obj
