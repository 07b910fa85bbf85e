"""HLLC flux of the Euler equations (hydro/flux/euler-hllc.lua,
euler-hllc.cl:13-240): HLL's 'Davis direct bounded' wave estimates, the
contact speed sStar, and the star-state fluxes of hllcMethod 2 (the default,
euler-hllc.lua:18; Toro 2012, variation 2) on the side of the contact, with
the HLL flux when sStar is not a number. No flux limiter."""
from .hll import HLL


class EulerHLLC(HLL):
    name = 'euler-hllc'
    hllcMethod = 2
