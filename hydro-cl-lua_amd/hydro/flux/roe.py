"""Roe flux (hydro/flux/roe.lua, roe.cl:17-163): flux = R (fluxEig) + the mean
of the two physical fluxes (roeUseFluxFromCons), with the flux limiter on the
characteristic jumps; the Harten entropy fix stays off as in roe.lua:10."""


class Roe:
    name = 'roe'
    usesFluxLimiter = True
    useEntropyFluxFix = False
