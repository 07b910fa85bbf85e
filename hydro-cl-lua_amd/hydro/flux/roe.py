"""Roe flux (hydro/flux/roe.lua, roe.cl:17-163): flux = R (fluxEig) + the mean
of the two physical fluxes (roeUseFluxFromCons), with the flux limiter on the
characteristic jumps; Harten's entropy fix (roe.cl:83-106) when
useEntropyFluxFix, off by default as in roe.lua:10."""


class Roe:
    name = 'roe'
    usesFluxLimiter = True

    def __init__(self, useEntropyFluxFix=False, entropyFluxFixLimitingFactor=.2):
        # Roe:init (roe.lua:13-16)
        self.useEntropyFluxFix = bool(useEntropyFluxFix)
        self.entropyFluxFixLimitingFactor = float(entropyFluxFixLimitingFactor)
