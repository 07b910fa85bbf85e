"""Flux plugins (hydro/flux/*.lua): Roe, HLL, Rusanov, HLLC and Marquina.
Each class names a flux the HIP stage kernels compute (csrc/fv_stage.h)
and carries the reference plugin's options for it."""


class Roe:
    """Roe flux (hydro/flux/roe.lua, roe.cl:17-163): flux = R (fluxEig) + the
    mean of the two physical fluxes (roeUseFluxFromCons), with the flux
    limiter on the characteristic jumps; Harten's entropy fix (roe.cl:83-106)
    when useEntropyFluxFix, off by default as in roe.lua:10."""
    name = 'roe'
    usesFluxLimiter = True

    def __init__(self, useEntropyFluxFix=False, entropyFluxFixLimitingFactor=.2):
        # Roe:init (roe.lua:13-16)
        self.useEntropyFluxFix = bool(useEntropyFluxFix)
        self.entropyFluxFixLimitingFactor = float(entropyFluxFixLimitingFactor)


class HLL:
    """HLL flux (hydro/flux/hll.lua, hll.cl:13-73) with the 'Davis direct
    bounded' wave estimates (hll.lua:10): sL = min(lambda_min(UL),
    lambda_min(Roe)), sR = max(lambda_max(UR), lambda_max(Roe)); F(UL) when
    0 <= sL, F(UR) when sR <= 0, else (sR FL - sL FR + sL sR (UR - UL)) /
    (sR - sL). No flux limiter (fvsolver.lua:63); the Euler equations."""
    name = 'hll'
    usesFluxLimiter = False
    hllCalcWaveMethod = 'Davis direct bounded'


class Rusanov:
    """Rusanov flux (hydro/flux/rusanov.lua, rusanov.cl:13-39):
    F = (FL + FR - a (UR - UL)) / 2. The reference assigns a =
    max(|lambda_min|, |lambda_max|) once per side in turn, so the right
    state's value is the one used; kept for parity. No flux limiter; the
    Euler equations."""
    name = 'rusanov'
    usesFluxLimiter = False


class EulerHLLC(HLL):
    """HLLC flux of the Euler equations (hydro/flux/euler-hllc.lua,
    euler-hllc.cl:13-240): HLL's 'Davis direct bounded' wave estimates, the
    contact speed sStar, and the star-state fluxes of hllcMethod 2 (the
    default, euler-hllc.lua:18; Toro 2012, variation 2) on the side of the
    contact, with the HLL flux when sStar is not a number. No flux limiter."""
    name = 'euler-hllc'
    hllcMethod = 2


class Marquina:
    """Marquina flux (hydro/flux/marquina.lua, marquina.cl:5-93; Marti et
    al. 1997, appendix A): each side's state and physical flux in that side's
    own characteristic basis (eigen_forCell, euler.cl:315-341); per wave,
    upwind when both sides' speeds share a sign, else a local Lax-Friedrichs
    split with the larger |lambda|. The reference's plugin carries the name
    'hll' (marquina.lua:9); it is selected here as 'marquina'. No flux
    limiter; the Euler equations."""
    name = 'marquina'
    usesFluxLimiter = False


FLUXES = {'roe': Roe, 'hll': HLL, 'rusanov': Rusanov, 'euler-hllc': EulerHLLC, 'marquina': Marquina}


def get(name_or_obj, **fluxArgs):
    """a flux plugin by name, built with the reference's args.fluxArgs"""
    if isinstance(name_or_obj, str):
        if name_or_obj not in FLUXES:
            raise ValueError(f'unknown flux {name_or_obj!r} (known: {sorted(FLUXES)})')
        return FLUXES[name_or_obj](**fluxArgs)
    return name_or_obj
