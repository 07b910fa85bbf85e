"""Marquina flux (hydro/flux/marquina.lua, marquina.cl:5-93; Marti et al.
1997, appendix A): each side's state and physical flux in that side's own
characteristic basis (eigen_forCell, euler.cl:315-341); per wave, upwind when
both sides' speeds share a sign, else a local Lax-Friedrichs split with the
larger |lambda|. The reference's plugin carries the name 'hll'
(marquina.lua:9); it is selected here as 'marquina'. No flux limiter; the
Euler equations."""


class Marquina:
    name = 'marquina'
    usesFluxLimiter = False
