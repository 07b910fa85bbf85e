"""Rusanov flux (hydro/flux/rusanov.lua, rusanov.cl:13-39):
F = (FL + FR - a (UR - UL)) / 2. The reference assigns a = max(|lambda_min|,
|lambda_max|) once per side in turn, so the right state's value is the one
used; kept for parity. No flux limiter; the Euler equations."""


class Rusanov:
    name = 'rusanov'
    usesFluxLimiter = False
