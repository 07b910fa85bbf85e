"""HLL flux (hydro/flux/hll.lua, hll.cl:13-73) with the 'Davis direct
bounded' wave estimates (hll.lua:10): sL = min(lambda_min(UL), lambda_min(Roe)),
sR = max(lambda_max(UR), lambda_max(Roe)); F(UL) when 0 <= sL, F(UR) when
sR <= 0, else (sR FL - sL FR + sL sR (UR - UL)) / (sR - sL). No flux limiter
(fvsolver.lua:63); the Euler equations."""


class HLL:
    name = 'hll'
    usesFluxLimiter = False
    hllCalcWaveMethod = 'Davis direct bounded'
