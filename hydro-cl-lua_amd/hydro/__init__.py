"""MI355X finite-volume hot path of hydro-cl-lua behind the reference's plugin
names (hydro.solver / hydro.eqn / hydro.flux / hydro.int)."""
