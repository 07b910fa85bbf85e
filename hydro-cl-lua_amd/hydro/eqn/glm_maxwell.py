"""GLM-Maxwell (hydro/eqn/glm-maxwell.lua): cons_t = {D, B, phi, psi,
rhoCharge, sigma, sqrt_1_eps, sqrt_1_mu} (glm-maxwell.lua:48-57), 8
integrated states and 8 waves, eigen_t = {sqrt_1_eps, sqrt_1_mu} (58-61),
divPhiWavespeed = divPsiWavespeed = 1 in units of the speed of light."""


class GLMMaxwell:
    name = 'glm_maxwell'
    numStates = 12
    numIntStates = 8
    numWaves = 8
    consVars = [('D', 'real3'), ('B', 'real3'), ('phi', 'real'), ('psi', 'real'), ('rhoCharge', 'real'),
                ('sigma', 'real'), ('sqrt_1_eps', 'real'), ('sqrt_1_mu', 'real')]
    eigenVars = [('sqrt_1_eps', 'real'), ('sqrt_1_mu', 'real')]
    initConds = ['Gaussian pulses']
    guiVars = {'divPhiWavespeed': 1., 'divPsiWavespeed': 1.}

    def __init__(self, **eqnArgs):
        self.guiVars = dict(GLMMaxwell.guiVars, **eqnArgs)
