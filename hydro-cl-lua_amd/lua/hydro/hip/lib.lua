--[[
hydro/hip/lib.lua — LuaJIT binding of the HIP library (include/hydro_hip.h),
the way the reference binds C (hydro/code/safecdef.lua): the header is plain
C, so its preprocessor lines are dropped, HYDRO_MAX_STATES is substituted,
the extern "C" guard's braces are removed, and the rest goes to ffi.cdef.

Copy lua/hydro/* into the hydro-cl-lua tree; HYDRO_HIP_ROOT (environment)
points at this repository (default: the current directory).

Also maps the reference's plugin names (the orders of app.lua:614-636,
int/all.lua and the eqn/flux/plm names) to the header's enums, and turns a
solver's `args` into a hydro_cfg.
--]]
local ffi = require 'ffi'

local root = os.getenv'HYDRO_HIP_ROOT' or '.'

local function readFile(path)
	local f = assert(io.open(path, 'rb'))
	local s = f:read'*a'
	f:close()
	return s
end

-- the header as ffi.cdef takes it (tests/test_lua_binding_cpu.py compiles the
-- same rewrite as C99)
local function cdefText(header)
	return (header
		:gsub('#[^\n]*\n', '\n')            -- #ifndef / #define / #include / #ifdef / #endif
		:gsub('HYDRO_MAX_STATES', '12')     -- its #define went with the line above
		:gsub('extern "C" {', '')           -- the C++ guard ...
		:gsub('\n}%s*\n', '\n'))            -- ... and its closing brace
end

ffi.cdef(cdefText(readFile(root..'/include/hydro_hip.h')))
local C = ffi.load(root..'/hydro-cl-lua_amd/lib/libhydro_hip.so')

local M = {C = C, cdefText = cdefText}

-- errors become Lua errors, as the reference's cl.assert does (choppedup.lua:24)
function M.check(rc)
	if rc ~= 0 then error(ffi.string(C.hydro_last_error()), 2) end
end

local function indexOf(list)
	local t = {}
	for i, name in ipairs(list) do t[name] = i - 1 end
	return t
end

-- app.lua:614-636 (the reference's 1-based index minus one)
M.limiters = indexOf{
	'donor cell', 'Lax-Wendroff', 'Beam-Warming', 'Fromm', 'CHARM', 'HCUS', 'HQUICK', 'Koren', 'minmod',
	'Oshker', 'ospre', 'smart', 'Sweby', 'UMIST', 'van Albada 1', 'van Albada 2', 'van Leer',
	'monotized central', 'superbee', 'Barth-Jespersen',
}
-- int/all.lua, int/fe.lua
M.integrators = indexOf{
	'forward Euler', 'Runge-Kutta 2', 'Runge-Kutta 2 Heun', 'Runge-Kutta 2 Ralston', 'Runge-Kutta 3',
	'Runge-Kutta 4', 'Runge-Kutta 4, 3/8ths rule', 'Runge-Kutta 2, TVD', 'Runge-Kutta 2, non-TVD',
	'Runge-Kutta 3, TVD', 'Runge-Kutta 4, TVD', 'Runge-Kutta 4, non-TVD',
}
M.equations = {euler = 0, ['glm-maxwell'] = 1, ['shallow-water'] = 2}
M.fluxes = {roe = 0, hll = 1, rusanov = 2, ['euler-hllc'] = 3, marquina = 4}
-- solver.usePLM; 'piecewise constant' (plm.cl:10-24) is the flux of no reconstruction
M.plms = {
	['piecewise constant'] = 0, ['plm cons'] = 1, ['plm prim'] = 2, ['plm eig'] = 3,
	['plm cons with flux'] = 4, ['plm athena'] = 5, ppm = 6, ['plm eig prim'] = 7, ['plm eig prim ref'] = 8,
}
M.boundaries = {periodic = 0, mirror = 1, freeflow = 2, fixed = 3}
M.faces = {'xmin', 'xmax', 'ymin', 'ymax', 'zmin', 'zmax'}

local function pick(map, name, what)
	local i = map[name]
	if i == nil then error('unknown '..what..' '..tostring(name)) end
	return i
end

--[[
The hydro_cfg of a solver's args (solverbase.lua:327-338, 775-812;
gridsolver.lua:54-136; fvsolver.lua:28-31): dim, gridSize, mins, maxs, eqn,
eqnArgs, flux, fluxArgs, fluxLimiter, usePLM, slopeLimiter, integrator, cfl,
fixedDT, boundary (a name or a table face -> name, or {name = 'fixed',
args = {...}} with the cons_t of the ghosts); real is 'double' or 'float'.
--]]
function M.cfg(args, real)
	local cfg = ffi.new'hydro_cfg'
	C.hydro_cfg_default(cfg)
	local dim = assert(args.dim, 'expected dim')
	cfg.dim = dim
	local gs = args.gridSize
	for i = 0, dim - 1 do
		cfg.size[i] = type(gs) == 'table' and (gs[i + 1] or gs[({'x', 'y', 'z'})[i + 1]]) or gs
		if args.mins then cfg.mins[i] = args.mins[i + 1] end
		if args.maxs then cfg.maxs[i] = args.maxs[i + 1] end
	end
	cfg.eqn = pick(M.equations, args.eqn or 'euler', 'eqn')
	cfg.flux = pick(M.fluxes, args.flux or 'roe', 'flux')
	local fluxArgs = args.fluxArgs or {}       -- Roe:init (roe.lua:10-16)
	if fluxArgs.useEntropyFluxFix then
		cfg.roe_entropy_fix = 1
		cfg.roe_entropy_fix_factor = fluxArgs.entropyFluxFixLimitingFactor or .2
	end
	cfg.flux_limiter = pick(M.limiters, args.fluxLimiter or 'donor cell', 'flux limiter')
	cfg.plm = args.usePLM and pick(M.plms, args.usePLM == true and 'plm cons' or args.usePLM, 'usePLM') or 0
	cfg.slope_limiter = pick(M.limiters, args.slopeLimiter or 'minmod', 'slope limiter')
	cfg.integrator = pick(M.integrators, args.integrator or 'forward Euler', 'integrator')
	cfg.cfl = args.cfl or .5
	if args.fixedDT then
		cfg.use_fixed_dt = 1
		cfg.fixed_dt = args.fixedDT
	end
	for i, face in ipairs(M.faces) do
		local bc = type(args.boundary) == 'table' and args.boundary[face] or args.boundary or 'freeflow'
		local name = type(bc) == 'table' and bc.name or bc
		cfg.boundary[i - 1] = pick(M.boundaries, name, 'boundary')
		if name == 'fixed' then
			local vals = assert(type(bc) == 'table' and bc.args, 'a fixed boundary needs the cons_t of its ghosts')
			for j, v in ipairs(vals) do cfg.boundary_fixed[i - 1][j - 1] = v end
		end
	end
	cfg.precision = (real or 'double') == 'float' and 1 or 0
	local e = args.eqnArgs or {}
	if e.heatCapacityRatio then cfg.gamma = e.heatCapacityRatio end   -- euler.lua:194
	if e.rhoMin then cfg.rho_min = e.rhoMin end
	if e.PMin then cfg.P_min = e.PMin end
	if e.divPhiWavespeed then cfg.div_phi_wavespeed = e.divPhiWavespeed end   -- glm-maxwell.lua:110-111
	if e.divPsiWavespeed then cfg.div_psi_wavespeed = e.divPsiWavespeed end
	if e.gravity then cfg.gravity = e.gravity end    -- shallow-water.lua:167
	cfg.fast_math = args.fastMath and 1 or 0
	return cfg
end

return M
