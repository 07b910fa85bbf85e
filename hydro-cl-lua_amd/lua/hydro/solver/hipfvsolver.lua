--[[
hydro/solver/hipfvsolver.lua — the finite-volume solver of hydro-cl-lua over
the HIP C ABI (hydro/hip/lib.lua), for LuaJIT, without lua-opencl.

It takes the reference solver's args (fvsolver.lua:28-31, gridsolver.lua:54-136,
solverbase.lua:327-338: dim, gridSize, mins, maxs, eqn, eqnArgs, flux,
fluxArgs, fluxLimiter, usePLM, slopeLimiter, integrator, cfl, fixedDT,
boundary) and has the members the app loop and scripts use: update, step(dt),
calcDT, boundary, resetState, constrainU, save(prefix), UBufObj:toCPU() /
:fromCPU(), t, dt, numCells, gridSize, sizeWithoutBorder, dim, mins, maxs,
numGhost, eqn (name, numStates, numIntStates), calcDisplayVarRangeAndAvgAndStdDev.

It implements that interface rather than inheriting SolverBase:init, because
every step of that constructor builds OpenCL objects: initCLDomainVars reads
the CL device (solverbase.lua:759-773), createSolverBuf / createInitCondBuf /
finalizeCLAllocs make CLBuffers (1076-1095, 1399-1430), the refresh*Program
calls compile programs (1184, 1670; gridsolver.lua:1229) and resetState runs
the initial-condition kernel (2094-2114). Here the state lives in a host FFI
array (cons_t per cell, ghosts included, x fastest: the reference's UBuf
layout) and on the device behind the handle; nothing is compiled at run time.

Initial state: the reference's initial conditions are OpenCL code, so a
headless run passes args.initState = function(x, y, z, U) that fills one
cons_t (U[0 .. numStates-1]) at the cell centre (coord.lua:1421-1436), or
args.initBuffer, a cons_t array of numCells entries.

Save: GridSolver:save (gridsolver.lua:1465-1470) writes every solver buffer
through saveBuffer; here the one buffer is UBuf, written by the same
image / FITS path in the same plane order (saveBuffer, gridsolver.lua:1410-1455).
--]]
local ffi = require 'ffi'
local class = require 'ext.class'
local hip = require 'hydro.hip.lib'

local C, check = hip.C, hip.check

-- equations: cons_t sizes (include/hydro_hip.h)
local eqnInfo = {
	euler = {numStates = 6, numIntStates = 5},
	['glm-maxwell'] = {numStates = 12, numIntStates = 8},
	['shallow-water'] = {numStates = 5, numIntStates = 4},
}

-- the host buffer that stands where the reference's CLBuffer UBufObj is
local HostBuffer = class()
function HostBuffer:init(solver)
	self.solver = solver
	self.count = solver.numCells * solver.eqn.numStates
	self.obj = ffi.new(solver.real..'[?]', self.count)
	self.size = ffi.sizeof(solver.real) * self.count
end
-- the device state with ghost cells filled by the boundary conditions
function HostBuffer:toCPU(dst)
	check(C.hydro_download_U(self.solver.hip, self.obj, self.size))
	if dst then ffi.copy(dst, self.obj, self.size) return dst end
	return self.obj
end
function HostBuffer:fromCPU(src)
	if src and src ~= self.obj then ffi.copy(self.obj, src, self.size) end
	check(C.hydro_upload_U(self.solver.hip, self.obj, self.size))
end

local HipFVSolver = class()
HipFVSolver.name = 'hipfvsolver'
HipFVSolver.numGhost = 2
HipFVSolver.t = 0

function HipFVSolver:init(args)
	self.app = args.app
	self.real = args.real or (args.app and args.app.real) or 'double'
	self.initState = args.initState
	self.initBuffer = args.initBuffer
	self.cfg = hip.cfg(args, self.real)
	self.dim = self.cfg.dim
	local eqnName = args.eqn or 'euler'
	self.eqn = setmetatable({name = eqnName}, {__index = assert(eqnInfo[eqnName], 'unknown eqn '..eqnName)})
	-- gridsolver.lua:60-96: sizes with and without the ghost layers
	self.sizeWithoutBorder, self.gridSize = {}, {}
	self.mins, self.maxs = {}, {}
	self.numCells = 1
	for i, x in ipairs{'x', 'y', 'z'} do
		local n = i <= self.dim and self.cfg.size[i - 1] or 1
		self.sizeWithoutBorder[x] = n
		self.gridSize[x] = i <= self.dim and n + 2 * self.numGhost or 1
		self.numCells = self.numCells * self.gridSize[x]
		self.mins[i], self.maxs[i] = self.cfg.mins[i - 1], self.cfg.maxs[i - 1]
	end
	local h = ffi.new'hydro_solver*[1]'
	check(C.hydro_create(self.cfg, h))
	self.hip = ffi.gc(h[0], C.hydro_destroy)
	assert(tonumber(C.hydro_num_cells(self.hip)) == self.numCells)
	self.UBufObj = HostBuffer(self)
	self.UBuf = self.UBufObj.obj
	self:resetState()
end

-- SolverBase:resetState (solverbase.lua:2094-2114): initial state, boundary,
-- constrainU, t = 0
function HipFVSolver:resetState()
	local U, S = self.UBufObj.obj, self.eqn.numStates
	if self.initBuffer then
		ffi.copy(U, self.initBuffer, self.UBufObj.size)
	elseif self.initState then
		ffi.fill(U, self.UBufObj.size)
		local gs, sw = self.gridSize, self.sizeWithoutBorder
		local g = self.numGhost
		for k = 0, gs.z - 1 do
			for j = 0, gs.y - 1 do
				for i = 0, gs.x - 1 do
					local inside = i >= g and i < gs.x - g
						and (self.dim < 2 or (j >= g and j < gs.y - g))
						and (self.dim < 3 or (k >= g and k < gs.z - g))
					if inside then
						-- cell centres (coord.lua:1421-1436)
						local function centre(c, a) return (c - g + .5) / sw[({'x', 'y', 'z'})[a]] * (self.maxs[a] - self.mins[a]) + self.mins[a] end
						local x = centre(i, 1)
						local y = self.dim > 1 and centre(j, 2) or 0
						local z = self.dim > 2 and centre(k, 3) or 0
						self.initState(x, y, z, U + S * (i + gs.x * (j + gs.y * k)))
					end
				end
			end
		end
	else
		error'hipfvsolver needs args.initState or args.initBuffer (the reference\'s initial conditions are OpenCL code)'
	end
	self.UBufObj:fromCPU()
	check(C.hydro_constrain_U(self.hip))
	check(C.hydro_set_time(self.hip, 0))
	self.t, self.dt = 0, 0
end

function HipFVSolver:constrainU() check(C.hydro_constrain_U(self.hip)) end

-- GridSolver:boundary (gridsolver.lua:1316): ghosts are derived from the
-- interior when they are read, so nothing is filled here
function HipFVSolver:boundary() check(C.hydro_boundary(self.hip)) end

-- SolverBase:calcDT (solverbase.lua:3004-3023)
function HipFVSolver:calcDT()
	local dt = ffi.new'double[1]'
	check(C.hydro_calc_dt(self.hip, dt))
	return dt[0]
end

-- SolverBase:step (solverbase.lua:3193-3238): one integrator step, t unchanged
function HipFVSolver:step(dt) check(C.hydro_step(self.hip, dt)) end

-- SolverBase:update (solverbase.lua:3026-3190): calcDT, step, t += dt; dt
-- stays on the device. n updates in one call (default 1).
function HipFVSolver:update(n)
	check(C.hydro_update(self.hip, n or 1))
	local t, dt = ffi.new'double[1]', ffi.new'double[1]'
	check(C.hydro_get_time(self.hip, t, dt))
	self.t, self.dt = t[0], dt[0]
end

-- calcDisplayVarRangeAndAvgAndStdDev (solverbase.lua:2971-3002) of the
-- display var 'U <field>': min, max over the ghosted grid, average and
-- standard deviation over the interior
function HipFVSolver:calcDisplayVarRangeAndAvgAndStdDev(field)
	local out = ffi.new'double[4]'
	check(C.hydro_var_stats(self.hip, field, out))
	return out[0], out[1], out[2], out[3]
end

-- GridSolver:save / saveBuffer (gridsolver.lua:1410-1470): <prefix>_UBuf.fits,
-- one plane per cons_t channel, x fastest then z then y
function HipFVSolver:save(prefix)
	local Image = require 'image'
	local src = self.UBufObj:toCPU()
	local width, height, depth = self.gridSize.x, self.gridSize.y, self.gridSize.z
	local channels = self.eqn.numStates
	local image = Image(width * depth, height, channels, self.real)
	local dst = image.buffer
	for ch = 0, channels - 1 do
		for k = 0, depth - 1 do
			for j = 0, height - 1 do
				for i = 0, width - 1 do
					dst[i + width * (k + depth * (j + height * ch))] = src[ch + channels * (i + width * (j + height * k))]
				end
			end
		end
	end
	image:save((prefix and prefix..'_' or '')..'UBuf.fits')
end

return HipFVSolver
