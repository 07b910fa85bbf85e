/*
 * hydro_hip.h — C ABI of the MI355X finite-volume hot path.
 *
 * Plain C: opaque handles, int status codes (0 = ok, <0 = error, message in
 * hydro_last_error()), caller-owned host arrays, library-owned device memory,
 * one HIP stream per handle. A handle is thread-compatible, not thread-safe.
 * The same header is meant for LuaJIT `ffi.cdef` + `ffi.load` (the reference
 * binds C the same way, hydro/code/safecdef.lua:1-16) and for Python ctypes.
 *
 * Each entry point names the reference interface it replaces (file:line,
 * relative to the hydro-cl-lua tree).
 */
#ifndef HYDRO_HIP_H
#define HYDRO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest cons_t of the supported equations (GLM-Maxwell: 12 reals). */
#define HYDRO_MAX_STATES 12

/* Equation plugins: hydro/eqn/euler.lua (cons_t = rho, m.xyz, ETotal, ePot;
 * 5 integrated, 5 waves), hydro/eqn/glm-maxwell.lua (cons_t = D.xyz, B.xyz,
 * phi, psi, rhoCharge, sigma, sqrt_1_eps, sqrt_1_mu; 8 integrated, 8 waves)
 * and hydro/eqn/shallow-water.lua in its depthInCell = false form (cons_t =
 * h, m.xyz, depth; 4 integrated, 4 waves; shallow-water.lua:34-86). */
enum { HYDRO_EQN_EULER = 0, HYDRO_EQN_GLM_MAXWELL = 1, HYDRO_EQN_SHALLOW_WATER = 2 };

/* Flux plugins: hydro/flux/roe.lua, hll.lua ('Davis direct bounded' wave
 * estimates, hll.lua:10), rusanov.lua, euler-hllc.lua (hllcMethod 2,
 * euler-hllc.lua:18, with HLL's wave estimates) and marquina.lua. All but Roe
 * take the Euler equations and ignore the flux limiter (only Roe uses one,
 * fvsolver.lua:63). */
enum { HYDRO_FLUX_ROE = 0, HYDRO_FLUX_HLL = 1, HYDRO_FLUX_RUSANOV = 2, HYDRO_FLUX_HLLC = 3,
	HYDRO_FLUX_MARQUINA = 4 };

/* Limiter table in the order of HydroCLApp.limiters (hydro/app.lua:614-636),
 * 0-based: the reference's index 1 ('donor cell') is 0 here, and a flux limiter
 * is active when flux_limiter != HYDRO_LIMITER_DONOR_CELL (the reference's
 * `solver.fluxLimiter > 1`, hydro/flux/roe.cl:12). */
enum {
	HYDRO_LIMITER_DONOR_CELL = 0,
	HYDRO_LIMITER_LAX_WENDROFF,
	HYDRO_LIMITER_BEAM_WARMING,
	HYDRO_LIMITER_FROMM,
	HYDRO_LIMITER_CHARM,
	HYDRO_LIMITER_HCUS,
	HYDRO_LIMITER_HQUICK,
	HYDRO_LIMITER_KOREN,
	HYDRO_LIMITER_MINMOD,
	HYDRO_LIMITER_OSHKER,
	HYDRO_LIMITER_OSPRE,
	HYDRO_LIMITER_SMART,
	HYDRO_LIMITER_SWEBY,
	HYDRO_LIMITER_UMIST,
	HYDRO_LIMITER_VAN_ALBADA_1,
	HYDRO_LIMITER_VAN_ALBADA_2,
	HYDRO_LIMITER_VAN_LEER,
	HYDRO_LIMITER_MONOTIZED_CENTRAL,
	HYDRO_LIMITER_SUPERBEE,
	HYDRO_LIMITER_BARTH_JESPERSEN,
	HYDRO_NUM_LIMITERS
};

/* Reconstruction (solver.usePLM): none, 'plm cons' (hydro/solver/plm.cl:27-91),
 * or one of the Euler-only forms below. */
enum {
	HYDRO_PLM_NONE = 0,
	HYDRO_PLM_CONS = 1,
	HYDRO_PLM_PRIM = 2,   /* Euler, Roe flux (plm.cl:191-254) */
	HYDRO_PLM_EIG = 3,    /* Euler, Roe flux (plm.cl:256-432) */
	HYDRO_PLM_CONS_FLUX = 4,  /* 'plm cons with flux': Euler, Roe flux (plm.cl:95-188) */
	HYDRO_PLM_ATHENA = 5, /* 'plm athena': Euler, Roe flux (plm.cl:782-879) */
	HYDRO_PLM_PPM = 6,    /* 'ppm': Euler, Roe flux (plm.cl:882-944) */
	HYDRO_PLM_EIG_PRIM = 7,      /* 'plm eig prim': Euler, Roe flux (plm.cl:536-701) */
	HYDRO_PLM_EIG_PRIM_REF = 8   /* 'plm eig prim ref': Euler, Roe flux (plm.cl:536-615, 703-776) */
};

/* Explicit integrators, hydro/int/fe.lua and the Shu-Osher tables of
 * hydro/int/all.lua:10-160 (same names). Implicit ones are not provided. */
enum {
	HYDRO_INT_FORWARD_EULER = 0,
	HYDRO_INT_RK2,
	HYDRO_INT_RK2_HEUN,
	HYDRO_INT_RK2_RALSTON,
	HYDRO_INT_RK3,
	HYDRO_INT_RK4,
	HYDRO_INT_RK4_3_8THS,
	HYDRO_INT_RK2_TVD,
	HYDRO_INT_RK2_NON_TVD,
	HYDRO_INT_RK3_TVD,
	HYDRO_INT_RK4_TVD,
	HYDRO_INT_RK4_NON_TVD,
	HYDRO_NUM_INTEGRATORS
};

/* Grid boundaries (hydro/solver/gridsolver.lua:567-780). Faces are ordered
 * xmin, xmax, ymin, ymax, zmin, zmax. FIXED writes boundary_fixed[face] (a full
 * cons_t) into the ghost cells, the reference's user `fixedCode`. */
enum { HYDRO_BC_PERIODIC = 0, HYDRO_BC_MIRROR = 1, HYDRO_BC_FREEFLOW = 2, HYDRO_BC_FIXED = 3 };

/* Precision of the stored state. F64: everything in double. F32: state and
 * stencil arithmetic in float; dt min-reduction, conserved-total sums and the
 * RK stage accumulation in double. */
enum { HYDRO_REAL_F64 = 0, HYDRO_REAL_F32 = 1 };

/* Solver construction arguments: the subset of the reference's `args`
 * (hydro/solver/solverbase.lua:327-338, 775-812; gridsolver.lua:54-136;
 * fvsolver.lua:28-31) that the finite-volume path reads. */
typedef struct hydro_cfg {
	int dim;                 /* 1, 2, 3 */
	int size[3];             /* interior cells per axis (args.gridSize); unused axes 1 */
	double mins[3];          /* args.mins (default -1) */
	double maxs[3];          /* args.maxs (default 1) */
	int eqn;                 /* HYDRO_EQN_* */
	int flux;                /* HYDRO_FLUX_* */
	int flux_limiter;        /* HYDRO_LIMITER_*, args.fluxLimiter */
	int plm;                 /* HYDRO_PLM_*, args.usePLM */
	int slope_limiter;       /* HYDRO_LIMITER_*, args.slopeLimiter */
	int integrator;          /* HYDRO_INT_* */
	double cfl;              /* args.cfl */
	int use_fixed_dt;        /* args.fixedDT given */
	double fixed_dt;
	int boundary[6];         /* HYDRO_BC_* per face */
	double boundary_fixed[6][HYDRO_MAX_STATES];
	int precision;           /* HYDRO_REAL_* */
	double gamma;            /* Euler heatCapacityRatio (euler.lua:194) */
	double rho_min;          /* Euler rhoMin (euler.lua:195) */
	double P_min;            /* Euler PMin (euler.lua:196) */
	double div_phi_wavespeed;/* GLM-Maxwell divPhiWavespeed (glm-maxwell.lua:110) */
	double div_psi_wavespeed;/* GLM-Maxwell divPsiWavespeed (glm-maxwell.lua:111) */
	int fast_math;           /* 0: IEEE division and sqrt, the reference's arithmetic;
	                          * 1: faithfully rounded hardware sequences (FP64: rcp +
	                          * Newton + residual correction; FP32: v_rcp / v_sqrt) */
	int roe_entropy_fix;     /* Roe flux, Euler: Harten's entropy fix (args.useEntropyFluxFix,
	                          * roe.lua:10-16, roe.cl:83-106; default off) */
	double roe_entropy_fix_factor; /* args.entropyFluxFixLimitingFactor (roe.lua:11, 0.2) */
	double gravity;          /* shallow water: gravity (shallow-water.lua:167, default 1) */
} hydro_cfg;

/* Reference defaults: mins -1, maxs 1 (gridsolver.lua:80-81), gamma 7/5,
 * rhoMin = PMin = 1e-7 (euler.lua:192-197), gravity 1 (shallow-water.lua:167),
 * cfl .5, boundaries freeflow (gridsolver.lua:124), limiters donor cell, Roe,
 * forward Euler, F64. */
void hydro_cfg_default(hydro_cfg *cfg);

/* Last error message of the calling thread ("" if none). */
const char *hydro_last_error(void);

/* Number of reals in cons_t (numStates) and integrated (numIntStates). */
int hydro_eqn_num_states(int eqn);
int hydro_eqn_num_int_states(int eqn);

/* The Shu-Osher table of an integrator, as the reference's `alphas` / `betas`
 * (hydro/int/all.lua:10-160, read by rk.lua:47-167): *stages rows, alpha and
 * beta row-major 4 x 4 (row i: the stage's weights of U^(k), k <= i). Returns
 * -1 for an unknown integrator. Host only. */
int hydro_integrator_table(int integrator, int *stages, double *alpha, double *beta);

/* Number of device GPUs visible (0 when none). No other call needs one. */
int hydro_device_count(void);

/* ---------------------------------------------------------------- plain solver */

typedef struct hydro_solver hydro_solver;

/* FiniteVolumeSolver(args) (fvsolver.lua, solverbase.lua:351-395): validates the
 * configuration, allocates the device state. Errors like the reference's
 * asserts (e.g. PLM together with a flux limiter, gridsolver.lua:119). */
int hydro_create(const hydro_cfg *cfg, hydro_solver **out);
int hydro_destroy(hydro_solver *h);

/* Stored cells including numGhost=2 per used side (solver.numCells). */
size_t hydro_num_cells(const hydro_solver *h);

/* UBufObj:fromCPU / :toCPU. The host array is the reference's layout: cons_t
 * (numStates reals of the solver's precision) per cell, x fastest, ghost cells
 * included. Upload takes the interior; ghost cells of the download are filled
 * from the boundary conditions, as after the reference's boundary(). */
int hydro_upload_U(hydro_solver *h, const void *U, size_t bytes);
int hydro_download_U(hydro_solver *h, void *U, size_t bytes);

/* SolverBase:constrainU (solverbase.lua:2116-2127, euler.cl:697-717). */
int hydro_constrain_U(hydro_solver *h);

/* GridSolver:boundary (gridsolver.lua:1316). Ghost values are derived from the
 * interior when read, so this only exists for call-compatibility. */
int hydro_boundary(hydro_solver *h);

/* SolverBase:calcDT (solverbase.lua:3004-3023): dt = cfl * min over interior
 * cells (or fixedDT). Blocking: returns dt to the host. */
int hydro_calc_dt(hydro_solver *h, double *dt);

/* SolverBase:step(dt) (solverbase.lua:3193-3238): one integrator step with a
 * host-given dt. Does not advance t. */
int hydro_step(hydro_solver *h, double dt);

/* nsteps x SolverBase:update (solverbase.lua:3026-3190): calcDT, step, t += dt,
 * with dt kept on the device (no per-step readback). Asynchronous on the
 * solver stream; hydro_get_time synchronizes. */
int hydro_update(hydro_solver *h, int nsteps);

/* solver.t and solver.dt after the last update (synchronizes). */
int hydro_get_time(hydro_solver *h, double *t, double *dt);
int hydro_set_time(hydro_solver *h, double t);

/* Sum over interior cells of each integrated cons_t field, in double with
 * compensated (Neumaier) summation and a fixed reduction order.
 * out has numIntStates entries. */
int hydro_conserved_totals(hydro_solver *h, double *out);

/* The same totals of a state the caller holds on the device, with no device
 * memory allocated in the call. No counterpart in the reference: this is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). U: num_fields planes of num_cells
 * reals (precision HYDRO_REAL_*), plane_stride reals apart, the layout of the
 * solver's interior state; out: num_fields doubles on the device; stream: a
 * hipStream_t, NULL for the default stream. Asynchronous on that stream. The
 * per-block partial sums go to a 192 KiB library-static device array (one
 * per GPU, kept until the process ends): num_fields is at most
 * HYDRO_MAX_STATES, and calls that overlap on one GPU race on that array.
 * Same partition and order as hydro_conserved_totals, so the same bits. */
int hydro_conserved_totals_device(const void *U, int precision, int num_fields, long long plane_stride,
	long long num_cells, double *out, void *stream);

/* Derived display variables of the Euler equations (euler.lua:240-260), for
 * hydro_var_stats: W = primFromCons(U) per cell, values in double. */
enum {
	HYDRO_VAR_P = 64,          /* 'P' */
	HYDRO_VAR_E_INT_SPEC,      /* 'eInt' = P / (gamma - 1) / rho */
	HYDRO_VAR_E_KIN_SPEC,      /* 'eKin' = |v|^2 / 2 */
	HYDRO_VAR_E_TOTAL_SPEC,    /* 'eTotal' = ETotal / rho */
	HYDRO_VAR_E_INT,           /* 'EInt' = P / (gamma - 1) */
	HYDRO_VAR_E_KIN,           /* 'EKin' = rho |v|^2 / 2 */
	HYDRO_VAR_S,               /* 'S' = P / rho^gamma */
	HYDRO_VAR_H,               /* 'H' = P gamma / (gamma - 1) */
	HYDRO_VAR_H_SPEC,          /* 'h' = H / rho */
	HYDRO_VAR_H_TOTAL,         /* 'HTotal' = P + ETotal */
	HYDRO_VAR_H_TOTAL_SPEC,    /* 'hTotal' = HTotal / rho */
	HYDRO_VAR_CS,              /* 'speed of sound' (calc_Cs, euler.cl:87-94) */
	HYDRO_VAR_MACH,            /* 'Mach number' = |v| / Cs */
	HYDRO_VAR_V_X,             /* 'v x', 'v y', 'v z' */
	HYDRO_VAR_V_Y,
	HYDRO_VAR_V_Z,
	HYDRO_VAR_END
};

/* The display statistics of one stored cons_t field (0 .. numIntStates +
 * static fields - 1), or of a derived Euler variable (HYDRO_VAR_*),
 * calcDisplayVarRangeAndAvgAndStdDev
 * (hydro/solver/solverbase.lua:2871-2994) of the display var 'U <field>':
 * out[0], out[1] = min and max over every cell of the ghosted grid, ghosts as
 * the boundary fills them; out[2] = average and out[3] = population standard
 * deviation over the interior (the reference zeroes the ghosts for its sums
 * and divides by the interior count). Computed on the device in double
 * (synchronizes). */
int hydro_var_stats(hydro_solver *h, int field, double *out);

/* Wait for the solver's stream. */
int hydro_sync(hydro_solver *h);

/* Device timing of nsteps updates with HIP events on the solver stream.
 * total_ms: whole update loop; stage_ms: sum of the fused stage kernel
 * durations (events around each launch); n_stage_launches: their count. */
typedef struct hydro_timing {
	double total_ms;
	double stage_ms;
	int n_stage_launches;
	double dt_ms;            /* standalone dt kernels (first step only) */
	int n_dt_launches;
} hydro_timing;
int hydro_time_updates(hydro_solver *h, int nsteps, int per_kernel_events, hydro_timing *out);

/* Launch shape of the fused stage kernel (no counterpart in the reference,
 * whose work sizes are fixed at gridsolver.lua:141-243): rows each 64-lane
 * wave marches, 1..32 (3D: planes per block, 1..16), and whether blocks are remapped so that
 * each XCD works on a contiguous band of strips. 0 keeps the current value.
 * Results do not depend on them. */
int hydro_set_launch(hydro_solver *h, int rows_per_wave, int xcd_remap);
int hydro_get_launch(const hydro_solver *h, int *rows_per_wave, int *xcd_remap);

/* Bytes of the device state held by the handle. */
size_t hydro_device_bytes(const hydro_solver *h);

/* ---------------------------------------------------------------- batched grids */

/* Many independent grids advanced together: the reference runs several solvers
 * from one app loop, each with its own update (hydro/app.lua:1573-1737); here
 * every stage of every grid of the batch is one persistent launch (one wave per
 * 64-column strip, strips of all grids dealt to the waves in turn), each grid
 * keeps its own dt (cfl * its own min, solverbase.lua:3004-3023) and time, and
 * three steps (one turn of the dt-slot rotation) can be replayed from a
 * captured hipGraph. All grids share eqn, dim, flux, limiters, plm, integrator,
 * precision, math and equation parameters; sizes, extents, boundaries and cfl
 * may differ. Results of a grid do not depend on the other grids. */
typedef struct hydro_batch hydro_batch;

int hydro_batch_create(const hydro_cfg *cfgs, int n, hydro_batch **out);
int hydro_batch_destroy(hydro_batch *b);
int hydro_batch_size(const hydro_batch *b);
size_t hydro_batch_device_bytes(const hydro_batch *b);
/* per-grid UBufObj:fromCPU / :toCPU, same layout as hydro_upload_U */
int hydro_batch_upload_U(hydro_batch *b, int i, const void *U, size_t bytes);
int hydro_batch_download_U(hydro_batch *b, int i, void *U, size_t bytes);
int hydro_batch_constrain_U(hydro_batch *b);
/* nsteps updates of every active grid; use_graph replays captured steps */
int hydro_batch_update(hydro_batch *b, int nsteps, int use_graph);
int hydro_batch_get_time(hydro_batch *b, int i, double *t, double *dt);
/* inactive grids are skipped by the launches (their state and time stay) */
int hydro_batch_set_active(hydro_batch *b, int i, int active);
int hydro_batch_sync(hydro_batch *b);
/* device time of nsteps updates (HIP events on the batch stream) */
int hydro_batch_time_updates(hydro_batch *b, int nsteps, int use_graph, double *ms);
/* Launch shape (no counterpart in the reference): rows each wave marches in y
 * (1..32) and waves per block (1, 2, 4); 0 keeps the current value. */
int hydro_batch_set_launch(hydro_batch *b, int rows_per_wave, int waves_per_block);
int hydro_batch_get_launch(const hydro_batch *b, int *rows_per_wave, int *waves_per_block);
/* Autotuner (no counterpart in the reference): times trial_steps updates for
 * each rows-per-wave / waves-per-block candidate (the state advances) and keeps
 * the fastest; reports the choice and its ms per step. */
int hydro_batch_autotune(hydro_batch *b, int trial_steps, int *rows_out, int *waves_out, double *ms_per_step);

/* ---------------------------------------------------------------- solve server */

/* Continuous batching of solve requests (no counterpart in the reference; its
 * app loop steps a fixed set of solvers, hydro/app.lua:1573-1737). Requests
 * (configuration + initial state + step count) queue on submit. Admission
 * takes queued requests in arrival order while their device state fits the
 * HBM budget and their packed state fits the pinned host staging (the smaller
 * of the budget and 256 MiB); ones that do not fit wait, smaller ones may pass
 * them. Requests of one kernel configuration form a pool, which advances on
 * its own stream in chunks of up to steps_per_tick steps: the joining
 * requests' uploads, persistent batched launches over every resident grid,
 * and, for the requests the chunk finishes, a non-finite check on the device
 * and the copy of the final state to the host. Pools run concurrently. A
 * result equals the one of the same request solved alone. The server runs the
 * Runge-Kutta integrators. */
typedef struct hydro_server hydro_server;

typedef struct hydro_server_info {
	int queued;              /* waiting for memory */
	int resident;            /* on the device */
	int finished;            /* done, result not yet fetched */
	int pools;               /* kernel configurations seen */
	size_t bytes_in_use;
	size_t budget;
	long long ticks;
} hydro_server_info;

int hydro_server_create(size_t hbm_budget_bytes, int steps_per_tick, hydro_server **out);
int hydro_server_destroy(hydro_server *s);
/* U: initial ghosted cons_t buffer as for hydro_upload_U. Fails when the
 * request alone needs more device memory than the budget, or more host
 * staging than the server has. */
int hydro_server_submit(hydro_server *s, const hydro_cfg *cfg, const void *U, size_t bytes, int nsteps,
	long long *id);
/* Completes the chunks that have ended (results unpacked, memory released),
 * admits, and starts the next chunk of every idle pool; it waits only when
 * none of these can happen, for the chunk started first. n_completed: requests
 * finished in this call. */
int hydro_server_tick(hydro_server *s, int *n_completed);
/* ids of finished requests whose results were not fetched yet (up to max) */
int hydro_server_poll(hydro_server *s, long long *ids, int max, int *n);
/* final state (ghosted, as hydro_download_U), time, and status: 0 ok,
 * 1 non-finite values in the state. Releases the request. */
int hydro_server_result(hydro_server *s, long long id, void *U, size_t bytes, double *t, int *status);
int hydro_server_info_get(hydro_server *s, hydro_server_info *info);
/* launch shape of the pools, present and future (default: 12 rows per wave
 * for 2D pools, one wave per block; rows 1..32); 0 keeps a value. No
 * counterpart in the reference. */
int hydro_server_set_launch(hydro_server *s, int rows_per_wave, int waves_per_block);

#ifdef __cplusplus
}
#endif

#endif
