/*
 * hydro_oracle.h — TEST INFRASTRUCTURE. CPU restatement of hydro-cl-lua's
 * structured-grid finite-volume path, used only by tests/, the smoke check and
 * bench.py's cpu_baseline leg as the checker. Never linked into the product.
 *
 * FP64, the reference's layout (AoS cons_t with numGhost=2 ghost cells, x
 * fastest) and the reference's kernel sequence with unfused buffers:
 * calcLR -> calcFlux -> calcDerivFromFlux -> (addSource) -> multAdd ->
 * boundary -> constrainU, and calcDT + min. OpenMP over cells.
 */
#ifndef HYDRO_ORACLE_H
#define HYDRO_ORACLE_H

#include <stddef.h>
#include "../include/hydro_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solver oracle_solver;

/* equation parameters the per-cell functions read from solver_t */
typedef struct oracle_params {
	double gamma, rho_min, P_min, div_phi, div_psi;
	int roe_fix;                 /* Roe, Euler: Harten's entropy fix (roe.cl:83-106) */
	double roe_fix_factor;
	double gravity;              /* shallow water */
} oracle_params;

oracle_solver *oracle_create(const hydro_cfg *cfg);
void oracle_destroy(oracle_solver *s);
const char *oracle_last_error(void);

size_t oracle_num_cells(const oracle_solver *s);
int oracle_num_states(const oracle_solver *s);
double *oracle_U(oracle_solver *s);           /* AoS, ghosts included */
void oracle_set_threads(int n);
int oracle_get_threads(void);

void oracle_boundary(oracle_solver *s);
/* calcDisplayVarRangeAndAvgAndStdDev (solverbase.lua:2871-2994) of cons_t
 * field `field` of the ghosted state: out = {min, max over all cells, average,
 * population standard deviation over the interior} */
/* a derived Euler display variable (HYDRO_VAR_*, euler.lua:240-260) of one cons_t */
double oracle_euler_display(const oracle_params *p, int id, const double *U);
void oracle_var_stats(const oracle_solver *s, int field, double *out);
void oracle_constrain_U(oracle_solver *s);      /* constrainU kernel + boundary */
double oracle_calc_dt(oracle_solver *s);        /* cfl * min (or fixed dt) */
void oracle_step(oracle_solver *s, double dt);  /* integrator step */
double oracle_update(oracle_solver *s);         /* calcDT, step, t += dt, boundary */
double oracle_time(const oracle_solver *s);
void oracle_set_time(oracle_solver *s, double t);
/* calcDeriv into deriv (AoS, G*S); deriv is cleared first */
void oracle_calc_deriv(oracle_solver *s, double dt, double *deriv);

/* ---- per-interface / per-cell functions, for the pinning tests ---- */
double oracle_limiter(int index, double r);
/* Euler */
void oracle_euler_prim_from_cons(const oracle_params *p, const double *U, double *W);
void oracle_euler_cons_from_prim(const oracle_params *p, const double *W, double *U);
void oracle_euler_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *eig);
void oracle_euler_left(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_euler_right(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_euler_flux_from_cons(const oracle_params *p, const double *U, int side, double *F);
void oracle_euler_waves(const oracle_params *p, const double *eig, int side, double *lambda);
double oracle_euler_dt_cell(const oracle_params *p, const double *U, int dim, const double *dx);
void oracle_euler_constrain(const oracle_params *p, double *U);
/* GLM-Maxwell */
void oracle_glm_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *eig);
void oracle_glm_left(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_glm_right(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_glm_flux_from_cons(const oracle_params *p, const double *U, int side, double *F);
void oracle_glm_waves(const oracle_params *p, const double *eig, int side, double *lambda);
/* shallow water (depth in cons_t: h, m.xyz, depth; eigen h, v.xyz, C, depth) */
void oracle_sw_prim_from_cons(const oracle_params *p, const double *U, double *W);
void oracle_sw_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *eig);
void oracle_sw_left(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_sw_right(const oracle_params *p, const double *eig, const double *X, int side, double *Y);
void oracle_sw_flux_from_cons(const oracle_params *p, const double *U, int side, double *F);
void oracle_sw_waves(const oracle_params *p, const double *eig, int side, double *lambda);
double oracle_sw_dt_cell(const oracle_params *p, const double *U, int dim, const double *dx);
/* HLL ('Davis direct bounded', hll.cl:13-73) and Rusanov (rusanov.cl:13-39)
 * fluxes of the Euler equations for one interface; F gets 6 reals. */
void oracle_hll_flux(const oracle_params *p, const double *UL, const double *UR, int side, double *F);
void oracle_rusanov_flux(const oracle_params *p, const double *UL, const double *UR, int side, double *F);
/* HLLC (euler-hllc.cl:13-240, hllcMethod 2) of the Euler equations */
void oracle_hllc_flux(const oracle_params *p, const double *UL, const double *UR, int side, double *F);
/* eigen_forCell (euler.cl:315-341) and the Marquina flux (marquina.cl:5-93) */
void oracle_euler_eigen_for_cell(const oracle_params *p, const double *U, double *eig);
void oracle_marquina_flux(const oracle_params *p, const double *UL, const double *UR, int side, double *F);

/* Roe flux for one interface (roe.cl:17-163). lim_args may be NULL when no
 * flux limiter: otherwise {UL_L, UR_L, UL_R, UR_R} pointers. */
void oracle_roe_flux(const oracle_params *p, int eqn, int flux_limiter, double dt_dx,
	const double *UL, const double *UR, const double *const *lim_args, int side, double *F);
/* plm cons edge states of one cell along one axis (plm.cl:27-91) */
/* 'plm prim' edge states (plm.cl:191-254), Euler; L and R get 6 reals */
void oracle_plm_prim(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	double *L, double *R);
/* 'plm eig' edge states (plm.cl:256-432), Euler, along axis side with
 * dt_dx = dt / dx; L and R get 6 reals */
void oracle_plm_eig(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	int side, double dt_dx, double *L, double *R);
/* 'plm cons with flux' edge states (plm.cl:95-188), Euler, along axis side
 * with dt_dx = dt / dx; L and R get 6 reals */
void oracle_plm_cons_flux(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	int side, double dt_dx, double *L, double *R);
/* 'plm athena' edge states (plm.cl:782-879), Euler, along axis side; L and R
 * get 6 reals */
void oracle_plm_athena(const oracle_params *p, const double *UL, const double *U, const double *UR,
	int side, double *L, double *R);
/* 'plm eig prim' (ref = 0) and 'plm eig prim ref' (ref = 1) edge states
 * (plm.cl:536-780), Euler, along axis side with dt_dx = dt / dx */
void oracle_plm_eig_prim(const oracle_params *p, int ref, const double *UL, const double *U, const double *UR,
	int side, double dt_dx, double *L, double *R);
/* 'ppm' edge states (plm.cl:882-944) */
void oracle_plm_ppm(int slope_limiter, int nint, int nstates,
	const double *UL, const double *U, const double *UR, double *L, double *R);
void oracle_plm_cons(int slope_limiter, int nint, int nstates,
	const double *UL, const double *U, const double *UR, double *L, double *R);

#ifdef __cplusplus
}
#endif

#endif
