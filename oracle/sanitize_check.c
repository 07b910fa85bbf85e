/*
 * sanitize_check.c — TEST INFRASTRUCTURE. Drives the CPU oracle through every
 * configuration it restates (equations, reconstructions, fluxes, limiters,
 * integrators, boundaries, 1D/2D/3D, vacuum and cold states) plus its
 * per-interface functions, so that a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer (`make -C oracle sanitize`) reports any
 * out-of-bounds access, use after free, leak or undefined arithmetic in the
 * checker. Exits 0 when every case ran; the sanitizers abort otherwise.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hydro_oracle.h"

static unsigned long long rng_state = 0x9E3779B97F4A7C15ull;
static double urand(double lo, double hi) {
	rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
	return lo + (hi - lo) * (double)(rng_state >> 11) / 9007199254740992.0;
}

static void cfg_base(hydro_cfg *c, int dim, int nx, int ny, int nz) {
	memset(c, 0, sizeof *c);
	c->dim = dim;
	c->size[0] = nx; c->size[1] = dim > 1 ? ny : 1; c->size[2] = dim > 2 ? nz : 1;
	for (int i = 0; i < 3; ++i) { c->mins[i] = -1; c->maxs[i] = 1; }
	c->eqn = HYDRO_EQN_EULER;
	c->integrator = HYDRO_INT_RK2_TVD;
	c->cfl = .4;
	for (int f = 0; f < 6; ++f) c->boundary[f] = HYDRO_BC_FREEFLOW;
	c->gamma = 1.4; c->rho_min = 1e-7; c->P_min = 1e-7;
	c->div_phi_wavespeed = 1; c->div_psi_wavespeed = 1;
	c->roe_entropy_fix_factor = .2;
}

/* kind 0: random admissible Euler states; 1: a gas expanding into vacuum;
 * 2: a cold gas (P below PMin); GLM-Maxwell: random fields, eps = mu = 1 */
static void fill(oracle_solver *s, const hydro_cfg *c, int kind) {
	double *U = oracle_U(s);
	size_t n = oracle_num_cells(s);
	int S = oracle_num_states(s);
	for (size_t i = 0; i < n; ++i) {
		double *u = U + i * S;
		if (c->eqn == HYDRO_EQN_GLM_MAXWELL) {
			for (int j = 0; j < 8; ++j) u[j] = urand(-1, 1);
			u[8] = 0; u[9] = urand(0, 1); u[10] = 1; u[11] = 1;
			continue;
		}
		double rho = urand(.1, 2), P = urand(.1, 2), v[3];
		for (int j = 0; j < 3; ++j) v[j] = urand(-.5, .5);
		if (kind == 1 && (i % (size_t)(c->size[0] + 4)) > (size_t)(c->size[0] + 4) / 2) { rho = 1e-9; P = 1e-9; }
		if (kind == 2) P = 1e-9;
		u[0] = rho;
		for (int j = 0; j < 3; ++j) u[1 + j] = rho * v[j];
		u[4] = rho * .5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + P / (c->gamma - 1);
		u[5] = 0;
	}
}

static int run_case(const hydro_cfg *c, int kind, int steps) {
	oracle_solver *s = oracle_create(c);
	if (!s) {
		fprintf(stderr, "oracle_create refused a case: %s\n", oracle_last_error());
		return 1;
	}
	fill(s, c, kind);
	oracle_boundary(s);
	oracle_constrain_U(s);
	for (int i = 0; i < steps; ++i) oracle_update(s);
	double st[4];
	oracle_var_stats(s, 0, st);
	double *d = (double *)malloc(oracle_num_cells(s) * oracle_num_states(s) * sizeof(double));
	oracle_calc_deriv(s, 1e-3, d);
	free(d);
	oracle_destroy(s);
	return 0;
}

int main(void) {
	int bad = 0, cases = 0;
	hydro_cfg c;
	oracle_set_threads(2);
	const int plms[] = {HYDRO_PLM_NONE, HYDRO_PLM_CONS, HYDRO_PLM_PRIM, HYDRO_PLM_EIG, HYDRO_PLM_CONS_FLUX,
		HYDRO_PLM_ATHENA, HYDRO_PLM_PPM, HYDRO_PLM_EIG_PRIM, HYDRO_PLM_EIG_PRIM_REF};
	for (int dim = 1; dim <= 3; ++dim) {
		for (size_t k = 0; k < sizeof plms / sizeof plms[0]; ++k) {
			cfg_base(&c, dim, 9, 7, 5);
			c.plm = plms[k];
			c.slope_limiter = HYDRO_LIMITER_MINMOD;
			bad |= run_case(&c, 0, 2); ++cases;
		}
		for (int flux = HYDRO_FLUX_ROE; flux <= HYDRO_FLUX_MARQUINA; ++flux) {
			cfg_base(&c, dim, 8, 6, 5);
			c.flux = flux;
			c.plm = HYDRO_PLM_CONS;
			c.slope_limiter = HYDRO_LIMITER_VAN_LEER;
			bad |= run_case(&c, 0, 2); ++cases;
			bad |= run_case(&c, 1, 2); ++cases;
		}
		for (int lim = 0; lim < HYDRO_NUM_LIMITERS; ++lim) {
			cfg_base(&c, dim, 7, 6, 4);
			c.flux_limiter = lim;
			c.boundary[0] = HYDRO_BC_PERIODIC; c.boundary[1] = HYDRO_BC_PERIODIC;
			c.boundary[2] = HYDRO_BC_MIRROR; c.boundary[3] = HYDRO_BC_FIXED;
			for (int j = 0; j < 5; ++j) c.boundary_fixed[3][j] = 1 + j;
			bad |= run_case(&c, 0, 1); ++cases;
		}
		cfg_base(&c, dim, 10, 6, 4);
		c.roe_entropy_fix = 1;
		bad |= run_case(&c, 0, 2); ++cases;
		cfg_base(&c, dim, 10, 6, 4);
		c.flux = HYDRO_FLUX_HLL;
		c.plm = HYDRO_PLM_CONS;
		bad |= run_case(&c, 2, 3); ++cases;
	}
	for (int integ = 0; integ < HYDRO_NUM_INTEGRATORS; ++integ) {
		cfg_base(&c, 2, 8, 8, 1);
		c.integrator = integ;
		c.plm = HYDRO_PLM_CONS;
		c.slope_limiter = HYDRO_LIMITER_SUPERBEE;
		bad |= run_case(&c, 0, 2); ++cases;
	}
	for (int dim = 1; dim <= 2; ++dim) {
		for (int plm = 0; plm < 3; ++plm) {
			cfg_base(&c, dim, 9, 7, 1);
			c.eqn = HYDRO_EQN_GLM_MAXWELL;
			c.plm = plm == 0 ? HYDRO_PLM_NONE : plm == 1 ? HYDRO_PLM_CONS : HYDRO_PLM_PPM;
			c.slope_limiter = HYDRO_LIMITER_MINMOD;
			for (int f = 0; f < 4; ++f) c.boundary[f] = f < 2 ? HYDRO_BC_PERIODIC : HYDRO_BC_MIRROR;
			bad |= run_case(&c, 0, 2); ++cases;
		}
		cfg_base(&c, dim, 9, 7, 1);
		c.eqn = HYDRO_EQN_GLM_MAXWELL;
		c.flux_limiter = HYDRO_LIMITER_SUPERBEE;
		bad |= run_case(&c, 0, 2); ++cases;
	}

	/* per-interface functions on random and edge-case states */
	oracle_params p = {1.4, 1e-7, 1e-7, 1, 1, 0, .2};
	for (int it = 0; it < 200; ++it) {
		double UL[12], UR[12], UM[12], UP[12], e[10], X[6], Y[6], F[12], L[12], R[12];
		double *st[4] = {UM, UL, UR, UP};
		for (int k = 0; k < 4; ++k) {
			double rho = it % 7 == 0 ? 1e-6 : urand(.1, 2), P = it % 11 == 0 ? 1e-9 : urand(.1, 2);
			double v[3] = {urand(-1, 1), urand(-1, 1), urand(-1, 1)};
			double W[6] = {rho, v[0], v[1], v[2], P, 0};
			oracle_euler_cons_from_prim(&p, W, st[k]);
		}
		for (int side = 0; side < 3; ++side) {
			oracle_euler_eigen_for_interface(&p, UL, UR, e);
			for (int j = 0; j < 6; ++j) X[j] = urand(-1, 1);
			oracle_euler_left(&p, e, X, side, Y);
			oracle_euler_right(&p, e, X, side, Y);
			oracle_euler_waves(&p, e, side, Y);
			oracle_euler_flux_from_cons(&p, UL, side, F);
			oracle_roe_flux(&p, HYDRO_EQN_EULER, HYDRO_LIMITER_SUPERBEE, .3, UL, UR,
				(const double *const *)st, side, F);
			oracle_hll_flux(&p, UL, UR, side, F);
			oracle_hllc_flux(&p, UL, UR, side, F);
			oracle_rusanov_flux(&p, UL, UR, side, F);
			oracle_marquina_flux(&p, UL, UR, side, F);
			oracle_plm_eig(&p, HYDRO_LIMITER_MINMOD, UM, UL, UR, side, .3, L, R);
			oracle_plm_eig_prim(&p, it & 1, UM, UL, UR, side, .3, L, R);
			oracle_plm_athena(&p, UM, UL, UR, side, L, R);
			oracle_plm_cons_flux(&p, HYDRO_LIMITER_VAN_LEER, UM, UL, UR, side, .3, L, R);
		}
		oracle_plm_prim(&p, HYDRO_LIMITER_SUPERBEE, UM, UL, UR, L, R);
		oracle_plm_cons(HYDRO_LIMITER_MINMOD, 5, 6, UM, UL, UR, L, R);
		oracle_plm_ppm(HYDRO_LIMITER_MINMOD, 5, 6, UM, UL, UR, L, R);
		double dx[3] = {.1, .2, .3};
		(void)oracle_euler_dt_cell(&p, UL, 3, dx);
		for (int id = HYDRO_VAR_P; id < HYDRO_VAR_END; ++id) (void)oracle_euler_display(&p, id, UL);
		oracle_euler_constrain(&p, UL);
		for (int j = 0; j < 12; ++j) { UL[j] = urand(-1, 1); UR[j] = urand(-1, 1); }
		UL[10] = UR[10] = UL[11] = UR[11] = 1;
		for (int side = 0; side < 3; ++side) {
			oracle_glm_eigen_for_interface(&p, UL, UR, e);
			double X8[12], Y8[12];
			for (int j = 0; j < 12; ++j) X8[j] = urand(-1, 1);
			oracle_glm_left(&p, e, X8, side, Y8);
			oracle_glm_right(&p, e, X8, side, Y8);
			oracle_glm_waves(&p, e, side, Y8);
			oracle_glm_flux_from_cons(&p, UL, side, F);
			oracle_roe_flux(&p, HYDRO_EQN_GLM_MAXWELL, HYDRO_LIMITER_DONOR_CELL, .3, UL, UR, NULL, side, F);
		}
	}
	for (int i = 0; i < HYDRO_NUM_LIMITERS; ++i)
		for (double r = -3; r <= 3; r += .25) (void)oracle_limiter(i, r);
	printf("sanitize_check: %d solver cases, %s\n", cases, bad ? "some refused" : "all ran");
	return bad;
}
