/*
 * hydro_oracle.c — TEST INFRASTRUCTURE ONLY (the checker; never the product).
 *
 * A plain-C, FP64 restatement of hydro-cl-lua's structured-grid finite-volume
 * path for Cartesian coordinates, as the reference's templates expand it:
 *   Euler eqn        hydro/eqn/euler.cl:1-173, 271-291, 343-542, 697-717
 *   GLM-Maxwell eqn  hydro/eqn/glm-maxwell.cl:96-380, glm-maxwell.lua:181-245
 *   Roe flux         hydro/flux/roe.cl:17-163
 *   HLL, Rusanov     hydro/flux/hll.cl:13-73 (hll.lua:10), rusanov.cl:13-39
 *   HLLC             hydro/flux/euler-hllc.cl:13-240 (hllcMethod 2)
 *   Marquina         hydro/flux/marquina.cl:5-93, eigen_forCell euler.cl:315-341
 *   plm cons, prim   hydro/solver/plm.cl:27-91, 191-254, 976-997
 *   calcFlux         hydro/solver/fvsolver.lua:55-199
 *   calcDerivFromFlux hydro/solver/fvsolver.cl:1-125
 *   calcDT           hydro/eqn/cl/calcDT.cl:33-73, eqn.lua:1182-1227,
 *                    solverbase.lua:3004-3023
 *   boundaries       hydro/solver/gridsolver.lua:567-780, 1070-1320
 *   multAdd(Into)    hydro/solver/solverbase.lua:1216-1248
 *   integrators      hydro/int/fe.lua, rk.lua, all.lua
 *   limiters         hydro/app.lua:614-636
 * Layout and order of operations follow the reference: AoS cons_t, numGhost=2,
 * unfused buffers, left-to-right evaluation as written, no FP contraction
 * (built with -ffp-contract=off). Normals are Cartesian (`int side`,
 * coord/coord.lua:2342-2410): normal_len = 1, normal_lj = e_(side+j-1)%3.
 *
 * Parity status of each function is recorded in the tests that pin it.
 */
#include "hydro_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double real;
#define NG 2

static __thread char g_err[512];
const char *oracle_last_error(void) { return g_err; }
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }

void oracle_set_threads(int n) {
#ifdef _OPENMP
	omp_set_num_threads(n);
#else
	(void)n;
#endif
}
int oracle_get_threads(void) {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

/* ------------------------------------------------------------------ limiters */
/* hydro/app.lua:614-636. OpenCL max/min on finite values are fmax/fmin.
 * van Leer: the reference's `max(0., r) * 2. / (1. + r)` (app.lua:631) is 0/0
 * at r = -1; this uses `r > 0 ? r * 2. / (1. + r) : 0.`, identical for r > 0
 * and 0 (up to the sign of zero) elsewhere. */
double oracle_limiter(int index, double r) {
	switch (index) {
	case 0: return 0.;
	case 1: return 1.;
	case 2: return r;
	case 3: return .5 * (1. + r);
	case 4: return fmax(0., r) * (3. * r + 1.) / ((r + 1.) * (r + 1.));
	case 5: return fmax(0., 1.5 * (r + fabs(r)) / (r + 2.));
	case 6: return fmax(0., 2. * (r + fabs(r)) / (r + 3.));
	case 7: return fmax(0., fmin(2. * r, fmin((1. + 2. * r) / 3., 2.)));
	case 8: return fmax(0., fmin(r, 1.));
	case 9: return fmax(0., fmin(r, 1.5));
	case 10: return .5 * (r * r + r) / (r * r + r + 1.);
	case 11: return fmax(0., fmin(2. * r, fmin(.25 + .75 * r, 4.)));
	case 12: return fmax(0., fmax(fmin(1.5 * r, 1.), fmin(r, 1.5)));
	case 13: return fmax(0., fmin(fmin(2. * r, .75 + .25 * r), fmin(.25 + .75 * r, 2.)));
	case 14: return (r * r + r) / (r * r + 1.);
	case 15: return 2. * r / (r * r + 1.);
	case 16: return r > 0. ? r * 2. / (1. + r) : 0.;
	case 17: return fmax(0., fmin(2., fmin(.5 * (1. + r), 2. * r)));
	case 18: return fmax(0., fmax(fmin(1., 2. * r), fmin(2., r)));
	case 19: return .5 * (r + 1.) * fmin(1., fmin(4. * r / (r + 1.), 4. / (r + 1.)));
	}
	return 0.;
}

/* -------------------------------------------------------------- normals */
/* normal_l<j><i>(n) = (n.side == (i-j)%3 ? 1. : 0.)  (coord.lua:2376-2385) */
static inline real nl(int j, int i, int side) {
	int m = ((i - j) % 3 + 3) % 3;
	return side == m ? 1. : 0.;
}
static inline real dot3(const real *a, const real *b) {
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ------------------------------------------------------------------ Euler */
/* cons: 0 rho, 1-3 m, 4 ETotal, 5 ePot; prim: 0 rho, 1-3 v, 4 P, 5 ePot */
/* eigen: 0 rho, 1-3 v, 4 hTotal, 5 Cs, 6 vSq, 7-9 vL (euler.lua:298-307) */
enum { E_RHO = 0, E_HT = 4, E_CS = 5, E_VSQ = 6, E_VL = 7 };

static inline real euler_calc_P(const oracle_params *p, const real *U) {
	/* euler.cl:63-83 */
	real EKin = U[0] < p->rho_min ? 0. : (.5 * dot3(U + 1, U + 1) / U[0]);
	return U[0] < p->rho_min ? 0. : ((p->gamma - 1.) * (U[4] - EKin));
}

void oracle_euler_prim_from_cons(const oracle_params *p, const double *U, double *W) {
	/* euler.cl:140-159: the rhoMin branch is overwritten unconditionally */
	real inv = 1. / U[0];
	W[0] = U[0];
	W[1] = U[1] * inv;
	W[2] = U[2] * inv;
	W[3] = U[3] * inv;
	W[4] = euler_calc_P(p, U);
	W[5] = U[5];
}

void oracle_euler_cons_from_prim(const oracle_params *p, const double *W, double *U) {
	/* euler.cl:163-173, calc_ETotal = calc_EKin + calc_EInt (euler.cl:41-75) */
	U[0] = W[0];
	U[1] = W[1] * W[0];
	U[2] = W[2] * W[0];
	U[3] = W[3] * W[0];
	U[4] = W[0] * (.5 * dot3(W + 1, W + 1)) + W[4] / (p->gamma - 1.);
	U[5] = W[5];
}

static inline real euler_calc_Cs(const oracle_params *p, const real *W) {
	/* euler.cl:87-94 */
	if (W[4] <= p->P_min) return 0.;
	if (W[0] < p->rho_min) return INFINITY;
	return sqrt(p->gamma * W[4] / W[0]);
}

void oracle_euler_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *e) {
	/* euler.cl:347-430 */
	const real rhoEpsilon = 1e-5;
	if (UL[0] < rhoEpsilon && UR[0] < rhoEpsilon) {
		for (int i = 0; i < 10; ++i) e[i] = 0.;
		return;
	}
	if (UL[0] < rhoEpsilon || UR[0] < rhoEpsilon) {
		const real *Us = UL[0] < rhoEpsilon ? UR : UL;
		real W[6];
		oracle_euler_prim_from_cons(p, Us, W);
		e[0] = Us[0];
		e[1] = W[1]; e[2] = W[2]; e[3] = W[3];
		e[7] = W[1]; e[8] = W[2]; e[9] = W[3];
		e[E_VSQ] = dot3(e + 1, e + 7);
		e[E_HT] = (W[4] + Us[4]) / W[0];
		e[E_CS] = euler_calc_Cs(p, W);
		return;
	}
	real WL[6], WR[6];
	oracle_euler_prim_from_cons(p, UL, WL);
	real sqrtRhoL = sqrt(WL[0]);
	real hTotalL = (WL[4] + UL[4]) / WL[0];
	oracle_euler_prim_from_cons(p, UR, WR);
	real sqrtRhoR = sqrt(WR[0]);
	real hTotalR = (WR[4] + UR[4]) / WR[0];
	real invDenom = 1. / (sqrtRhoL + sqrtRhoR);
	e[0] = sqrtRhoL * sqrtRhoR;
	real v[3];
	for (int i = 0; i < 3; ++i)
		v[i] = WL[1 + i] * (sqrtRhoL * invDenom) + WR[1 + i] * (sqrtRhoR * invDenom);
	real hTotal = invDenom * (sqrtRhoL * hTotalL + sqrtRhoR * hTotalR);
	real vSq = dot3(v, v);
	real eKin = .5 * vSq;
	real h = hTotal - eKin;
	if (h < rhoEpsilon) {
		e[E_HT] = eKin;
		e[E_CS] = 0.;
	} else {
		e[E_HT] = hTotal;
		real CsSq = h < rhoEpsilon ? 0. : (p->gamma - 1.) * h;
		e[E_CS] = sqrt(CsSq);
	}
	e[1] = v[0]; e[2] = v[1]; e[3] = v[2];
	e[E_VSQ] = vSq;
	e[7] = v[0]; e[8] = v[1]; e[9] = v[2];
}

void oracle_euler_left(const oracle_params *p, const double *e, const double *X, int s, double *Y) {
	/* euler.cl:434-489 */
	if (e[0] < p->rho_min) {
		for (int i = 0; i < 5; ++i) Y[i] = X[i];
		return;
	}
	const real *v = e + 1, *vL = e + 7;
	real vn[3] = {v[s], v[(s + 1) % 3], v[(s + 2) % 3]};
	real nLen = 1.;
	real inv_nLen = 1. / nLen;
	real Cs = e[E_CS], vSq = e[E_VSQ];
	real denom = 2. * Cs * Cs;
	real invDenom = 1. / denom;
	real g1 = p->gamma - 1.;
	Y[0] = (X[0] * (.5 * g1 * vSq + Cs * vn[0] * inv_nLen)
		+ X[1] * (-g1 * vL[0] - Cs * nl(1, 1, s))
		+ X[2] * (-g1 * vL[1] - Cs * nl(1, 2, s))
		+ X[3] * (-g1 * vL[2] - Cs * nl(1, 3, s))
		+ X[4] * g1) * invDenom;
	Y[1] = (X[0] * (denom - g1 * vSq)
		+ X[1] * 2. * g1 * vL[0]
		+ X[2] * 2. * g1 * vL[1]
		+ X[3] * 2. * g1 * vL[2]
		+ X[4] * -2. * g1) * invDenom;
	Y[2] = X[0] * -vn[1]
		+ X[1] * nl(2, 1, s)
		+ X[2] * nl(2, 2, s)
		+ X[3] * nl(2, 3, s);
	Y[3] = X[0] * -vn[2]
		+ X[1] * nl(3, 1, s)
		+ X[2] * nl(3, 2, s)
		+ X[3] * nl(3, 3, s);
	Y[4] = (X[0] * (.5 * g1 * vSq - Cs * vn[0] * inv_nLen)
		+ X[1] * (-g1 * vL[0] + Cs * nl(1, 1, s))
		+ X[2] * (-g1 * vL[1] + Cs * nl(1, 2, s))
		+ X[3] * (-g1 * vL[2] + Cs * nl(1, 3, s))
		+ X[4] * g1) * invDenom;
}

void oracle_euler_right(const oracle_params *p, const double *e, const double *X, int s, double *Y) {
	/* euler.cl:493-542 */
	if (e[0] < p->rho_min) {
		for (int i = 0; i < 5; ++i) Y[i] = X[i];
	} else {
		const real *v = e + 1;
		real vn[3] = {v[s], v[(s + 1) % 3], v[(s + 2) % 3]};
		real nLen = 1.;
		real inv_nLen = 1. / nLen;
		real Cs = e[E_CS], hTotal = e[E_HT], vSq = e[E_VSQ];
		Y[0] = X[0] + X[1] + X[4];
		for (int i = 0; i < 3; ++i) {
			Y[1 + i] = X[0] * (v[i] - Cs * nl(1, 1 + i, s))
				+ X[1] * v[i]
				+ X[2] * nl(2, 1 + i, s)
				+ X[3] * nl(3, 1 + i, s)
				+ X[4] * (v[i] + Cs * nl(1, 1 + i, s));
		}
		Y[4] = X[0] * (hTotal - Cs * vn[0] * inv_nLen)
			+ X[1] * .5 * vSq
			+ X[2] * vn[1]
			+ X[3] * vn[2]
			+ X[4] * (hTotal + Cs * vn[0] * inv_nLen);
	}
	Y[5] = 0;
}

void oracle_euler_flux_from_cons(const oracle_params *p, const double *U, int s, double *F) {
	/* euler.cl:273-291 */
	real W[6];
	oracle_euler_prim_from_cons(p, U, W);
	real v_n = W[1 + s];
	F[0] = U[0] * v_n;
	for (int i = 0; i < 3; ++i) F[1 + i] = U[1 + i] * v_n + nl(1, 1 + i, s) * W[4];
	real HTotal = U[4] + W[4];
	F[4] = HTotal * v_n;
	F[5] = 0;
}

void oracle_euler_waves(const oracle_params *p, const double *e, int s, double *lambda) {
	/* euler.lua:309-325 */
	(void)p;
	real Cs_nLen = 1. * e[E_CS];
	real v_n = e[1 + s];
	lambda[0] = v_n - Cs_nLen;
	lambda[1] = v_n;
	lambda[2] = v_n;
	lambda[3] = v_n;
	lambda[4] = v_n + Cs_nLen;
}

double oracle_euler_dt_cell(const oracle_params *p, const double *U, int dim, const double *dx) {
	/* calcDT.cl:38-73 with euler.lua:346-373 */
	real P = euler_calc_P(p, U);
	real Cs;
	if (P <= p->P_min) Cs = 0.;
	else if (U[0] < p->rho_min) Cs = INFINITY;
	else Cs = sqrt(p->gamma * P / U[0]);
	real dt = INFINITY;
	for (int s = 0; s < dim; ++s) {
		real d = dx[s];
		if (d > 0) {
			real Cs_nLen = Cs * 1.;
			real v_n = U[0] < p->rho_min ? 0. : U[1 + s] / U[0];
			real lmin = v_n - Cs_nLen;
			real lmax = v_n + Cs_nLen;
			real absLambdaMax = fmax(fabs(lmin), fabs(lmax));
			absLambdaMax = fmax(1e-9, absLambdaMax);
			dt = fmin(dt, d / absLambdaMax);
		}
	}
	return dt;
}

void oracle_euler_constrain(const oracle_params *p, double *U) {
	/* euler.cl:697-717 */
	if (U[0] < p->rho_min) U[0] = p->rho_min;
	real W[6];
	oracle_euler_prim_from_cons(p, U, W);
	if (W[4] < p->P_min) W[4] = p->P_min;
	oracle_euler_cons_from_prim(p, W, U);
}

/* ------------------------------------------------------------ GLM-Maxwell */
/* cons: 0-2 D, 3-5 B, 6 phi, 7 psi, 8 rhoCharge, 9 sigma, 10 sqrt_1_eps, 11 sqrt_1_mu */
/* eigen: 0 sqrt_1_eps, 1 sqrt_1_mu */

void oracle_glm_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *e) {
	/* glm-maxwell.cl:138-153 */
	(void)p;
	e[0] = (UL[10] + UR[10]) * .5;
	e[1] = (UL[11] + UR[11]) * .5;
}

void oracle_glm_left(const oracle_params *p, const double *eig, const double *X, int s, double *Y) {
	/* glm-maxwell.cl:175-252 */
	(void)p;
	real n_l[3] = {nl(1, 1, s), nl(1, 2, s), nl(1, 3, s)};
	real n2_l[3] = {nl(2, 1, s), nl(2, 2, s), nl(2, 3, s)};
	real n3_l[3] = {nl(3, 1, s), nl(3, 2, s), nl(3, 3, s)};
	real e = eig[0], m = eig[1];
	real tmp1 = 1. / e;
	real tmp2 = tmp1 * n_l[0];
	real tmp3 = 1. / 2.;
	real tmp4 = X[0] * tmp2;
	real tmp5 = tmp3 * tmp4;
	real tmp7 = n_l[1] * tmp1;
	real tmp9 = X[1] * tmp7;
	real tmp11 = tmp9 * tmp3;
	real tmp13 = n_l[2] * tmp1;
	real tmp15 = X[2] * tmp13;
	real tmp17 = tmp15 * tmp3;
	real tmp20 = X[6] * tmp1;
	real tmp22 = tmp20 * tmp3;
	real tmp26 = X[3] * tmp2;
	real tmp27 = tmp26 * tmp3;
	real tmp31 = X[4] * tmp7;
	real tmp33 = tmp31 * tmp3;
	real tmp37 = X[5] * tmp13;
	real tmp39 = tmp37 * tmp3;
	real tmp42 = X[7] * tmp1;
	real tmp44 = tmp42 * tmp3;
	real tmp45 = m * n2_l[2];
	real tmp47 = X[5] * tmp45;
	real tmp48 = m * n2_l[1];
	real tmp50 = X[4] * tmp48;
	real tmp51 = m * n2_l[0];
	real tmp53 = X[3] * tmp51;
	real tmp54 = e * n3_l[2];
	real tmp56 = X[2] * tmp54;
	real tmp57 = e * n3_l[1];
	real tmp59 = X[1] * tmp57;
	real tmp60 = e * n3_l[0];
	real tmp62 = X[0] * tmp60;
	real tmp63 = tmp59 * tmp3;
	real tmp64 = tmp62 * tmp3;
	real tmp65 = tmp56 * tmp3;
	real tmp67 = tmp53 * tmp3;
	real tmp69 = tmp50 * tmp3;
	real tmp71 = tmp47 * tmp3;
	real tmp73 = e * n2_l[0];
	real tmp75 = X[0] * tmp73;
	real tmp76 = tmp75 * tmp3;
	real tmp77 = e * n2_l[1];
	real tmp79 = X[1] * tmp77;
	real tmp81 = tmp79 * tmp3;
	real tmp82 = e * n2_l[2];
	real tmp84 = X[2] * tmp82;
	real tmp86 = tmp84 * tmp3;
	real tmp87 = m * n3_l[0];
	real tmp89 = X[3] * tmp87;
	real tmp90 = m * n3_l[1];
	real tmp92 = X[4] * tmp90;
	real tmp93 = m * n3_l[2];
	real tmp95 = X[5] * tmp93;
	real tmp96 = tmp92 * tmp3;
	real tmp97 = tmp95 * tmp3;
	real tmp98 = tmp89 * tmp3;
	Y[0] = tmp22 + -tmp5 - tmp11 - tmp17;
	Y[1] = tmp44 + -tmp27 - tmp33 - tmp39;
	Y[2] = tmp71 + tmp69 + tmp67 + tmp65 + tmp63 + tmp64;
	Y[3] = tmp98 + tmp96 + tmp97 + -tmp76 - tmp81 - tmp86;
	Y[4] = tmp67 + tmp69 + tmp71 + -tmp64 - tmp63 - tmp65;
	Y[5] = tmp97 + tmp96 + tmp98 + tmp86 + tmp81 + tmp76;
	Y[6] = tmp22 + tmp17 + tmp11 + tmp5;
	Y[7] = tmp44 + tmp39 + tmp33 + tmp27;
}

void oracle_glm_right(const oracle_params *p, const double *eig, const double *X, int s, double *Y) {
	/* glm-maxwell.cl:257-298 */
	(void)p;
	real n_l[3] = {nl(1, 1, s), nl(1, 2, s), nl(1, 3, s)};
	real n2_l[3] = {nl(2, 1, s), nl(2, 2, s), nl(2, 3, s)};
	real n3_l[3] = {nl(3, 1, s), nl(3, 2, s), nl(3, 3, s)};
	real e = eig[0], m = eig[1];
	real tmp1 = e * n_l[0];
	real tmp3 = 1. / e;
	real tmp4 = tmp3 * n3_l[0];
	real tmp6 = n2_l[0] * tmp3;
	real tmp22 = e * n_l[1];
	real tmp25 = n3_l[1] * tmp3;
	real tmp27 = n2_l[1] * tmp3;
	real tmp43 = e * n_l[2];
	real tmp46 = n3_l[2] * tmp3;
	real tmp48 = n2_l[2] * tmp3;
	real tmp66 = 1. / m;
	real tmp67 = tmp66 * n2_l[0];
	real tmp69 = n3_l[0] * tmp66;
	real tmp88 = n2_l[1] * tmp66;
	real tmp90 = n3_l[1] * tmp66;
	real tmp109 = n2_l[2] * tmp66;
	real tmp111 = n3_l[2] * tmp66;
	Y[0] = X[6] * tmp1 + X[5] * tmp6 + X[2] * tmp4 - X[3] * tmp6 - X[4] * tmp4 + -X[0] * tmp1;
	Y[1] = X[6] * tmp22 + X[5] * tmp27 + X[2] * tmp25 - X[3] * tmp27 - X[4] * tmp25 + -X[0] * tmp22;
	Y[2] = X[6] * tmp43 + X[5] * tmp48 + X[2] * tmp46 - X[3] * tmp48 - X[4] * tmp46 + -X[0] * tmp43;
	Y[3] = X[7] * tmp1 + X[5] * tmp69 + X[4] * tmp67 + X[3] * tmp69 + X[2] * tmp67 + -X[1] * tmp1;
	Y[4] = X[7] * tmp22 + X[5] * tmp90 + X[4] * tmp88 + X[3] * tmp90 + X[2] * tmp88 + -X[1] * tmp22;
	Y[5] = X[7] * tmp43 + X[5] * tmp111 + X[4] * tmp109 + X[3] * tmp111 + X[2] * tmp109 + -X[1] * tmp43;
	Y[6] = X[6] * e + X[0] * e;
	Y[7] = X[7] * e + X[1] * e;
	for (int i = 8; i < 12; ++i) Y[i] = 0;
}

void oracle_glm_flux_from_cons(const oracle_params *p, const double *U, int s, double *F) {
	/* glm-maxwell.cl:99-130, unit_m_per_s = 1 */
	real a2 = U[10] * U[10], b2 = U[11] * U[11];
	real E[3] = {U[0] * a2, U[1] * a2, U[2] * a2};
	real H[3] = {U[3] * b2, U[4] * b2, U[5] * b2};
	real nx = nl(1, 1, s), ny = nl(1, 2, s), nz = nl(1, 3, s);
	real phi = U[6], psi = U[7], cphi = p->div_phi, cpsi = p->div_psi;
	F[0] = H[1] * nz - H[2] * ny + nx * phi * cphi;
	F[3] = E[2] * ny - E[1] * nz + nx * psi * cpsi;
	F[1] = H[2] * nx - H[0] * nz + ny * phi * cphi;
	F[4] = E[0] * nz - E[2] * nx + ny * psi * cpsi;
	F[2] = H[0] * ny - H[1] * nx + nz * phi * cphi;
	F[5] = E[1] * nx - E[0] * ny + nz * psi * cpsi;
	F[6] = U[s] * cphi;
	F[7] = U[3 + s] * cpsi;
	F[8] = F[9] = F[10] = F[11] = 0;
}

void oracle_glm_waves(const oracle_params *p, const double *eig, int s, double *l) {
	/* glm-maxwell.lua:181-214 */
	(void)s;
	real v = eig[0] * eig[1] * (1. / 1.);
	l[0] = -p->div_phi; l[1] = -p->div_psi;
	l[2] = -v; l[3] = -v; l[4] = v; l[5] = v;
	l[6] = p->div_phi; l[7] = p->div_psi;
}

static real glm_dt_cell(const oracle_params *p, const real *U, int dim, const real *dx) {
	/* calcDT.cl:38-73 with glm-maxwell.lua:216-245 */
	real dt = INFINITY;
	for (int s = 0; s < dim; ++s) {
		if (dx[s] > 0) {
			real v_p_abs = U[10] * U[11] * (1. / 1.);
			real wmax = fmax(fmax(p->div_psi, p->div_phi), v_p_abs);
			real lmin = wmax, lmax = -wmax;
			real absLambdaMax = fmax(fabs(lmin), fabs(lmax));
			absLambdaMax = fmax(1e-9, absLambdaMax);
			dt = fmin(dt, dx[s] / absLambdaMax);
		}
	}
	return dt;
}

/* ----------------------------------------------------------- shallow water */
/* hydro/eqn/shallow-water.lua / .cl with depthInCell = false (the depth in
 * cons_t, not integrated) and extraTermInFlux = false, Cartesian normals:
 * cons 0 h, 1-3 m, 4 depth; prim 0 h, 1-3 v, 4 depth; eigen 0 h, 1-3 v, 4 C,
 * 5 depth (shallow-water.lua:293-306). */
enum { SW_H = 0, SW_C = 4, SW_DEPTH = 5 };

void oracle_sw_prim_from_cons(const oracle_params *p, const double *U, double *W) {
	/* shallow-water.cl:3-15 */
	W[0] = U[0];
	real inv = 1. / U[0];
	for (int i = 0; i < 3; ++i) W[1 + i] = U[1 + i] * inv;
	W[4] = U[4];
}

void oracle_sw_eigen_for_interface(const oracle_params *p, const double *UL, const double *UR, double *e) {
	/* shallow-water.cl:228-265: Roe velocity with sqrt(h) weights, the
	 * arithmetic mean height */
	real WL[5], WR[5];
	oracle_sw_prim_from_cons(p, UL, WL);
	oracle_sw_prim_from_cons(p, UR, WR);
	real sqrtHL = sqrt(WL[0]), sqrtHR = sqrt(WR[0]);
	real invDenom = 1. / (sqrtHL + sqrtHR);
	e[SW_H] = .5 * (UL[0] + UR[0]);
	for (int i = 0; i < 3; ++i) e[1 + i] = WL[1 + i] * (sqrtHL * invDenom) + WR[1 + i] * (sqrtHR * invDenom);
	e[SW_DEPTH] = .5 * (UR[4] + UL[4]);
	real CSq = p->gravity * e[SW_H];
	e[SW_C] = sqrt(CSq);
}

void oracle_sw_left(const oracle_params *p, const double *e, const double *X, int s, double *Y) {
	/* shallow-water.cl:286-318 (nLen = 1; nL, n2L, n3L the Cartesian l1..l3 rows) */
	real nLen = 1.;
	real nLx = nl(1, 1, s), nLy = nl(1, 2, s), nLz = nl(1, 3, s);
	real n2Lx = nl(2, 1, s), n2Ly = nl(2, 2, s), n2Lz = nl(2, 3, s);
	real n3Lx = nl(3, 1, s), n3Ly = nl(3, 2, s), n3Lz = nl(3, 3, s);
	real v_n = e[1 + s];
	real invDenom = 1. / (2. * nLen * e[SW_H] * e[SW_C]);
	real tmp1 = e[SW_C] * X[0];
	real tmp2 = nLen * tmp1;
	real tmp3 = v_n * X[0];
	real tmp5 = -tmp3;
	real tmp6 = nLz * X[3];
	real tmp8 = nLy * X[2];
	real tmp10 = nLx * X[1];
	real tmp13 = e[1] * X[0];
	real tmp15 = e[2] * X[0];
	real tmp17 = e[3] * X[0];
	real tmp29 = nLen * e[SW_C];
	real tmp31 = invDenom * tmp29;
	Y[0] = invDenom * (tmp10 + tmp8 + tmp6 + tmp5 + -tmp2);
	Y[1] = 2. * tmp31 * (n2Lx * X[1] + n2Ly * X[2] + n2Lz * X[3] + -n2Lx * tmp13 + -n2Lz * tmp17 + -n2Ly * tmp15);
	Y[2] = 2. * tmp31 * (n3Lx * X[1] + n3Ly * X[2] + n3Lz * X[3] + -n3Lx * tmp13 + -n3Lz * tmp17 + -n3Ly * tmp15);
	Y[3] = invDenom * (tmp10 + tmp8 + tmp6 + tmp2 + tmp5);
}

void oracle_sw_right(const oracle_params *p, const double *e, const double *X, int s, double *Y) {
	/* shallow-water.cl:322-362 (n2Len = n3Len = nLen = 1) */
	real nU[3] = {nl(1, 1, s), nl(1, 2, s), nl(1, 3, s)};
	real n2U[3] = {nl(2, 1, s), nl(2, 2, s), nl(2, 3, s)};
	real n3U[3] = {nl(3, 1, s), nl(3, 2, s), nl(3, 3, s)};
	real nLen = 1., n2Len = 1., n3Len = 1.;
	real h = e[SW_H], C = e[SW_C];
	real invDenom = 1. / C;
	real tmp5 = n3Len * n3Len;
	real tmp6 = n2Len * n2Len;
	real tmp7 = tmp5 * X[0];
	real tmp8 = C * C;
	real tmp9 = tmp6 * tmp7;
	real tmp10 = tmp8 * tmp9;
	real tmp13 = X[3] * tmp5;
	real tmp15 = tmp13 * tmp6;
	real tmp16 = tmp15 * tmp8;
	real tmp17 = C * X[0];
	real tmp19 = nLen * tmp17;
	real tmp21 = tmp19 * tmp5;
	real tmp22 = tmp21 * tmp6;
	real tmp24 = C * X[1];
	real tmp26 = nLen * tmp24;
	real tmp27 = tmp26 * tmp5;
	real tmp28 = C * X[2];
	real tmp30 = nLen * tmp28;
	real tmp31 = tmp30 * tmp6;
	real tmp32 = C * X[3];
	real tmp34 = nLen * tmp32;
	real tmp36 = tmp34 * tmp5;
	real tmp37 = tmp36 * tmp6;
	real tmp51 = nLen * tmp5;
	real tmp53 = tmp51 * tmp6;
	Y[0] = invDenom * (X[3] + -X[0]) * h * C;
	for (int i = 0; i < 3; ++i)
		Y[1 + i] = invDenom * (h * (nU[i] * tmp10 + nU[i] * tmp16 + -e[1 + i] * tmp22 + n2U[i] * tmp27
			+ e[1 + i] * tmp37 + n3U[i] * tmp31)) / tmp53;
}

void oracle_sw_flux_from_cons(const oracle_params *p, const double *U, int s, double *F) {
	/* shallow-water.cl:186-206 (extraTermInFlux = false) */
	real m_n = U[1 + s];
	real v_n = m_n / U[0];
	real hP = .5 * p->gravity * U[0] * (U[0]);
	F[0] = m_n;
	for (int i = 0; i < 3; ++i) F[1 + i] = U[1 + i] * v_n + nl(1, 1 + i, s) * hP;
	F[4] = 0.;
}

void oracle_sw_waves(const oracle_params *p, const double *e, int s, double *l) {
	/* eigenWaveCodePrefix / eigenWaveCode (shallow-water.lua:247-275) */
	real CSq = p->gravity * e[SW_H];
	real C = sqrt(CSq);
	real C_nLen = C * 1.;
	real v_n = e[1 + s];
	if (C <= 0.) { C_nLen = 0.; v_n = 0.; }
	l[0] = v_n - C_nLen;
	l[1] = v_n;
	l[2] = v_n;
	l[3] = v_n + C_nLen;
}

double oracle_sw_dt_cell(const oracle_params *p, const double *U, int dim, const double *dx) {
	/* calcDT.cl:38-73 with consWaveCodeMinMaxAllSides (shallow-water.lua:318-354) */
	real CSq = p->gravity * U[0];
	real C = sqrt(CSq);
	real W[5];
	oracle_sw_prim_from_cons(p, U, W);
	real dt = INFINITY;
	for (int s = 0; s < dim; ++s) {
		if (dx[s] > 0) {
			real C_nLen = C * 1.;
			real v_n = W[1 + s];
			if (C <= 0.) { C_nLen = 0.; v_n = 0.; }
			real lmin = v_n - C_nLen, lmax = v_n + C_nLen;
			real absLambdaMax = fmax(fabs(lmin), fabs(lmax));
			absLambdaMax = fmax(1e-9, absLambdaMax);
			dt = fmin(dt, dx[s] / absLambdaMax);
		}
	}
	return dt;
}

/* ------------------------------------------------------------------- Roe */

static void eqn_eigen(const oracle_params *p, int eqn, const real *UL, const real *UR, real *e) {
	if (eqn == HYDRO_EQN_EULER) oracle_euler_eigen_for_interface(p, UL, UR, e);
	else if (eqn == HYDRO_EQN_SHALLOW_WATER) oracle_sw_eigen_for_interface(p, UL, UR, e);
	else oracle_glm_eigen_for_interface(p, UL, UR, e);
}
static void eqn_left(const oracle_params *p, int eqn, const real *e, const real *X, int s, real *Y) {
	if (eqn == HYDRO_EQN_EULER) oracle_euler_left(p, e, X, s, Y);
	else if (eqn == HYDRO_EQN_SHALLOW_WATER) oracle_sw_left(p, e, X, s, Y);
	else oracle_glm_left(p, e, X, s, Y);
}
static void eqn_right(const oracle_params *p, int eqn, const real *e, const real *X, int s, real *Y) {
	if (eqn == HYDRO_EQN_EULER) oracle_euler_right(p, e, X, s, Y);
	else if (eqn == HYDRO_EQN_SHALLOW_WATER) oracle_sw_right(p, e, X, s, Y);
	else oracle_glm_right(p, e, X, s, Y);
}
static void eqn_flux(const oracle_params *p, int eqn, const real *U, int s, real *F) {
	if (eqn == HYDRO_EQN_EULER) oracle_euler_flux_from_cons(p, U, s, F);
	else if (eqn == HYDRO_EQN_SHALLOW_WATER) oracle_sw_flux_from_cons(p, U, s, F);
	else oracle_glm_flux_from_cons(p, U, s, F);
}
static void eqn_waves(const oracle_params *p, int eqn, const real *e, int s, real *l) {
	if (eqn == HYDRO_EQN_EULER) oracle_euler_waves(p, e, s, l);
	else if (eqn == HYDRO_EQN_SHALLOW_WATER) oracle_sw_waves(p, e, s, l);
	else oracle_glm_waves(p, e, s, l);
}
static int eqn_S(int eqn) { return eqn == HYDRO_EQN_EULER ? 6 : eqn == HYDRO_EQN_SHALLOW_WATER ? 5 : 12; }
static int eqn_q(int eqn) { return eqn == HYDRO_EQN_EULER ? 5 : eqn == HYDRO_EQN_SHALLOW_WATER ? 4 : 8; }

void oracle_roe_flux(const oracle_params *p, int eqn, int flux_limiter, double dt_dx,
	const double *UL, const double *UR, const double *const *lim, int s, double *F)
{
	/* roe.cl:17-163 with roeUseFluxFromCons = true (eqn.lua:46, glm-maxwell.lua:25) */
	int S = eqn_S(eqn), q = eqn_q(eqn), nw = q;
	int useLim = flux_limiter != HYDRO_LIMITER_DONOR_CELL && lim;
	real eig[10], eigL[10], eigR[10];
	eqn_eigen(p, eqn, UL, UR, eig);
	real lambdas[8];
	eqn_waves(p, eqn, eig, s, lambdas);
	real dU[12], dUL[12], dUR[12];
	for (int j = 0; j < S; ++j) {
		dU[j] = UR[j] - UL[j];
		if (useLim) {
			dUL[j] = lim[1][j] - lim[0][j];
			dUR[j] = lim[3][j] - lim[2][j];
		}
	}
	real dW[8], dWL[8], dWR[8], fluxEig[8];
	eqn_left(p, eqn, eig, dU, s, dW);
	if (useLim) {
		eqn_eigen(p, eqn, lim[0], lim[1], eigL);
		eqn_eigen(p, eqn, lim[2], lim[3], eigR);
		eqn_left(p, eqn, eigL, dUL, s, dWL);
		eqn_left(p, eqn, eigR, dUR, s, dWR);
	}
	for (int j = 0; j < nw; ++j) {
		real lambda = lambdas[j];
		if (p->roe_fix && eqn == HYDRO_EQN_EULER) {
			/* roe.cl:83-106 (OpenCL sign: 0 at 0) */
			real dws = p->roe_fix_factor * eig[E_CS];
			real sgn = lambda > 0 ? 1. : lambda < 0 ? -1. : lambda;
			if (fabs(lambda) < dws) lambda = sgn * .5 * (lambda * lambda / dws + dws);
		}
		fluxEig[j] = 0.;
		real sgnLambda = lambda >= 0 ? 1 : -1;
		if (useLim) {
			real rEig;
			if (dW[j] == 0) rEig = 0;
			else rEig = lambda >= 0 ? dWL[j] / dW[j] : dWR[j] / dW[j];
			real phi = oracle_limiter(flux_limiter, rEig);
			fluxEig[j] -= .5 * lambda * dW[j] * (sgnLambda + phi * (lambda * dt_dx - sgnLambda));
		} else {
			fluxEig[j] -= .5 * lambda * dW[j] * (sgnLambda);
		}
	}
	eqn_right(p, eqn, eig, fluxEig, s, F);
	real FL[12], FR[12];
	eqn_flux(p, eqn, UL, s, FL);
	eqn_flux(p, eqn, UR, s, FR);
	for (int j = 0; j < q; ++j) F[j] += .5 * (FL[j] + FR[j]);
}

/* ------------------------------------------------------------ HLL, Rusanov */

static real euler_cs_from_cons(const oracle_params *p, const real *U) {
	/* calc_Cs_fromCons (euler.cl:96-112) */
	real P = euler_calc_P(p, U);
	if (P <= p->P_min) return 0.;
	if (U[0] < p->rho_min) return INFINITY;
	return sqrt(p->gamma * P / U[0]);
}

static void euler_cons_wave_min_max(const oracle_params *p, const real *U, int s, real *lmin, real *lmax) {
	/* consWaveCodeMinMax (eqn.lua:1134-1146) with euler.lua:329-337:
	 * v_n = m_n / rho (0 below rhoMin), v_n -+ Cs */
	real Cs = euler_cs_from_cons(p, U);
	real v_n = U[0] < p->rho_min ? 0. : U[1 + s] / U[0];
	*lmin = v_n - Cs;
	*lmax = v_n + Cs;
}

void oracle_hll_flux(const oracle_params *p, const double *UL, const double *UR, int s, double *F) {
	/* hll.cl:13-73, 'Davis direct bounded' (hll.lua:10); Euler */
	real eig[10];
	oracle_euler_eigen_for_interface(p, UL, UR, eig);
	real lambdas[5];
	oracle_euler_waves(p, eig, s, lambdas);
	real lambdaIntMin = lambdas[0], lambdaIntMax = lambdas[4];
	real lambdaLMin, lambdaLMax, lambdaRMin, lambdaRMax;
	euler_cons_wave_min_max(p, UL, s, &lambdaLMin, &lambdaLMax);
	euler_cons_wave_min_max(p, UR, s, &lambdaRMin, &lambdaRMax);
	real sL = fmin(lambdaLMin, lambdaIntMin);
	real sR = fmax(lambdaRMax, lambdaIntMax);
	real FL[6], FR[6];
	oracle_euler_flux_from_cons(p, UL, s, FL);
	oracle_euler_flux_from_cons(p, UR, s, FR);
	for (int j = 0; j < 5; ++j) {
		if (0 <= sL) F[j] = FL[j];
		else if (sR <= 0) F[j] = FR[j];
		else
			/* also the NaN case, where the reference leaves the flux buffer's
			 * earlier contents */
			F[j] = (sR * FL[j] - sL * FR[j] + sL * sR * (UR[j] - UL[j])) / (sR - sL);
	}
	F[5] = 0;
}

void oracle_hllc_flux(const oracle_params *p, const double *UL, const double *UR, int s, double *F) {
	/* euler-hllc.cl:13-240: hllcMethod 2 (euler-hllc.lua:18) with the 'Davis
	 * direct bounded' estimates it inherits from hll.lua:10. Momenta in the
	 * normal basis (n, t1, t2) = (e_s, e_(s+1)%3, e_(s+2)%3) (coord.lua:2393-2408) */
	real WL[6], WR[6];
	oracle_euler_prim_from_cons(p, UL, WL);
	oracle_euler_prim_from_cons(p, UR, WR);
	real eig[10];
	oracle_euler_eigen_for_interface(p, UL, UR, eig);
	real lambdas[5];
	oracle_euler_waves(p, eig, s, lambdas);
	real lambdaIntMin = lambdas[0], lambdaIntMax = lambdas[4];
	real lambdaLMin, lambdaLMax, lambdaRMin, lambdaRMax;
	euler_cons_wave_min_max(p, UL, s, &lambdaLMin, &lambdaLMax);
	euler_cons_wave_min_max(p, UR, s, &lambdaRMin, &lambdaRMax);
	real sL = fmin(lambdaLMin, lambdaIntMin);
	real sR = fmax(lambdaRMax, lambdaIntMax);
	int a = s, b = (s + 1) % 3, c = (s + 2) % 3;
	real vnL[3] = {WL[1 + a], WL[1 + b], WL[1 + c]};
	real vnR[3] = {WR[1 + a], WR[1 + b], WR[1 + c]};
	real sStar = (WR[0] * vnR[0] * (sR - vnR[0]) - WL[0] * vnL[0] * (sL - vnL[0]) + WL[4] - WR[4])
		/ (WR[0] * (sR - vnR[0]) - WL[0] * (sL - vnL[0]));
	real FL[6], FR[6];
	oracle_euler_flux_from_cons(p, UL, s, FL);
	oracle_euler_flux_from_cons(p, UR, s, FR);
	real fn[3];
	if (0 <= sL) {
		for (int j = 0; j < 5; ++j) F[j] = FL[j];
	} else if (sL <= 0. && 0. <= sStar) {
		real PLR = .5 * (WL[4] + WR[4] + WL[0] * (sL - vnL[0]) * (sStar - vnL[0])
			+ WR[0] * (sR - vnR[0]) * (sStar - vnR[0]));
		F[0] = (sL * UL[0] - FL[0]) * sStar / (sL - sStar);
		real Umn[3] = {UL[1 + a], UL[1 + b], UL[1 + c]}, Fmn[3] = {FL[1 + a], FL[1 + b], FL[1 + c]};
		fn[0] = ((sL * Umn[0] - Fmn[0]) * sStar + sL * PLR) / (sL - sStar);
		fn[1] = sStar * (sL * Umn[1] - Fmn[1]) / (sL - sStar);
		fn[2] = sStar * (sL * Umn[2] - Fmn[2]) / (sL - sStar);
		F[1 + a] = fn[0]; F[1 + b] = fn[1]; F[1 + c] = fn[2];
		F[4] = (sStar * (sL * UL[4] - FL[4]) + sL * PLR * sStar) / (sL - sStar);
	} else if (sStar <= 0. && 0. <= sR) {
		real PLR = .5 * (WL[4] + WR[4] + WL[0] * (sL - vnL[0]) * (sStar - vnL[0])
			+ WR[0] * (sR - vnR[0]) * (sStar - vnR[0]));
		F[0] = sStar * (sR * UR[0] - FR[0]) / (sR - sStar);
		real Umn[3] = {UR[1 + a], UR[1 + b], UR[1 + c]}, Fmn[3] = {FR[1 + a], FR[1 + b], FR[1 + c]};
		fn[0] = (sStar * (sR * Umn[0] - Fmn[0]) + sR * PLR) / (sR - sStar);
		fn[1] = sStar * (sR * Umn[1] - Fmn[1]) / (sR - sStar);
		fn[2] = sStar * (sR * Umn[2] - Fmn[2]) / (sR - sStar);
		F[1 + a] = fn[0]; F[1 + b] = fn[1]; F[1 + c] = fn[2];
		F[4] = (sStar * (sR * UR[4] - FR[4]) + sR * PLR * sStar) / (sR - sStar);
	} else if (sR <= 0) {
		for (int j = 0; j < 5; ++j) F[j] = FR[j];
	} else {
		/* sL <= 0 <= sR with sStar not a number: HLL (also the NaN case,
		 * where the reference leaves the buffer's earlier contents) */
		for (int j = 0; j < 5; ++j)
			F[j] = (sR * FL[j] - sL * FR[j] + sL * sR * (UR[j] - UL[j])) / (sR - sL);
	}
	F[5] = 0;
}

void oracle_euler_eigen_for_cell(const oracle_params *p, const double *U, double *e) {
	/* eigen_forCell (euler.cl:315-341): the cell's own state, no averaging */
	real W[6];
	oracle_euler_prim_from_cons(p, U, W);
	real vSq = dot3(W + 1, W + 1);
	real eKin = .5 * vSq;
	real hTotal = (W[4] + U[4]) / W[0];   /* calc_hTotal = calc_HTotal / rho */
	real CsSq = (p->gamma - 1.) * (hTotal - eKin);
	e[E_RHO] = W[0];
	e[1] = W[1]; e[2] = W[2]; e[3] = W[3];
	e[E_HT] = hTotal;
	e[E_CS] = sqrt(CsSq);
	e[E_VSQ] = vSq;
	e[E_VL] = W[1]; e[E_VL + 1] = W[2]; e[E_VL + 2] = W[3];
}

void oracle_marquina_flux(const oracle_params *p, const double *UL, const double *UR, int s, double *F) {
	/* marquina.cl:5-93 */
	real FL[6], FR[6];
	oracle_euler_flux_from_cons(p, UL, s, FL);
	oracle_euler_flux_from_cons(p, UR, s, FR);
	real eL[10], eR[10];
	oracle_euler_eigen_for_cell(p, UL, eL);
	oracle_euler_eigen_for_cell(p, UR, eR);
	real omegaL[5], omegaR[5], phiL[5], phiR[5];
	oracle_euler_left(p, eL, UL, s, omegaL);
	oracle_euler_left(p, eR, UR, s, omegaR);
	oracle_euler_left(p, eL, FL, s, phiL);
	oracle_euler_left(p, eR, FR, s, phiR);
	real lL[5], lR[5];
	oracle_euler_waves(p, eL, s, lL);
	oracle_euler_waves(p, eR, s, lR);
	real pp[5], pm[5];
	for (int k = 0; k < 5; ++k) {
		if (lL[k] * lR[k] > 0) {
			if (lL[k] > 0) {
				pp[k] = phiL[k];
				pm[k] = 0;
			} else {
				pp[k] = 0;
				pm[k] = phiR[k];
			}
		} else {
			real alpha = fmax(fabs(lL[k]), fabs(lR[k]));
			pp[k] = (phiL[k] + alpha * omegaL[k]) * .5;
			pm[k] = (phiR[k] - alpha * omegaR[k]) * .5;
		}
	}
	real YL[6], YR[6];
	oracle_euler_right(p, eL, pp, s, YL);
	oracle_euler_right(p, eR, pm, s, YR);
	for (int k = 0; k < 5; ++k) F[k] = YL[k] + YR[k];
	F[5] = 0;
}

void oracle_rusanov_flux(const oracle_params *p, const double *UL, const double *UR, int s, double *F) {
	/* rusanov.cl:13-39: lambdaMax is assigned for the left state, then for the
	 * right state, so the right state's value is used */
	real lambdaMax, lmin, lmax;
	euler_cons_wave_min_max(p, UL, s, &lmin, &lmax);
	lambdaMax = fmax(fabs(lmin), fabs(lmax));
	euler_cons_wave_min_max(p, UR, s, &lmin, &lmax);
	lambdaMax = fmax(fabs(lmin), fabs(lmax));
	real FL[6], FR[6];
	oracle_euler_flux_from_cons(p, UL, s, FL);
	oracle_euler_flux_from_cons(p, UR, s, FR);
	for (int j = 0; j < 5; ++j) F[j] = .5 * (FL[j] + FR[j] - lambdaMax * (UR[j] - UL[j]));
	F[5] = 0;
}

void oracle_plm_prim(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	double *L, double *R)
{
	/* plm.cl:191-254 ('plm prim', Euler): slopes of the primitive variables,
	 * r = dWL / dWR always, back to cons at the edges */
	real W[6], WL[6], WR[6], nWL[6], nWR[6];
	oracle_euler_prim_from_cons(p, U, W);
	oracle_euler_prim_from_cons(p, UL, WL);
	oracle_euler_prim_from_cons(p, UR, WR);
	for (int j = 0; j < 6; ++j) nWL[j] = nWR[j] = W[j];
	for (int j = 0; j < 5; ++j) {
		real dWR = WR[j] - W[j];
		real dWL = W[j] - WL[j];
		real r = dWR == 0 ? 0 : (dWL / dWR);
		real phi = oracle_limiter(lim, r);
		real sigma = phi * dWR;
		nWL[j] -= .5 * sigma;
		nWR[j] += .5 * sigma;
	}
	oracle_euler_cons_from_prim(p, nWL, L);
	oracle_euler_cons_from_prim(p, nWR, R);
}

void oracle_plm_eig(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	int s, double dt_dx, double *L, double *R)
{
	/* plm.cl:256-432 ('plm eig', the form its '#if 1' enables), Euler,
	 * Cartesian: the slopes dUL = U - UL, dUR = UR - U in the characteristic
	 * variables of the cell's own eigensystem (eigen_forCell), limited with
	 * r = dL / dR (0 when dR == 0) to sigma = phi(r) dR; a wave with speed >= 0
	 * keeps sigma on the left edge only, a wave with speed <= 0 on the right
	 * edge only; back to cons, UHalfL = U - sL / 2, UHalfR = U + sR / 2; both
	 * edges then add (dt/dx)/2 (F(UHalfR) - F(UHalfL)). The center slope's
	 * transform (dUCEig) is computed there but used by no output. */
	real dUL[6], dUR[6];
	for (int j = 0; j < 5; ++j) {
		dUL[j] = U[j] - UL[j];
		dUR[j] = UR[j] - U[j];
	}
	dUL[5] = dUR[5] = 0;
	real eig[10];
	oracle_euler_eigen_for_cell(p, U, eig);
	real dULEig[5], dUREig[5];
	oracle_euler_left(p, eig, dUL, s, dULEig);
	oracle_euler_left(p, eig, dUR, s, dUREig);
	real wave[5];
	oracle_euler_waves(p, eig, s, wave);
	for (int j = 0; j < 5; ++j) {
		real rEig = dUREig[j] == 0 ? 0 : (dULEig[j] / dUREig[j]);
		real phi = oracle_limiter(lim, rEig);
		real sigma = phi * dUREig[j];
		dULEig[j] = sigma;
		dUREig[j] = sigma;
		if (wave[j] >= 0) dUREig[j] = 0;
		if (wave[j] <= 0) dULEig[j] = 0;
	}
	real sL[6], sR[6];
	oracle_euler_right(p, eig, dULEig, s, sL);
	oracle_euler_right(p, eig, dUREig, s, sR);
	real UHalfL[6], UHalfR[6];
	for (int j = 0; j < 6; ++j) UHalfL[j] = UHalfR[j] = U[j];
	for (int j = 0; j < 5; ++j) {
		UHalfL[j] -= .5 * sL[j];
		UHalfR[j] += .5 * sR[j];
	}
	real FHalfL[6], FHalfR[6];
	oracle_euler_flux_from_cons(p, UHalfL, s, FHalfL);
	oracle_euler_flux_from_cons(p, UHalfR, s, FHalfR);
	for (int j = 0; j < 6; ++j) {
		L[j] = UHalfL[j];
		R[j] = UHalfR[j];
	}
	for (int j = 0; j < 5; ++j) {
		real dF = FHalfR[j] - FHalfL[j];
		L[j] += .5 * dt_dx * dF;
		R[j] += .5 * dt_dx * dF;
	}
}

void oracle_plm_cons_flux(const oracle_params *p, int lim, const double *UL, const double *U, const double *UR,
	int s, double dt_dx, double *L, double *R)
{
	/* plm.cl:95-188 ('plm cons with flux'), Euler, Cartesian: per field
	 * r = dUL / dUR (0 when dUR == 0), sigma = phi(r) dUR, UHalfL = U - sigma/2,
	 * UHalfR = U + sigma/2; both edges then add (dt/dx)/2 (F(UHalfR) - F(UHalfL)) */
	real UHalfL[6], UHalfR[6];
	for (int j = 0; j < 6; ++j) UHalfL[j] = UHalfR[j] = U[j];
	for (int j = 0; j < 5; ++j) {
		real dUL = U[j] - UL[j];
		real dUR = UR[j] - U[j];
		real r = dUR == 0 ? 0 : (dUL / dUR);
		real phi = oracle_limiter(lim, r);
		real sigma = phi * dUR;
		UHalfL[j] -= .5 * sigma;
		UHalfR[j] += .5 * sigma;
	}
	real FHalfL[6], FHalfR[6];
	oracle_euler_flux_from_cons(p, UHalfL, s, FHalfL);
	oracle_euler_flux_from_cons(p, UHalfR, s, FHalfR);
	for (int j = 0; j < 6; ++j) {
		L[j] = UHalfL[j];
		R[j] = UHalfR[j];
	}
	for (int j = 0; j < 5; ++j) {
		real dF = FHalfR[j] - FHalfL[j];
		L[j] += .5 * dt_dx * dF;
		R[j] += .5 * dt_dx * dF;
	}
}

static real ocl_sign(real x) {
	/* OpenCL sign(): 1 for x > 0, -1 for x < 0, the zero itself for +-0, 0 for NaN */
	if (x > 0) return 1.;
	if (x < 0) return -1.;
	return x == x ? x : 0.;
}

void oracle_plm_athena(const oracle_params *p, const double *UL, const double *U, const double *UR,
	int s, double *L, double *R)
{
	/* plm.cl:782-879 ('plm athena'; a second branch of the same name at
	 * plm.cl:950 is never reached), Euler, Cartesian. The differences of the
	 * primitives go through the cons eigensystem of the cell as they are
	 * (the reference's own TODO asks whether it should be dA/dW). OpenCL
	 * min/max on reals are fmin/fmax here (they differ only for NaN and in
	 * the sign of a zero tie). */
	real eig[10];
	oracle_euler_eigen_for_cell(p, U, eig);
	real W[6], WL[6], WR[6];
	oracle_euler_prim_from_cons(p, U, W);
	oracle_euler_prim_from_cons(p, UL, WL);
	oracle_euler_prim_from_cons(p, UR, WR);
	real dWL[6], dWR[6], dWC[6], dWG[6];
	for (int j = 0; j < 5; ++j) {
		dWL[j] = W[j] - WL[j];
		dWR[j] = WR[j] - W[j];
		dWC[j] = .5 * (WR[j] - WL[j]);
		dWG[j] = (dWL[j] * dWR[j]) <= 0. ? 0. : (2. * dWL[j] * dWR[j] / (dWL[j] + dWR[j]));
	}
	dWL[5] = dWR[5] = dWC[5] = dWG[5] = 0.;
	real dal[5], dar[5], dac[5], dag[5];
	oracle_euler_left(p, eig, dWL, s, dal);
	oracle_euler_left(p, eig, dWR, s, dar);
	oracle_euler_left(p, eig, dWC, s, dac);
	oracle_euler_left(p, eig, dWG, s, dag);
	real da[5];
	for (int j = 0; j < 5; ++j) {
		da[j] = 0;
		if (dal[j] * dar[j] > 0) {
			real lim_slope1 = fmin(fabs(dal[j]), fabs(dar[j]));
			real lim_slope2 = fmin(fabs(dac[j]), fabs(dag[j]));
			da[j] = ocl_sign(dac[j]) * fmin(2. * lim_slope1, lim_slope2);
		}
	}
	real dWm[6];
	oracle_euler_right(p, eig, da, s, dWm);
	real Wlv[6], Wrv[6];
	for (int j = 0; j < 5; ++j) {
		Wlv[j] = W[j] - .5 * dWm[j];
		Wrv[j] = W[j] + .5 * dWm[j];
		real C = Wrv[j] + Wlv[j];
		Wlv[j] = fmax(fmin(W[j], WL[j]), Wlv[j]);
		Wlv[j] = fmin(fmax(W[j], WL[j]), Wlv[j]);
		Wrv[j] = C - Wlv[j];
		Wrv[j] = fmax(fmin(W[j], WR[j]), Wrv[j]);
		Wrv[j] = fmin(fmax(W[j], WR[j]), Wrv[j]);
		Wlv[j] = C - Wrv[j];
	}
	Wlv[5] = Wrv[5] = W[5];
	/* the cell's left edge takes Wrv, its right edge Wlv, as stored there */
	oracle_euler_cons_from_prim(p, Wrv, L);
	oracle_euler_cons_from_prim(p, Wlv, R);
}

static void euler_apply_dU_dW(const oracle_params *p, const real *WA, const real *W, real *U) {
	/* euler.cl:179-195, Cartesian (vL = v); prims as (rho, v, P, ePot) */
	U[0] = W[0];
	for (int i = 0; i < 3; ++i) U[1 + i] = WA[1 + i] * W[0] + W[1 + i] * WA[0];
	U[4] = W[0] * .5 * dot3(WA + 1, WA + 1) + WA[0] * dot3(W + 1, WA + 1) + W[4] / (p->gamma - 1.);
	U[5] = W[5];
}

static void euler_apply_dW_dU(const oracle_params *p, const real *WA, const real *U, real *W) {
	/* euler.cl:201-223, Cartesian: the rhoMin test is on U's (the
	 * difference's) density */
	W[0] = U[0];
	if (U[0] < p->rho_min) {
		W[1] = W[2] = W[3] = 0.;
		W[4] = 0.;
	} else {
		for (int i = 0; i < 3; ++i) W[1 + i] = U[1 + i] * (1. / WA[0]) - WA[1 + i] * (U[0] / WA[0]);
		W[4] = (p->gamma - 1.) * (.5 * dot3(WA + 1, WA + 1) * U[0] - dot3(U + 1, WA + 1) + U[4]);
	}
	W[5] = U[5];
}

void oracle_plm_eig_prim(const oracle_params *p, int ref, const double *UL, const double *U, const double *UR,
	int s, double dt_dx, double *L, double *R)
{
	/* plm.cl:536-780 ('plm eig prim', ref = 0; 'plm eig prim ref', ref = 1),
	 * Euler, Cartesian */
	real W[6], WL[6], WR[6];
	oracle_euler_prim_from_cons(p, U, W);
	oracle_euler_prim_from_cons(p, UL, WL);
	oracle_euler_prim_from_cons(p, UR, WR);
	real dWL[6], dWR[6], dWC[6];
	for (int j = 0; j < 5; ++j) {
		dWL[j] = W[j] - WL[j];
		dWR[j] = WR[j] - W[j];
		dWC[j] = .5 * (WR[j] - WL[j]);
	}
	dWL[5] = dWR[5] = dWC[5] = 0.;
	real eig[10];
	oracle_euler_eigen_for_cell(p, U, eig);
	real tmp[6], dWLEig[5], dWREig[5], dWCEig[5];
	euler_apply_dU_dW(p, W, dWL, tmp);
	oracle_euler_left(p, eig, tmp, s, dWLEig);
	euler_apply_dU_dW(p, W, dWR, tmp);
	oracle_euler_left(p, eig, tmp, s, dWREig);
	euler_apply_dU_dW(p, W, dWC, tmp);
	oracle_euler_left(p, eig, tmp, s, dWCEig);
	real dWMEig[5];
	for (int j = 0; j < 5; ++j)
		dWMEig[j] = dWLEig[j] * dWREig[j] < 0. ? 0.
			: ((dWCEig[j] >= 0. ? 1. : -1.) * 2. * fmin(fmin(fabs(dWLEig[j]), fabs(dWREig[j])), fabs(dWCEig[j])));
	real wave[5];
	oracle_euler_waves(p, eig, s, wave);
	real aL[5], aR[5], sL[6], sR[6], W2L[6], W2R[6];
	if (!ref) {
		for (int j = 0; j < 5; ++j) {
			aL[j] = wave[j] < 0 ? 0 : dWMEig[j] * .5 * (1. - wave[j] * dt_dx);
			aR[j] = wave[j] > 0 ? 0 : dWMEig[j] * .5 * (1. + wave[j] * dt_dx);
		}
		tmp[5] = 0;
		oracle_euler_right(p, eig, aL, s, tmp);
		euler_apply_dW_dU(p, W, tmp, sL);
		oracle_euler_right(p, eig, aR, s, tmp);
		euler_apply_dW_dU(p, W, tmp, sR);
		for (int j = 0; j < 5; ++j) {
			W2L[j] = W[j] + sL[j];
			W2R[j] = W[j] - sR[j];
		}
	} else {
		/* eigenWaveCodeMinMax (eqn.lua:1108-1120): the first and last waves */
		real waveMin = wave[0], waveMax = wave[4];
		waveMin = fmin(0., waveMin);
		waveMax = fmax(0., waveMax);
		real dWM[6];
		tmp[5] = 0;
		oracle_euler_right(p, eig, dWMEig, s, tmp);
		euler_apply_dW_dU(p, W, tmp, dWM);
		real WLRef[5], WRRef[5];
		for (int j = 0; j < 5; ++j) {
			WLRef[j] = W[j] + .5 * (1. - dt_dx * waveMax) * dWM[j];
			WRRef[j] = W[j] - .5 * (1. + dt_dx * waveMin) * dWM[j];
		}
		for (int j = 0; j < 5; ++j) {
			aL[j] = wave[j] < 0 ? 0 : (dWMEig[j] * dt_dx * (waveMax - wave[j]));
			aR[j] = wave[j] > 0 ? 0 : (dWMEig[j] * dt_dx * (waveMin - wave[j]));
		}
		oracle_euler_right(p, eig, aL, s, tmp);
		euler_apply_dW_dU(p, W, tmp, sL);
		oracle_euler_right(p, eig, aR, s, tmp);
		euler_apply_dW_dU(p, W, tmp, sR);
		for (int j = 0; j < 5; ++j) {
			W2R[j] = WRRef[j] + .5 * sR[j];
			W2L[j] = WLRef[j] + .5 * sL[j];
		}
	}
	W2L[5] = W2R[5] = W[5];
	oracle_euler_cons_from_prim(p, W2L, L);
	oracle_euler_cons_from_prim(p, W2R, R);
}

void oracle_plm_ppm(int lim, int nint, int S, const double *UL, const double *U, const double *UR,
	double *L, double *R)
{
	/* plm.cl:882-944 ('ppm'): with kappa = 1/3 and phi = phi(dUL / dUR),
	 * uL_{i+1/2} = U + phi/4 ((1 - kappa) dUL + (1 + kappa) dUR) is the right
	 * edge, uR_{i-1/2} = U - phi/4 ((1 - kappa) dUR + (1 + kappa) dUL) the left;
	 * the differences with the second neighbours feed no output */
	real const kappa = 1. / 3.;
	for (int j = 0; j < S; ++j) L[j] = R[j] = U[j];
	for (int j = 0; j < nint; ++j) {
		real dUL = U[j] - UL[j];
		real dUR = UR[j] - U[j];
		real r = dUR == 0 ? 0 : dUL / dUR;
		real phi = oracle_limiter(lim, r);
		R[j] += .25 * phi * ((1. - kappa) * dUL + (1. + kappa) * dUR);
		L[j] -= .25 * phi * ((1. - kappa) * dUR + (1. + kappa) * dUL);
	}
}

void oracle_plm_cons(int lim, int nint, int S, const double *UL, const double *U, const double *UR,
	double *L, double *R)
{
	/* plm.cl:27-91 ('plm cons') */
	for (int j = 0; j < S; ++j) L[j] = R[j] = U[j];
	for (int j = 0; j < nint; ++j) {
		real dUR = UR[j] - U[j];
		real dUL = U[j] - UL[j];
		real sigma;
		real dUC = .5 * (dUR - dUL);
		if (dUC >= 0) {
			real r = dUR == 0 ? 0 : (dUL / dUR);
			sigma = oracle_limiter(lim, r) * dUR;
		} else {
			real r = dUL == 0 ? 0 : (dUR / dUL);
			sigma = oracle_limiter(lim, r) * dUL;
		}
		R[j] += .5 * sigma;
		L[j] -= .5 * sigma;
	}
}

/* ------------------------------------------------------------ integrators */
/* hydro/int/all.lua:10-160, rows are stages, columns k */
typedef struct { int order; double a[4][4], b[4][4]; } rk_table;
static int rk_table_for(int integ, rk_table *t) {
	memset(t, 0, sizeof *t);
	switch (integ) {
	case HYDRO_INT_RK2: t->order = 2;
		t->a[0][0] = 1; t->a[1][0] = 1; t->b[0][0] = .5; t->b[1][1] = 1; return 1;
	case HYDRO_INT_RK2_HEUN: t->order = 2;
		t->a[0][0] = 1; t->a[1][0] = 1; t->b[0][0] = 1; t->b[1][0] = .5; t->b[1][1] = .5; return 1;
	case HYDRO_INT_RK2_RALSTON: t->order = 2;
		t->a[0][0] = 1; t->a[1][0] = 1; t->b[0][0] = 2. / 3.; t->b[1][0] = 1. / 4.; t->b[1][1] = 3. / 4.; return 1;
	case HYDRO_INT_RK3: t->order = 3;
		t->a[0][0] = t->a[1][0] = t->a[2][0] = 1;
		t->b[0][0] = .5; t->b[1][0] = -1; t->b[1][1] = 2;
		t->b[2][0] = 1. / 6.; t->b[2][1] = 2. / 6.; t->b[2][2] = 1. / 6.; return 1;
	case HYDRO_INT_RK4: t->order = 4;
		for (int i = 0; i < 4; ++i) t->a[i][0] = 1;
		t->b[0][0] = .5; t->b[1][1] = .5; t->b[2][2] = 1;
		t->b[3][0] = 1. / 6.; t->b[3][1] = 2. / 6.; t->b[3][2] = 2. / 6.; t->b[3][3] = 1. / 6.; return 1;
	case HYDRO_INT_RK4_3_8THS: t->order = 4;
		for (int i = 0; i < 4; ++i) t->a[i][0] = 1;
		t->b[0][0] = 1. / 3.; t->b[1][0] = -1. / 3.;
		t->b[2][0] = 1; t->b[2][1] = -1; t->b[2][2] = 1;
		t->b[3][0] = 1. / 8.; t->b[3][1] = 3. / 8.; t->b[3][2] = 3. / 8.; t->b[3][3] = 1. / 8.; return 1;
	case HYDRO_INT_RK2_TVD: t->order = 2;
		t->a[0][0] = 1; t->a[1][0] = .5; t->a[1][1] = .5;
		t->b[0][0] = 1; t->b[1][1] = .5; return 1;
	case HYDRO_INT_RK2_NON_TVD: t->order = 2;
		t->a[0][0] = 1; t->a[1][0] = 1;
		t->b[0][0] = -20; t->b[1][0] = 41. / 40.; t->b[1][1] = -1. / 40.; return 1;
	case HYDRO_INT_RK3_TVD: t->order = 3;
		t->a[0][0] = 1; t->a[1][0] = 3. / 4.; t->a[1][1] = 1. / 4.; t->a[2][0] = 1. / 3.; t->a[2][2] = 2. / 3.;
		t->b[0][0] = 1; t->b[1][1] = 1. / 4.; t->b[2][2] = 2. / 3.; return 1;
	case HYDRO_INT_RK4_TVD: t->order = 4;
		t->a[0][0] = 1;
		t->a[1][0] = 649. / 1600.; t->a[1][1] = 951. / 1600.;
		t->a[2][0] = 53989. / 2500000.; t->a[2][1] = 4806213. / 20000000.; t->a[2][2] = 23619. / 32000.;
		t->a[3][0] = 1. / 5.; t->a[3][1] = 6127. / 30000.; t->a[3][2] = 7873. / 30000.; t->a[3][3] = 1. / 3.;
		t->b[0][0] = .5;
		t->b[1][0] = -10890423. / 25193600.; t->b[1][1] = 5000. / 7873;
		t->b[2][0] = -102261. / 5000000.; t->b[2][1] = -5121. / 20000.; t->b[2][2] = 7873. / 10000.;
		t->b[3][0] = 1. / 10.; t->b[3][1] = 1. / 6.; t->b[3][3] = 1. / 6.; return 1;
	case HYDRO_INT_RK4_NON_TVD: t->order = 4;
		for (int i = 0; i < 3; ++i) t->a[i][0] = 1;
		t->a[3][0] = -1. / 3.; t->a[3][1] = 1. / 3.; t->a[3][2] = 2. / 3.; t->a[3][3] = 1. / 3.;
		t->b[0][0] = .5; t->b[1][1] = .5; t->b[2][2] = 1; t->b[3][3] = 1. / 6.; return 1;
	}
	return 0;
}

/* ---------------------------------------------------------------- solver */

struct oracle_solver {
	hydro_cfg cfg;
	oracle_params p;
	int dim, S, q;
	int gs[3];
	long step[3];
	size_t G;
	real dx[3];
	real *U, *ULR, *flux, *deriv, *dtBuf;
	real *UBufs[4], *derivBufs[4];
	rk_table rk;
	int needU[4], needD[4];
	double t;
};

size_t oracle_num_cells(const oracle_solver *s) { return s->G; }
int oracle_num_states(const oracle_solver *s) { return s->S; }
double *oracle_U(oracle_solver *s) { return s->U; }
double oracle_time(const oracle_solver *s) { return s->t; }
void oracle_set_time(oracle_solver *s, double t) { s->t = t; }

static real *zalloc(size_t n) {
	real *p = (real *)calloc(n ? n : 1, sizeof(real));
	return p;
}

oracle_solver *oracle_create(const hydro_cfg *cfg) {
	if (cfg->dim < 1 || cfg->dim > 3) { set_err("dim must be 1..3"); return NULL; }
	if (cfg->eqn != HYDRO_EQN_EULER && cfg->eqn != HYDRO_EQN_GLM_MAXWELL && cfg->eqn != HYDRO_EQN_SHALLOW_WATER) {
		set_err("bad eqn"); return NULL;
	}
	if (cfg->plm && cfg->flux_limiter != HYDRO_LIMITER_DONOR_CELL) {
		set_err("are you sure you want to use flux and slope limiters at the same time?");
		return NULL;
	}
	if (cfg->flux < HYDRO_FLUX_ROE || cfg->flux > HYDRO_FLUX_MARQUINA) {
		set_err("bad flux"); return NULL;
	}
	if (cfg->plm > HYDRO_PLM_CONS && cfg->plm != HYDRO_PLM_PPM && cfg->eqn != HYDRO_EQN_EULER) {
		set_err("the reconstructions other than plm cons and ppm are built for the Euler equations"); return NULL;
	}
	if (cfg->flux != HYDRO_FLUX_ROE && cfg->eqn != HYDRO_EQN_EULER) {
		set_err("HLL, Rusanov, HLLC and Marquina take the Euler equations"); return NULL;
	}
	oracle_solver *s = (oracle_solver *)calloc(1, sizeof *s);
	s->cfg = *cfg;
	s->p.gamma = cfg->gamma; s->p.rho_min = cfg->rho_min; s->p.P_min = cfg->P_min;
	s->p.div_phi = cfg->div_phi_wavespeed; s->p.div_psi = cfg->div_psi_wavespeed;
	s->p.roe_fix = cfg->roe_entropy_fix; s->p.roe_fix_factor = cfg->roe_entropy_fix_factor;
	s->p.gravity = cfg->gravity;
	s->dim = cfg->dim;
	s->S = eqn_S(cfg->eqn);
	s->q = eqn_q(cfg->eqn);
	for (int i = 0; i < 3; ++i) {
		int n = i < s->dim ? cfg->size[i] : 1;
		s->gs[i] = i < s->dim ? n + 2 * NG : 1;
		/* gridsolver.lua:404-411: dx over all three axes, sizeWithoutBorder */
		s->dx[i] = (cfg->maxs[i] - cfg->mins[i]) / (double)n;
	}
	s->step[0] = 1; s->step[1] = s->gs[0]; s->step[2] = (long)s->gs[0] * s->gs[1];
	s->G = (size_t)s->gs[0] * s->gs[1] * s->gs[2];
	size_t GS = s->G * s->S;
	s->U = zalloc(GS);
	s->ULR = zalloc(GS * s->dim * 2);
	s->flux = zalloc(GS * s->dim);
	s->deriv = zalloc(GS);
	s->dtBuf = zalloc(s->G);
	if (cfg->integrator != HYDRO_INT_FORWARD_EULER) {
		if (!rk_table_for(cfg->integrator, &s->rk)) { set_err("unsupported integrator"); oracle_destroy(s); return NULL; }
		int o = s->rk.order;
		for (int i = 0; i < o; ++i) {
			for (int m = i; m < o; ++m) {
				if (s->rk.a[m][i] != 0) s->needU[i] = 1;
				if (s->rk.b[m][i] != 0) s->needD[i] = 1;
			}
			if (s->needU[i]) s->UBufs[i] = zalloc(GS);
			if (s->needD[i]) s->derivBufs[i] = zalloc(GS);
		}
	}
	return s;
}

void oracle_destroy(oracle_solver *s) {
	if (!s) return;
	free(s->U); free(s->ULR); free(s->flux); free(s->deriv); free(s->dtBuf);
	for (int i = 0; i < 4; ++i) { free(s->UBufs[i]); free(s->derivBufs[i]); }
	free(s);
}

/* SETBOUNDS(lhs, rhs) over used axes (gridsolver.lua:277-320) */
#define FOR_CELLS(s, lhs, rhs, ...) do { \
	int lo0 = (lhs), hi0 = (s)->gs[0] - (rhs); \
	int lo1 = (s)->dim > 1 ? (lhs) : 0, hi1 = (s)->dim > 1 ? (s)->gs[1] - (rhs) : 1; \
	int lo2 = (s)->dim > 2 ? (lhs) : 0, hi2 = (s)->dim > 2 ? (s)->gs[2] - (rhs) : 1; \
	_Pragma("omp parallel for collapse(2) schedule(static)") \
	for (int k = lo2; k < hi2; ++k) \
	for (int jj = lo1; jj < hi1; ++jj) \
	for (int i = lo0; i < hi0; ++i) { \
		long index = i + (s)->step[1] * jj + (s)->step[2] * k; \
		int iv[3] = {i, jj, k}; (void)iv; \
		__VA_ARGS__ \
	} \
} while (0)

/* ---------------------------------------------------------------- boundary */
static void bc_apply(oracle_solver *s, real *buf, int side, int minmax, long dst, long src) {
	int bc = s->cfg.boundary[2 * side + minmax];
	int S = s->S;
	real *d = buf + dst * S;
	switch (bc) {
	case HYDRO_BC_PERIODIC:
	case HYDRO_BC_FREEFLOW:
		memcpy(d, buf + src * S, S * sizeof(real));
		break;
	case HYDRO_BC_MIRROR:
		memcpy(d, buf + src * S, S * sizeof(real));
		/* reflectVars.mirror (eqn.lua:366-370): vector component along side */
		if (s->cfg.eqn == HYDRO_EQN_EULER || s->cfg.eqn == HYDRO_EQN_SHALLOW_WATER) {
			d[1 + side] = -1 * d[1 + side];
		} else {
			d[side] = -1 * d[side];
			d[3 + side] = -1 * d[3 + side];
		}
		break;
	case HYDRO_BC_FIXED:
		for (int j = 0; j < S; ++j) d[j] = s->cfg.boundary_fixed[2 * side + minmax][j];
		break;
	}
}

static long bc_src(oracle_solver *s, int side, int minmax, int j) {
	int bc = s->cfg.boundary[2 * side + minmax];
	int gsz = s->gs[side], n = gsz - 2 * NG;
	if (minmax == 0) {
		switch (bc) {
		case HYDRO_BC_PERIODIC: return NG + (j - NG + 2 * n) % n;
		case HYDRO_BC_MIRROR: return 2 * NG - 1 - j;
		case HYDRO_BC_FREEFLOW: return NG;
		}
	} else {
		switch (bc) {
		case HYDRO_BC_PERIODIC: return NG + (NG - 1 - j) % n;
		case HYDRO_BC_MIRROR: return gsz - NG - 1 - j;
		case HYDRO_BC_FREEFLOW: return gsz - NG - 1;
		}
	}
	return 0;
}
static long bc_dst(oracle_solver *s, int side, int minmax, int j) {
	int bc = s->cfg.boundary[2 * side + minmax];
	int gsz = s->gs[side];
	if (minmax == 0) return j;
	if (bc == HYDRO_BC_PERIODIC) return gsz - 1 - j;
	return gsz - NG + j;
}

static void boundary_buf(oracle_solver *s, real *buf) {
	/* gridsolver.lua:1070-1213 codegen, launched x then y then z (1272-1314).
	 * Each pass covers every index of the other axes, ghosts included. */
	for (int side = 0; side < s->dim; ++side) {
		int a = side == 0 ? 1 : 0, b = side == 2 ? 1 : 2;
		int na = s->gs[a], nb = s->gs[b];
		#pragma omp parallel for collapse(2) schedule(static)
		for (int ib = 0; ib < nb; ++ib)
		for (int ia = 0; ia < na; ++ia) {
			long base = ia * s->step[a] + ib * s->step[b];
			for (int j = 0; j < NG; ++j) {
				for (int minmax = 0; minmax < 2; ++minmax) {
					long dst = base + bc_dst(s, side, minmax, j) * s->step[side];
					long src = base + bc_src(s, side, minmax, j) * s->step[side];
					bc_apply(s, buf, side, minmax, dst, src);
				}
			}
		}
	}
}

void oracle_boundary(oracle_solver *s) { boundary_buf(s, s->U); }

double oracle_euler_display(const oracle_params *p, int id, const double *U) {
	/* euler.lua:240-260 with W = primFromCons(U) (euler.cl:140-159) and the
	 * formulas of euler.cl:1-94 */
	real W[6];
	oracle_euler_prim_from_cons(p, U, W);
	real g = p->gamma;
	real vSq = dot3(W + 1, W + 1);
	switch (id) {
	case HYDRO_VAR_P: return W[4];
	case HYDRO_VAR_E_INT_SPEC: return W[4] / (g - 1.) / W[0];
	case HYDRO_VAR_E_KIN_SPEC: return .5 * vSq;
	case HYDRO_VAR_E_TOTAL_SPEC: return U[4] / W[0];
	case HYDRO_VAR_E_INT: return W[4] / (g - 1.);
	case HYDRO_VAR_E_KIN: return W[0] * (.5 * vSq);
	case HYDRO_VAR_S: return W[4] / pow(W[0], g);
	case HYDRO_VAR_H: return W[4] * (g / (g - 1.));
	case HYDRO_VAR_H_SPEC: return W[4] * (g / (g - 1.)) / W[0];
	case HYDRO_VAR_H_TOTAL: return W[4] + U[4];
	case HYDRO_VAR_H_TOTAL_SPEC: return (W[4] + U[4]) / W[0];
	case HYDRO_VAR_CS: return euler_calc_Cs(p, W);
	case HYDRO_VAR_MACH: return sqrt(vSq) / euler_calc_Cs(p, W);
	case HYDRO_VAR_V_X: return W[1];
	case HYDRO_VAR_V_Y: return W[2];
	case HYDRO_VAR_V_Z: return W[3];
	default: return NAN;
	}
}

static real var_value(const oracle_solver *s, long c, int field) {
	if (field >= HYDRO_VAR_P) return oracle_euler_display(&s->p, field, s->U + c * s->S);
	return s->U[c * s->S + field];
}

void oracle_var_stats(const oracle_solver *s, int field, double *out) {
	/* solverbase.lua:2871-2994: reduceMin/reduceMax over the whole buffer;
	 * reduceSum over the buffer with zeroed ghosts, divided by
	 * volumeWithoutBorder; subtractAndSquare on the interior (solverbase.lua:
	 * 1269-1278) for the standard deviation */
	long n = 1, ni = 1;
	for (int a = 0; a < 3; ++a) { n *= s->gs[a]; if (a < s->dim) ni *= s->gs[a] - 2 * NG; }
	real mn = INFINITY, mx = -INFINITY, sum = 0;
	for (long c = 0; c < n; ++c) {
		real v = var_value(s, c, field);
		mn = fmin(mn, v);
		mx = fmax(mx, v);
	}
	int lo[3], hi[3];
	for (int a = 0; a < 3; ++a) {
		lo[a] = a < s->dim ? NG : 0;
		hi[a] = a < s->dim ? s->gs[a] - NG : 1;
	}
	/* serial loops: one fixed summation order */
	for (int k = lo[2]; k < hi[2]; ++k)
		for (int j = lo[1]; j < hi[1]; ++j)
			for (int i = lo[0]; i < hi[0]; ++i) sum += var_value(s, i + s->step[1] * j + s->step[2] * k, field);
	real avg = sum / (real)ni;
	real sq = 0;
	for (int k = lo[2]; k < hi[2]; ++k)
		for (int j = lo[1]; j < hi[1]; ++j)
			for (int i = lo[0]; i < hi[0]; ++i) {
				real d = var_value(s, i + s->step[1] * j + s->step[2] * k, field) - avg;
				sq += d * d;
			}
	out[0] = mn; out[1] = mx; out[2] = avg; out[3] = sqrt(sq / (real)ni);
}

void oracle_constrain_U(oracle_solver *s) {
	/* solverbase.lua:2116-2127: only equations with a constrainU module */
	if (s->cfg.eqn != HYDRO_EQN_EULER) return;
	int S = s->S;
	FOR_CELLS(s, 0, 0, {
		oracle_euler_constrain(&s->p, s->U + index * S);
	});
	oracle_boundary(s);
}

double oracle_calc_dt(oracle_solver *s) {
	if (s->cfg.use_fixed_dt) return s->cfg.fixed_dt;
	int S = s->S;
	real *dtBuf = s->dtBuf;
	FOR_CELLS(s, 0, 0, {
		dtBuf[index] = INFINITY;
		int oob = 0;
		for (int d = 0; d < s->dim; ++d) if (iv[d] < NG || iv[d] >= s->gs[d] - NG) oob = 1;
		if (!oob) {
			const real *U = s->U + index * S;
			dtBuf[index] = s->cfg.eqn == HYDRO_EQN_EULER ? oracle_euler_dt_cell(&s->p, U, s->dim, s->dx)
				: s->cfg.eqn == HYDRO_EQN_SHALLOW_WATER ? oracle_sw_dt_cell(&s->p, U, s->dim, s->dx)
				: glm_dt_cell(&s->p, U, s->dim, s->dx);
		}
	});
	real m = INFINITY;
	#pragma omp parallel for reduction(min:m)
	for (long i = 0; i < (long)s->G; ++i) if (dtBuf[i] < m) m = dtBuf[i];
	return s->cfg.cfl * m;
}

static void calc_lr(oracle_solver *s, const real *U, real dt) {
	/* calcLR (fvsolver.lua:232-234) with the stage's dt */
	int S = s->S, dim = s->dim;
	FOR_CELLS(s, 1, 1, {
		for (int side = 0; side < dim; ++side) {
			real *lr = s->ULR + (side + dim * index) * 2 * S;
			if (s->cfg.plm == HYDRO_PLM_EIG_PRIM || s->cfg.plm == HYDRO_PLM_EIG_PRIM_REF)
				oracle_plm_eig_prim(&s->p, s->cfg.plm == HYDRO_PLM_EIG_PRIM_REF,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					side, dt / s->dx[side], lr, lr + S);
			else if (s->cfg.plm == HYDRO_PLM_ATHENA)
				oracle_plm_athena(&s->p,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					side, lr, lr + S);
			else if (s->cfg.plm == HYDRO_PLM_PPM)
				oracle_plm_ppm(s->cfg.slope_limiter, s->q, S,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					lr, lr + S);
			else if (s->cfg.plm == HYDRO_PLM_CONS_FLUX)
				oracle_plm_cons_flux(&s->p, s->cfg.slope_limiter,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					side, dt / s->dx[side], lr, lr + S);
			else if (s->cfg.plm == HYDRO_PLM_EIG)
				oracle_plm_eig(&s->p, s->cfg.slope_limiter,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					side, dt / s->dx[side], lr, lr + S);
			else if (s->cfg.plm == HYDRO_PLM_PRIM)
				oracle_plm_prim(&s->p, s->cfg.slope_limiter,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					lr, lr + S);
			else
				oracle_plm_cons(s->cfg.slope_limiter, s->q, S,
					U + (index - s->step[side]) * S, U + index * S, U + (index + s->step[side]) * S,
					lr, lr + S);
		}
	});
}

static void calc_flux(oracle_solver *s, const real *U, real dt) {
	int S = s->S, dim = s->dim;
	int plm = s->cfg.plm;
	int lim = s->cfg.flux_limiter;
	FOR_CELLS(s, NG, NG - 1, {
		for (int side = 0; side < dim; ++side) {
			long indexL = index - s->step[side];
			real *F = s->flux + (side + dim * index) * S;
			/* cell_area<side> for Cartesian: product of the other used dx */
			real area = 1.;
			for (int o = 0; o < dim; ++o) if (o != side) area *= s->dx[o];
			if (area <= 1e-7) {
				for (int j = 0; j < S; ++j) F[j] = 0;
				continue;
			}
			const real *UL, *UR;
			if (plm) {
				UL = s->ULR + (side + dim * indexL) * 2 * S + S;
				UR = s->ULR + (side + dim * index) * 2 * S;
			} else {
				UL = U + indexL * S;
				UR = U + index * S;
			}
			real dt_dx = dt / s->dx[side];
			const real *limargs[4] = {
				U + (indexL - s->step[side]) * S, U + indexL * S,
				U + index * S, U + (index + s->step[side]) * S};
			if (s->cfg.flux == HYDRO_FLUX_HLL) oracle_hll_flux(&s->p, UL, UR, side, F);
			else if (s->cfg.flux == HYDRO_FLUX_RUSANOV) oracle_rusanov_flux(&s->p, UL, UR, side, F);
			else if (s->cfg.flux == HYDRO_FLUX_HLLC) oracle_hllc_flux(&s->p, UL, UR, side, F);
			else if (s->cfg.flux == HYDRO_FLUX_MARQUINA) oracle_marquina_flux(&s->p, UL, UR, side, F);
			else
				oracle_roe_flux(&s->p, s->cfg.eqn, lim, dt_dx, UL, UR,
					lim != HYDRO_LIMITER_DONOR_CELL ? limargs : NULL, side, F);
		}
	});
}

static void calc_deriv_from_flux(oracle_solver *s, real *deriv) {
	int S = s->S, dim = s->dim, q = s->q;
	real volume = 1.;
	for (int o = 0; o < dim; ++o) volume *= s->dx[o];
	FOR_CELLS(s, NG, NG, {
		real *dv = deriv + index * S;
		for (int side = 0; side < dim; ++side) {
			const real *fL = s->flux + (side + dim * index) * S;
			const real *fR = s->flux + (side + dim * (index + s->step[side])) * S;
			real areaL = 1.;
			for (int o = 0; o < dim; ++o) if (o != side) areaL *= s->dx[o];
			real areaR = areaL;
			if (volume > 1e-7) {
				real invVolume = 1. / volume;
				if (areaL <= 1e-7) areaL = 0.;
				if (areaR <= 1e-7) areaR = 0.;
				areaL *= invVolume;
				areaR *= invVolume;
				for (int j = 0; j < q; ++j) dv[j] -= fR[j] * areaR - fL[j] * areaL;
			}
		}
	});
}

static void glm_add_source(oracle_solver *s, real *deriv, const real *Ubuf) {
	/* glm-maxwell.cl:317-380 (the eps-gradient loop writes grad_1_mu as the
	 * reference does; grad_1_eps stays zero) */
	int S = s->S, dim = s->dim;
	FOR_CELLS(s, NG, NG, {
		real *dv = deriv + index * S;
		const real *U = Ubuf + index * S;
		real c = U[10] * U[10] * U[9];
		for (int i = 0; i < 3; ++i) dv[i] = dv[i] - U[i] * c;
		real grad_1_mu[3] = {0, 0, 0}, grad_1_eps[3] = {0, 0, 0};
		for (int j = 0; j < dim; ++j) {
			const real *Up = U + s->step[j] * S, *Um = U - s->step[j] * S;
			grad_1_mu[j] = (Up[11] * Up[11] - Um[11] * Um[11]) * (1. / s->dx[j]);
		}
		for (int j = 0; j < dim; ++j) {
			const real *Up = U + s->step[j] * S, *Um = U - s->step[j] * S;
			grad_1_mu[j] = (Up[10] * Up[10] - Um[10] * Um[10]) * (1. / s->dx[j]);
		}
		real _1_sqrt_det_g = 1. / 1.;
		for (int j = 0; j < dim; ++j) {
			real F[12];
			oracle_glm_flux_from_cons(&s->p, U, j, F);
			for (int i = 0; i < 6; ++i) F[i] = F[i] * _1_sqrt_det_g;
			dv[j] = dv[j] - dot3(F, grad_1_mu);
			dv[3 + j] = dv[3 + j] - dot3(F + 3, grad_1_eps);
		}
		dv[6] = dv[6] + U[8] * s->p.div_phi;
	});
}

static void sw_add_source(oracle_solver *s, real *deriv, const real *Ubuf) {
	/* shallow-water.cl:366-428, the enabled block: the second-order central
	 * difference of depth (makePartial1 with numGhost = 2: coefficient 1,
	 * denominator 2 dx; makepartial.lua:13-20, 85-110, eqn.lua:1345-1350),
	 * divided by the basis lengths (1) and raised (identity), times g h,
	 * added to the momentum derivative */
	int S = s->S, dim = s->dim;
	FOR_CELLS(s, NG, NG, {
		real *dv = deriv + index * S;
		const real *U = Ubuf + index * S;
		real partial_depth_l[3] = {0., 0., 0.};
		for (int j = 0; j < dim; ++j) {
			const real *Up = U + s->step[j] * S, *Um = U - s->step[j] * S;
			partial_depth_l[j] = (0. + (Up[4] - Um[4]) * 1.) * (1. / (s->dx[j] * 2.));
		}
		real e_depth_u[3];
		for (int j = 0; j < 3; ++j) e_depth_u[j] = partial_depth_l[j] / 1.;
		real gh = s->p.gravity * U[0];
		for (int j = 0; j < 3; ++j) dv[1 + j] = dv[1 + j] + e_depth_u[j] * gh;
	});
}

void oracle_calc_deriv(oracle_solver *s, double dt, double *deriv) {
	/* FiniteVolumeSolver:calcDeriv (fvsolver.lua:225-302) + addSource
	 * (solverbase.lua:3208-3214); the derivative buffer is cleared first
	 * (int.lua:6-8) */
	memset(deriv, 0, s->G * s->S * sizeof(real));
	if (s->cfg.plm) calc_lr(s, s->U, dt);
	calc_flux(s, s->U, dt);
	calc_deriv_from_flux(s, deriv);
	if (s->cfg.eqn == HYDRO_EQN_GLM_MAXWELL) glm_add_source(s, deriv, s->U);
	if (s->cfg.eqn == HYDRO_EQN_SHALLOW_WATER) sw_add_source(s, deriv, s->U);
}

static void mult_add(oracle_solver *s, real *a, const real *b, const real *c, real d) {
	/* solverbase.lua:1227-1242 */
	int S = s->S, q = s->q;
	FOR_CELLS(s, NG, NG, {
		for (int j = 0; j < q; ++j) a[index * S + j] = b[index * S + j] + c[index * S + j] * d;
	});
}
static void mult_add_into(oracle_solver *s, real *a, const real *b, real c) {
	/* solverbase.lua:1213-1225 */
	int S = s->S, q = s->q;
	FOR_CELLS(s, NG, NG, {
		for (int j = 0; j < q; ++j) a[index * S + j] += b[index * S + j] * c;
	});
}

void oracle_step(oracle_solver *s, double dt) {
	size_t bytes = s->G * s->S * sizeof(real);
	if (s->cfg.integrator == HYDRO_INT_FORWARD_EULER) {
		/* fe.lua:33-49 */
		oracle_calc_deriv(s, dt, s->deriv);
		mult_add_into(s, s->U, s->deriv, dt);
		oracle_boundary(s);
		oracle_constrain_U(s);
		return;
	}
	/* rk.lua:47-167 */
	rk_table *t = &s->rk;
	int o = t->order;
	if (s->needU[0]) memcpy(s->UBufs[0], s->U, bytes);
	if (s->needD[0]) oracle_calc_deriv(s, dt, s->derivBufs[0]);
	for (int i = 1; i <= o; ++i) {
		memset(s->U, 0, bytes);
		/* The reference clears the whole UBuf (rk.lua:155) and multAdd writes
		 * only numIntStates fields, which zeroes the non-integrated fields
		 * every stage (Euler ePot; GLM-Maxwell's rhoCharge, sigma, sqrt_1_eps
		 * and sqrt_1_mu, where it makes the next eigen transform divide by 0).
		 * This build carries those fields through unchanged. */
		if (s->S > s->q) {
			int S = s->S, q = s->q;
			const real *U0 = s->UBufs[0];
			#pragma omp parallel for schedule(static)
			for (long c = 0; c < (long)s->G; ++c)
				for (int j = q; j < S; ++j) s->U[c * S + j] = U0[c * S + j];
		}
		for (int k = 0; k < i; ++k)
			if (t->a[i - 1][k] != 0) mult_add(s, s->U, s->U, s->UBufs[k], t->a[i - 1][k]);
		for (int k = 0; k < i; ++k)
			if (t->b[i - 1][k] != 0) mult_add(s, s->U, s->U, s->derivBufs[k], t->b[i - 1][k] * dt);
		oracle_boundary(s);
		oracle_constrain_U(s);
		if (i < o) {
			if (s->needU[i]) memcpy(s->UBufs[i], s->U, bytes);
			if (s->needD[i]) oracle_calc_deriv(s, dt, s->derivBufs[i]);
		}
	}
}

double oracle_update(oracle_solver *s) {
	/* solverbase.lua:3026-3190 */
	double dt = oracle_calc_dt(s);
	oracle_step(s, dt);
	s->t = s->t + dt;
	oracle_boundary(s);
	return dt;
}
