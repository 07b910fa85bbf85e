// fv_glm.h — GLM-Maxwell equations (hydro/eqn/glm-maxwell.cl, glm-maxwell.lua)
// for the fused stage kernels.
//
// cons_t: D.xyz (0-2), B.xyz (3-5), phi (6), psi (7) are integrated; rhoCharge
// (8), sigma (9), sqrt_1_eps (10), sqrt_1_mu (11) are static per-cell fields
// kept in their own planes. The transforms keep the reference's expressions
// term for term, including the products with the normal components, so the
// arithmetic is the reference's; unit_m_per_s and coord_sqrt_det_g are 1 for
// Cartesian grids in the reference's default units.
#pragma once

#include "fv_common.h"
#include "fv_math.h"

namespace hydro {

// PLMK: HYDRO_PLM_CONS, or HYDRO_PLM_PPM (plm.cl:882-944), which like
// 'plm cons' takes the integrated fields' differences only.
template <bool FAST, int PLMK = HYDRO_PLM_CONS>
struct GLMT {
template <typename R>
struct Eqn {
	using M = Math<R, FAST>;
	static constexpr bool FASTM = FAST;
	static constexpr int FLUXID = HYDRO_FLUX_ROE;
	static constexpr bool PLM_PRIM = false;
	static constexpr bool PLM_EIG = false;
	static constexpr bool PLM_CONS_FLUX = false;
	static constexpr bool PLM_ATHENA = false;
	static constexpr bool PLM_PPM = PLMK == HYDRO_PLM_PPM;
	static constexpr bool PLM_EIG_PRIM = false;
	static constexpr bool PLM_EIG_PRIM_REF = false;
	static constexpr int Q = 8;
	static constexpr int NSTAT = 4;
	static constexpr int NW = 8;
	static constexpr int NSTATES = 12;
	static constexpr bool SOURCE = true;
	__device__ static constexpr bool active(int f) { return f < 12; }

	struct Par {
		R divPhi, divPsi;
	};
	struct Prim {};
	struct Eig {
		R e, m;   // sqrt_1_eps, sqrt_1_mu at the interface
	};

	__device__ static __forceinline__ Par par(const StageArgs<R> &a) {
		Par p;
		p.divPhi = a.divPhi;
		p.divPsi = a.divPsi;
		return p;
	}

	__device__ static __forceinline__ Prim prim(const Par &, const R *) { return Prim{}; }

	// eigen_forInterface (glm-maxwell.cl:138-153)
	__device__ static __forceinline__ Eig eigen(const Par &, const R *UL, const R *UR, const Prim &, const Prim &) {
		Eig g;
		g.e = (UL[10] + UR[10]) * R(.5);
		g.m = (UL[11] + UR[11]) * R(.5);
		return g;
	}

	// normal_l<j> components for axis S (coord.lua:2376-2385)
	template <int S, int J, int I>
	__device__ static constexpr R nl() { return (((I - J) % 3 + 3) % 3) == S ? R(1) : R(0); }

	// eigen_leftTransform (glm-maxwell.cl:175-252)
	template <int S>
	__device__ static __forceinline__ void left(const Par &, const Eig &g, const R *X, R *Y) {
		if constexpr (FAST) {
			// the same sums without their zero terms: with the normal along
			// axis A and the tangents B, C of coord.lua's normal_l
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			const R e = g.e, m = g.m;
			const R h = M::rcp(e) * R(.5);
			const R mB = X[3 + B] * m, mC = X[3 + C] * m;
			const R eB = X[B] * e, eC = X[C] * e;
			Y[0] = h * (X[6] - X[A]);
			Y[1] = h * (X[7] - X[3 + A]);
			Y[2] = R(.5) * (mB + eC);
			Y[3] = R(.5) * (mC - eB);
			Y[4] = R(.5) * (mB - eC);
			Y[5] = R(.5) * (mC + eB);
			Y[6] = h * (X[6] + X[A]);
			Y[7] = h * (X[7] + X[3 + A]);
			return;
		}
		const R n_l[3] = {nl<S, 1, 1>(), nl<S, 1, 2>(), nl<S, 1, 3>()};
		const R n2_l[3] = {nl<S, 2, 1>(), nl<S, 2, 2>(), nl<S, 2, 3>()};
		const R n3_l[3] = {nl<S, 3, 1>(), nl<S, 3, 2>(), nl<S, 3, 3>()};
		const R e = g.e, m = g.m;
		R tmp1 = M::rcp(e);
		R tmp2 = tmp1 * n_l[0];
		R tmp3 = R(1) / R(2);
		R tmp4 = X[0] * tmp2;
		R tmp5 = tmp3 * tmp4;
		R tmp7 = n_l[1] * tmp1;
		R tmp9 = X[1] * tmp7;
		R tmp11 = tmp9 * tmp3;
		R tmp13 = n_l[2] * tmp1;
		R tmp15 = X[2] * tmp13;
		R tmp17 = tmp15 * tmp3;
		R tmp20 = X[6] * tmp1;
		R tmp22 = tmp20 * tmp3;
		R tmp26 = X[3] * tmp2;
		R tmp27 = tmp26 * tmp3;
		R tmp31 = X[4] * tmp7;
		R tmp33 = tmp31 * tmp3;
		R tmp37 = X[5] * tmp13;
		R tmp39 = tmp37 * tmp3;
		R tmp42 = X[7] * tmp1;
		R tmp44 = tmp42 * tmp3;
		R tmp45 = m * n2_l[2];
		R tmp47 = X[5] * tmp45;
		R tmp48 = m * n2_l[1];
		R tmp50 = X[4] * tmp48;
		R tmp51 = m * n2_l[0];
		R tmp53 = X[3] * tmp51;
		R tmp54 = e * n3_l[2];
		R tmp56 = X[2] * tmp54;
		R tmp57 = e * n3_l[1];
		R tmp59 = X[1] * tmp57;
		R tmp60 = e * n3_l[0];
		R tmp62 = X[0] * tmp60;
		R tmp63 = tmp59 * tmp3;
		R tmp64 = tmp62 * tmp3;
		R tmp65 = tmp56 * tmp3;
		R tmp67 = tmp53 * tmp3;
		R tmp69 = tmp50 * tmp3;
		R tmp71 = tmp47 * tmp3;
		R tmp73 = e * n2_l[0];
		R tmp75 = X[0] * tmp73;
		R tmp76 = tmp75 * tmp3;
		R tmp77 = e * n2_l[1];
		R tmp79 = X[1] * tmp77;
		R tmp81 = tmp79 * tmp3;
		R tmp82 = e * n2_l[2];
		R tmp84 = X[2] * tmp82;
		R tmp86 = tmp84 * tmp3;
		R tmp87 = m * n3_l[0];
		R tmp89 = X[3] * tmp87;
		R tmp90 = m * n3_l[1];
		R tmp92 = X[4] * tmp90;
		R tmp93 = m * n3_l[2];
		R tmp95 = X[5] * tmp93;
		R tmp96 = tmp92 * tmp3;
		R tmp97 = tmp95 * tmp3;
		R tmp98 = tmp89 * tmp3;
		Y[0] = tmp22 + -tmp5 - tmp11 - tmp17;
		Y[1] = tmp44 + -tmp27 - tmp33 - tmp39;
		Y[2] = tmp71 + tmp69 + tmp67 + tmp65 + tmp63 + tmp64;
		Y[3] = tmp98 + tmp96 + tmp97 + -tmp76 - tmp81 - tmp86;
		Y[4] = tmp67 + tmp69 + tmp71 + -tmp64 - tmp63 - tmp65;
		Y[5] = tmp97 + tmp96 + tmp98 + tmp86 + tmp81 + tmp76;
		Y[6] = tmp22 + tmp17 + tmp11 + tmp5;
		Y[7] = tmp44 + tmp39 + tmp33 + tmp27;
	}

	// eigen_rightTransform (glm-maxwell.cl:257-298)
	template <int S>
	__device__ static __forceinline__ void right(const Par &, const Eig &g, const R *X, R *Y) {
		if constexpr (FAST) {
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			const R e = g.e;
			const R ie = M::rcp(e), im = M::rcp(g.m);
			Y[A] = e * (X[6] - X[0]);
			Y[B] = ie * (X[5] - X[3]);
			Y[C] = ie * (X[2] - X[4]);
			Y[3 + A] = e * (X[7] - X[1]);
			Y[3 + B] = im * (X[4] + X[2]);
			Y[3 + C] = im * (X[5] + X[3]);
			Y[6] = e * (X[6] + X[0]);
			Y[7] = e * (X[7] + X[1]);
			return;
		}
		const R n_l[3] = {nl<S, 1, 1>(), nl<S, 1, 2>(), nl<S, 1, 3>()};
		const R n2_l[3] = {nl<S, 2, 1>(), nl<S, 2, 2>(), nl<S, 2, 3>()};
		const R n3_l[3] = {nl<S, 3, 1>(), nl<S, 3, 2>(), nl<S, 3, 3>()};
		const R e = g.e, m = g.m;
		R tmp1 = e * n_l[0];
		R tmp3 = M::rcp(e);
		R tmp4 = tmp3 * n3_l[0];
		R tmp6 = n2_l[0] * tmp3;
		R tmp22 = e * n_l[1];
		R tmp25 = n3_l[1] * tmp3;
		R tmp27 = n2_l[1] * tmp3;
		R tmp43 = e * n_l[2];
		R tmp46 = n3_l[2] * tmp3;
		R tmp48 = n2_l[2] * tmp3;
		R tmp66 = M::rcp(m);
		R tmp67 = tmp66 * n2_l[0];
		R tmp69 = n3_l[0] * tmp66;
		R tmp88 = n2_l[1] * tmp66;
		R tmp90 = n3_l[1] * tmp66;
		R tmp109 = n2_l[2] * tmp66;
		R tmp111 = n3_l[2] * tmp66;
		Y[0] = X[6] * tmp1 + X[5] * tmp6 + X[2] * tmp4 - X[3] * tmp6 - X[4] * tmp4 + -X[0] * tmp1;
		Y[1] = X[6] * tmp22 + X[5] * tmp27 + X[2] * tmp25 - X[3] * tmp27 - X[4] * tmp25 + -X[0] * tmp22;
		Y[2] = X[6] * tmp43 + X[5] * tmp48 + X[2] * tmp46 - X[3] * tmp48 - X[4] * tmp46 + -X[0] * tmp43;
		Y[3] = X[7] * tmp1 + X[5] * tmp69 + X[4] * tmp67 + X[3] * tmp69 + X[2] * tmp67 + -X[1] * tmp1;
		Y[4] = X[7] * tmp22 + X[5] * tmp90 + X[4] * tmp88 + X[3] * tmp90 + X[2] * tmp88 + -X[1] * tmp22;
		Y[5] = X[7] * tmp43 + X[5] * tmp111 + X[4] * tmp109 + X[3] * tmp111 + X[2] * tmp109 + -X[1] * tmp43;
		Y[6] = X[6] * e + X[0] * e;
		Y[7] = X[7] * e + X[1] * e;
	}

	// fluxFromCons (glm-maxwell.cl:99-130)
	template <int S>
	__device__ static __forceinline__ void flux(const Par &p, const R *U, const Prim &, R *F) {
		R a2 = U[10] * U[10], b2 = U[11] * U[11];
		if constexpr (FAST) {
			// F_D = H x n + n phi divPhi, F_B = n psi divPsi - E x n, n along A
			constexpr int A = S, B = (S + 1) % 3, C = (S + 2) % 3;
			F[A] = U[6] * p.divPhi;
			F[B] = U[3 + C] * b2;
			F[C] = -(U[3 + B] * b2);
			F[3 + A] = U[7] * p.divPsi;
			F[3 + B] = -(U[C] * a2);
			F[3 + C] = U[B] * a2;
			F[6] = U[S] * p.divPhi;
			F[7] = U[3 + S] * p.divPsi;
			return;
		}
		R E0 = U[0] * a2, E1 = U[1] * a2, E2 = U[2] * a2;
		R H0 = U[3] * b2, H1 = U[4] * b2, H2 = U[5] * b2;
		const R nx = nl<S, 1, 1>(), ny = nl<S, 1, 2>(), nz = nl<S, 1, 3>();
		R phi = U[6], psi = U[7];
		F[0] = H1 * nz - H2 * ny + nx * phi * p.divPhi;
		F[3] = E2 * ny - E1 * nz + nx * psi * p.divPsi;
		F[1] = H2 * nx - H0 * nz + ny * phi * p.divPhi;
		F[4] = E0 * nz - E2 * nx + ny * psi * p.divPsi;
		F[2] = H0 * ny - H1 * nx + nz * phi * p.divPhi;
		F[5] = E1 * nx - E0 * ny + nz * psi * p.divPsi;
		F[6] = U[S] * p.divPhi;
		F[7] = U[3 + S] * p.divPsi;
	}

	// eigenWaveCode (glm-maxwell.lua:181-214)
	template <int S>
	__device__ static __forceinline__ void waves(const Par &p, const Eig &g, R *l) {
		R v = g.e * g.m * (R(1) / R(1));
		l[0] = -p.divPhi; l[1] = -p.divPsi;
		l[2] = -v; l[3] = -v; l[4] = v; l[5] = v;
		l[6] = p.divPhi; l[7] = p.divPsi;
	}

	__device__ static __forceinline__ void constrain(double, double, double, double *) {}

	// addSource (glm-maxwell.cl:317-380) for one interior cell: U the cell's
	// cons_t, em / ep the sqrt_1_eps of its -1 / +1 neighbours per axis. The
	// reference's second gradient loop stores into grad_1_mu (so grad_1_mu ends
	// as the sqrt_1_eps gradient and grad_1_eps stays zero); kept as is.
	static constexpr int SRC_FIELD = 10;   // the source reads the neighbours' sqrt_1_eps
	template <int DIM>
	__device__ static __forceinline__ void source(const Par &p, const GridDesc<R> &gd, const R *U, const R *em,
		const R *ep, R *d)
	{
		const R *invdx = gd.invdx;
		R c = U[10] * U[10] * U[9];
		d[0] = d[0] - U[0] * c;
		d[1] = d[1] - U[1] * c;
		d[2] = d[2] - U[2] * c;
		R g[3] = {R(0), R(0), R(0)};
		const R z[3] = {R(0), R(0), R(0)};
#pragma unroll
		for (int j = 0; j < DIM; ++j) g[j] = (ep[j] * ep[j] - em[j] * em[j]) * invdx[j];
		Prim W;
		R F[8];
		if constexpr (FAST) {
			// the same sums without the zero gradients (axes beyond DIM, and
			// the B rows' zero grad_1_eps)
#pragma unroll
			for (int i = 0; i < DIM; ++i) {
				if (i == 0) flux<0>(p, U, W, F);
				if (i == 1) flux<1>(p, U, W, F);
				if (i == 2) flux<2>(p, U, W, F);
				R s = F[0] * g[0];
				if constexpr (DIM > 1) s = s + F[1] * g[1];
				if constexpr (DIM > 2) s = s + F[2] * g[2];
				d[i] = d[i] - s;
			}
			d[6] = d[6] + U[8] * p.divPhi;
			return;
		}
		flux<0>(p, U, W, F);
		d[0] = d[0] - (F[0] * g[0] + F[1] * g[1] + F[2] * g[2]);
		d[3] = d[3] - (F[3] * z[0] + F[4] * z[1] + F[5] * z[2]);
		if constexpr (DIM > 1) {
			flux<1>(p, U, W, F);
			d[1] = d[1] - (F[0] * g[0] + F[1] * g[1] + F[2] * g[2]);
			d[4] = d[4] - (F[3] * z[0] + F[4] * z[1] + F[5] * z[2]);
		}
		if constexpr (DIM > 2) {
			flux<2>(p, U, W, F);
			d[2] = d[2] - (F[0] * g[0] + F[1] * g[1] + F[2] * g[2]);
			d[5] = d[5] - (F[3] * z[0] + F[4] * z[1] + F[5] * z[2]);
		}
		d[6] = d[6] + U[8] * p.divPhi;
	}

	// calcDTCell with the GLM wave bounds (glm-maxwell.lua:216-245)
	template <int DIM>
	__device__ static __forceinline__ double dt_cell(const EqnParD &pd, const double *dx, const double *U) {
		using MD = Math<double, FAST>;
		const double divPhi = pd.divPhi, divPsi = pd.divPsi;
		double dt = INFINITY;
		for (int s = 0; s < DIM; ++s) {
			if (dx[s] > 0.) {
				double v_p_abs = U[10] * U[11] * (1. / 1.);
				double wmax = fmax(fmax(divPsi, divPhi), v_p_abs);
				double absLambdaMax = fmax(fabs(wmax), fabs(-wmax));
				absLambdaMax = fmax(1e-9, absLambdaMax);
				dt = fmin(dt, MD::div(dx[s], absLambdaMax));
			}
		}
		return dt;
	}

	// mirror reflects D and B along the axis (reflectVars.mirror, eqn.lua:366-370)
	__device__ static __forceinline__ bool reflects(int field, int axis) { return field == axis || field == 3 + axis; }
};
};

} // namespace hydro
