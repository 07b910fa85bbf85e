// fv_limiters.h — flux / slope limiter table of hydro/app.lua:614-636.
//
// LIM >= 0 selects one limiter at compile time; LIM == -1 dispatches on the
// runtime index (uniform across the wave). van Leer uses
// `r > 0 ? r * 2 / (1 + r) : 0`, equal to the reference's
// `max(0., r) * 2. / (1. + r)` except at r = -1, where the reference divides
// 0 by 0.
#pragma once

#include "fv_common.h"
#include "fv_math.h"

namespace hydro {

template <typename R>
__device__ __forceinline__ R limiter_rt(int index, R r) {
	const R z = R(0), one = R(1), two = R(2);
	switch (index) {
	case 0: return z;
	case 1: return one;
	case 2: return r;
	case 3: return R(.5) * (one + r);
	case 4: return fmax(z, r) * (R(3) * r + one) / ((r + one) * (r + one));
	case 5: return fmax(z, R(1.5) * (r + fabs(r)) / (r + two));
	case 6: return fmax(z, two * (r + fabs(r)) / (r + R(3)));
	case 7: return fmax(z, fmin(two * r, fmin((one + two * r) / R(3), two)));
	case 8: return fmax(z, fmin(r, one));
	case 9: return fmax(z, fmin(r, R(1.5)));
	case 10: return R(.5) * (r * r + r) / (r * r + r + one);
	case 11: return fmax(z, fmin(two * r, fmin(R(.25) + R(.75) * r, R(4))));
	case 12: return fmax(z, fmax(fmin(R(1.5) * r, one), fmin(r, R(1.5))));
	case 13: return fmax(z, fmin(fmin(two * r, R(.75) + R(.25) * r), fmin(R(.25) + R(.75) * r, two)));
	case 14: return (r * r + r) / (r * r + one);
	case 15: return two * r / (r * r + one);
	case 16: return r > z ? r * two / (one + r) : z;
	case 17: return fmax(z, fmin(two, fmin(R(.5) * (one + r), two * r)));
	case 18: return fmax(z, fmax(fmin(one, two * r), fmin(two, r)));
	case 19: return R(.5) * (r + one) * fmin(one, fmin(R(4) * r / (r + one), R(4) / (r + one)));
	}
	return z;
}

template <int LIM, typename R>
__device__ __forceinline__ R limiter(int index, R r) {
	if constexpr (LIM == HYDRO_LIMITER_MINMOD) return fmax(R(0), fmin(r, R(1)));
	else if constexpr (LIM == HYDRO_LIMITER_SUPERBEE) return fmax(R(0), fmax(fmin(R(1), R(2) * r), fmin(R(2), r)));
	else if constexpr (LIM == HYDRO_LIMITER_VAN_LEER) return r > R(0) ? r * R(2) / (R(1) + r) : R(0);
	else if constexpr (LIM == HYDRO_LIMITER_MONOTIZED_CENTRAL) return fmax(R(0), fmin(R(2), fmin(R(.5) * (R(1) + r), R(2) * r)));
	else return limiter_rt<R>(index, r);
}

// phi(num / den) * den, the limited quantity both the slope limiter
// (plm.cl:56-75) and the Roe flux limiter (roe.cl:120-142) need, for the fast
// arithmetic policy. For limiters that are piecewise linear in r the product
// needs no division: with a = |num|, b = |den| and num, den of one sign,
// phi(a / b) * b is a max/min of a, b, 2a, 2b, (a + b) / 2 carrying den's sign,
// and 0 when the signs differ or either is 0 (r <= 0). van Leer needs one
// division, 2ab / (a + b). Other limiters form r with the policy's division.
// The results differ from the literal phi(num / den) * den by the rounding of
// r * den only.
__host__ __device__ constexpr bool limited_symmetric(int LIM) {
	return LIM == HYDRO_LIMITER_MINMOD || LIM == HYDRO_LIMITER_SUPERBEE
		|| LIM == HYDRO_LIMITER_MONOTIZED_CENTRAL || LIM == HYDRO_LIMITER_VAN_LEER;
}

template <int LIM, bool FAST, typename R>
__device__ __forceinline__ R limited(int index, R num, R den) {
	using M = Math<R, FAST>;
	if constexpr (FAST && LIM == HYDRO_LIMITER_MINMOD) {
		// minmod(a, b): the operand nearer 0 when both have one sign, else 0,
		// without compares: max(min(a, b), 0) + min(max(a, b), 0)
		return fmax(fmin(num, den), R(0)) + fmin(fmax(num, den), R(0));
	} else if constexpr (FAST && LIM == HYDRO_LIMITER_VAN_LEER) {
		// (num |den| + |num| den) / (|num| + |den|) without sign tests: with
		// one sign it is 2ab / (a + b) carrying den's sign (the two products
		// are one rounded value each, so the sum doubles it exactly, as 2a * b
		// does); with opposite signs the products cancel to +0. The products
		// are separate statements so that neither is fused into the sum. Both
		// 0, or a NaN: |num| + |den| is not positive and the result is 0.
		const R p1 = num * fabs(den);
		const R p2 = fabs(num) * den;
		const R n = p1 + p2;
		const R d = fabs(num) + fabs(den);
		const R q = M::div(n, d);
		return d > R(0) ? q : R(0);
	} else if constexpr (FAST && limited_symmetric(LIM)) {
		const bool same = (num > R(0) && den > R(0)) || (num < R(0) && den < R(0));
		const R a = fabs(num), b = fabs(den);
		R m;
		if constexpr (LIM == HYDRO_LIMITER_MINMOD) m = fmin(a, b);
		else if constexpr (LIM == HYDRO_LIMITER_SUPERBEE) m = fmax(fmin(b, R(2) * a), fmin(R(2) * b, a));
		else if constexpr (LIM == HYDRO_LIMITER_MONOTIZED_CENTRAL) m = fmin(R(2) * b, fmin(R(.5) * (a + b), R(2) * a));
		else m = M::div(R(2) * a * b, same ? a + b : R(1));
		return same ? copysign(m, den) : R(0);
	} else {
		R r = den == R(0) ? R(0) : (FAST ? M::div(num, den) : num / den);
		return limiter<LIM>(index, r) * den;
	}
}

} // namespace hydro
