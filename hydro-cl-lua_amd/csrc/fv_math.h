// fv_math.h — division / square-root policy of the stage kernels.
//
// EXACT: IEEE division and sqrt, i.e. the arithmetic of the reference's
// kernels (correctly rounded; the compiler expands them into the scaled
// v_div_* / v_rsq sequences).
// FAST: short hardware sequences. FP64: v_rcp_f64 + one Newton step + a
// residual correction for x / y (6 instructions), v_rsq_f64 + one Goldschmidt
// step + correction for sqrt; both are faithfully rounded and matched the
// correctly rounded results on 4M random operands on gfx950. Where the divisor
// is known to be a positive O(1) quantity (densities, sound speeds) the
// reciprocal is shared between divisions. FP32: v_rcp_f32 / v_sqrt_f32.
// General divisions (limiter ratios, FP64 and FP32) fall back to IEEE division
// when any lane of the wave has a zero, huge, tiny or non-finite divisor.
#pragma once

#include <hip/hip_runtime.h>

namespace hydro {

// Wave-wide votes over the active lanes, from the lane predicate's own
// compare mask (HIP's __all / __any first turn the predicate into an integer
// and compare it again).
__device__ __forceinline__ bool wave_all(bool p) { return __builtin_amdgcn_ballot_w64(!p) == 0; }
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

template <typename R, bool FAST>
struct Math;

template <>
struct Math<double, false> {
	__device__ static __forceinline__ double rcp(double y) { return 1. / y; }
	__device__ static __forceinline__ double div(double x, double y) { return x / y; }
	// x / y given inv = rcp(y): exact division
	__device__ static __forceinline__ double div_inv(double x, double y, double) { return x / y; }
	__device__ static __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
	__device__ static __forceinline__ double sqrt_pos(double x) { return __builtin_sqrt(x); }
	__device__ static __forceinline__ double div_pos(double x, double y) { return x / y; }
};

template <>
struct Math<double, true> {
	__device__ static __forceinline__ double rcp(double y) {
		double r = __builtin_amdgcn_rcp(y);
		double e = __builtin_fma(-y, r, 1.);
		r = __builtin_fma(e, r, r);
		e = __builtin_fma(-y, r, 1.);
		return __builtin_fma(e, r, r);
	}
	// x / y given inv = rcp(y): x * inv, within 1.5 ulp of the quotient (inv
	// is faithful; the residual correction bought half an ulp for two FMAs)
	__device__ static __forceinline__ double div_inv(double x, double, double inv) { return x * inv; }
	// positive, normal divisors
	__device__ static __forceinline__ double div_pos(double x, double y) {
		double r = __builtin_amdgcn_rcp(y);
		double e = __builtin_fma(-y, r, 1.);
		r = __builtin_fma(e, r, r);
		double q = x * r;
		e = __builtin_fma(-y, q, x);
		return __builtin_fma(e, r, q);
	}
	__device__ static __forceinline__ double div(double x, double y) {
		double q = div_pos(x, y);
		double ay = __builtin_fabs(y);
		bool ok = ay > 0x1p-900 && ay < 0x1p+900;
		if (!wave_all(ok)) q = ok ? q : x / y;
		return q;
	}
	// positive, normal arguments (no guard: values outside that range are
	// never used by the callers)
	__device__ static __forceinline__ double sqrt_pos(double x) {
		double y = __builtin_amdgcn_rsq(x);
		double s = x * y, h = .5 * y;
		double r = __builtin_fma(-s, h, .5);
		s = __builtin_fma(s, r, s);
		h = __builtin_fma(h, r, h);
		double d = __builtin_fma(-s, s, x);
		return __builtin_fma(d, h, s);
	}
	// s = sqrt(x) and hh = 1 / (2 sqrt(x)) of a positive, normal x: two
	// Goldschmidt steps from v_rsq_f64 (each squares the relative error)
	__device__ static __forceinline__ void sqrt_and_half_rsqrt(double x, double &s, double &hh) {
		double y = __builtin_amdgcn_rsq(x);
		s = x * y;
		hh = .5 * y;
		double r = __builtin_fma(-s, hh, .5);
		s = __builtin_fma(s, r, s);
		hh = __builtin_fma(hh, r, hh);
		r = __builtin_fma(-s, hh, .5);
		s = __builtin_fma(s, r, s);
		hh = __builtin_fma(hh, r, hh);
	}
	__device__ static __forceinline__ double sqrt_(double x) {
		double y = __builtin_amdgcn_rsq(x);
		double s = x * y, h = .5 * y;
		double r = __builtin_fma(-s, h, .5);
		s = __builtin_fma(s, r, s);
		h = __builtin_fma(h, r, h);
		double d = __builtin_fma(-s, s, x);
		double out = __builtin_fma(d, h, s);
		bool ok = x > 0x1p-900 && x < 0x1p+900;
		if (!wave_all(ok)) out = ok ? out : __builtin_sqrt(x);
		return out;
	}
};

template <>
struct Math<float, false> {
	__device__ static __forceinline__ float rcp(float y) { return 1.f / y; }
	__device__ static __forceinline__ float div(float x, float y) { return x / y; }
	__device__ static __forceinline__ float div_inv(float x, float y, float) { return x / y; }
	__device__ static __forceinline__ float sqrt_(float x) { return __builtin_sqrtf(x); }
	__device__ static __forceinline__ float sqrt_pos(float x) { return __builtin_sqrtf(x); }
	__device__ static __forceinline__ float div_pos(float x, float y) { return x / y; }
};

template <>
struct Math<float, true> {
	__device__ static __forceinline__ float rcp(float y) { return __builtin_amdgcn_rcpf(y); }
	// v_rcp_f32 flushes a denormal divisor to an infinite reciprocal (0 * inf
	// = NaN for x -> 0): such waves take IEEE division, as the FP64 form does
	__device__ static __forceinline__ float div(float x, float y) {
		float q = x * __builtin_amdgcn_rcpf(y);
		float ay = __builtin_fabsf(y);
		bool ok = ay > 0x1p-100f && ay < 0x1p+100f;
		if (!wave_all(ok)) q = ok ? q : x / y;
		return q;
	}
	__device__ static __forceinline__ float div_inv(float x, float, float inv) { return x * inv; }
	__device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
	__device__ static __forceinline__ float sqrt_pos(float x) { return __builtin_amdgcn_sqrtf(x); }
	__device__ static __forceinline__ float div_pos(float x, float y) { return x * __builtin_amdgcn_rcpf(y); }
	__device__ static __forceinline__ void sqrt_and_half_rsqrt(float x, float &s, float &hh) {
		float y = __builtin_amdgcn_rsqf(x);
		s = x * y;
		hh = .5f * y;
	}
};

} // namespace hydro
