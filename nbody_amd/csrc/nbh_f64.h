// nbh_f64.h -- binary64 square root and quotient for operand ranges where
// the general sequences' scaling steps are identities (shared by the FP64
// direct-sum and neighbour-scheme kernels).
#pragma once

#include <hip/hip_runtime.h>

// Correctly rounded binary64 square root and quotient without the range steps
// of the general sequences.  The compiler's sqrt is rsq + two Goldschmidt
// refinements + two Newton corrections, wrapped in a scaling by 2^+-256 for
// inputs below 2^-767 and a pass-through for 0 and inf; its division is rcp +
// two Newton refinements of the reciprocal, the quotient and one fma
// correction, wrapped in v_div_scale / v_div_fmas / v_div_fixup, which scale
// operands whose exponents are near the ends of the range and handle 0, inf and
// NaN.  For the operands of the pair term -- r2 in [1e-8, 1.3e13] (coordinates of
// magnitude <= 1e6) and numerators -m1 m2 with masses of magnitude in [1e-60,
// 1e60] -- those wrappers are identities, so the bare sequences below give the
// same bits (checked against the CPU restatement by the tests).  Tiles holding
// any other operand take the general sequences.
__device__ inline double sqrt_rn(double x)
{
	const double y = __builtin_amdgcn_rsq(x);
	double g = x * y;
	double h = y * 0.5;
	const double r = fma(-h, g, 0.5);
	g = fma(g, r, g);
	double d = fma(-g, g, x);
	h = fma(h, r, h);
	g = fma(d, h, g);
	d = fma(-g, g, x);
	return fma(d, h, g);
}

__device__ inline double div_rn(double a, double b)
{
	double r = __builtin_amdgcn_rcp(b);
	double e = fma(-b, r, 1.0);
	r = fma(r, e, r);
	e = fma(-b, r, 1.0);
	r = fma(r, e, r);
	const double q = a * r;
	const double rem = fma(-b, q, a);
	return fma(rem, r, q);
}

__device__ inline bool coord_in_range(double v) { return fabs(v) <= 1e6; }
__device__ inline bool mass_in_range(double m) { return fabs(m) >= 1e-60 && fabs(m) <= 1e60; }

