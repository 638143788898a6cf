/* Elementary functions used by the SCIARA-fv2 device program, written with
 * IEEE +, -, *, / (no fused multiply-add; the library and the oracle are
 * built with -ffp-contract=off), frexp and ldexp (both exact), so they give
 * the same bits on the x86-64 host and on the GPU.  Accuracy is a few ulp;
 * they are not glibc's routines, and results that involve them differ from a
 * glibc build by that much (SCIARA-fv2 calls pow, atan and cos,
 * OpenCAL/examples/cal_sciara-fv2/source/Sciara.cpp:18-21,61,85,153-154).
 *
 * Methods: exp by Cody-Waite reduction with a two-part ln 2 and a degree-13
 * Taylor polynomial on |r| <= ln2/2; log by mantissa reduction to
 * [sqrt(1/2), sqrt(2)) and the atanh series in s = (m-1)/(m+1); atan by
 * reciprocal and pi/6 reductions to |y| <= tan(pi/12) and its Taylor series;
 * cos by the Taylor series of cos on [0, pi/4] and of sin on the complement. */
#ifndef OPENCAL_PORTABLE_MATH_H
#define OPENCAL_PORTABLE_MATH_H

#include <math.h>

#if defined(__HIPCC__)
#define PM_FN __host__ __device__ static inline
#else
#define PM_FN static inline
#endif

#define PM_LN2_HI 6.93147180369123816490e-01 /* leading bits of ln 2 (k * hi exact) */
#define PM_LN2_LO 1.90821492927058770002e-10 /* ln 2 - PM_LN2_HI */
#define PM_INV_LN2 1.44269504088896338700e+00
#define PM_LN10 2.30258509299404568402e+00
#define PM_PI_2 1.57079632679489661923e+00
#define PM_PI_2_HI 1.57079632679489655800e+00 /* pi/2 rounded to double */
#define PM_PI_2_LO 6.12323399573676588613e-17 /* pi/2 - PM_PI_2_HI */
#define PM_PI_6 5.23598775598298873077e-01
#define PM_SQRT3 1.73205080756887729353e+00
#define PM_TAN_PI_12 2.67949192431122706473e-01
#define PM_SQRT2 1.41421356237309504880e+00

/* exp(x + xl) for |xl| << ulp(x) */
PM_FN double pm_exp2(double x, double xl)
{
	double kd, hi, lo, r, p;
	long k;
	if (x > 709.0) return HUGE_VAL;
	if (x < -745.0) return 0.0;
	kd = x * PM_INV_LN2;
	kd = kd >= 0.0 ? kd + 0.5 : kd - 0.5;
	k = (long)kd;
	hi = x - (double)k * PM_LN2_HI;
	lo = (double)k * PM_LN2_LO;
	r = (hi - lo) + xl;
	/* sum_{n=0}^{13} r^n / n! */
	p = 1.0 / 6227020800.0;
	p = p * r + 1.0 / 479001600.0;
	p = p * r + 1.0 / 39916800.0;
	p = p * r + 1.0 / 3628800.0;
	p = p * r + 1.0 / 362880.0;
	p = p * r + 1.0 / 40320.0;
	p = p * r + 1.0 / 5040.0;
	p = p * r + 1.0 / 720.0;
	p = p * r + 1.0 / 120.0;
	p = p * r + 1.0 / 24.0;
	p = p * r + 1.0 / 6.0;
	p = p * r + 0.5;
	p = p * r + 1.0;
	p = p * r + 1.0;
	return ldexp(p, (int)k);
}

PM_FN double pm_exp(double x) { return pm_exp2(x, 0.0); }

/* a * b = hi + lo exactly (Dekker's product with Veltkamp splitting) */
PM_FN void pm_two_prod(double a, double b, double* hi, double* lo)
{
	const double split = 134217729.0; /* 2^27 + 1 */
	double p = a * b;
	double ta = split * a, a1 = ta - (ta - a), a2 = a - a1;
	double tb = split * b, b1 = tb - (tb - b), b2 = b - b1;
	*hi = p;
	*lo = ((a1 * b1 - p) + a1 * b2 + a2 * b1) + a2 * b2;
}

PM_FN double pm_log(double x)
{
	int e;
	double m, s, s2, t;
	if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);
	m = frexp(x, &e); /* x = m * 2^e, m in [0.5, 1) */
	m = m * 2.0;
	e = e - 1;
	if (m > PM_SQRT2) {
		m = m * 0.5;
		e = e + 1;
	}
	s = (m - 1.0) / (m + 1.0);
	s2 = s * s;
	/* 2 s (1 + s2/3 + s2^2/5 + ... + s2^11/23) */
	t = 1.0 / 23.0;
	t = t * s2 + 1.0 / 21.0;
	t = t * s2 + 1.0 / 19.0;
	t = t * s2 + 1.0 / 17.0;
	t = t * s2 + 1.0 / 15.0;
	t = t * s2 + 1.0 / 13.0;
	t = t * s2 + 1.0 / 11.0;
	t = t * s2 + 1.0 / 9.0;
	t = t * s2 + 1.0 / 7.0;
	t = t * s2 + 1.0 / 5.0;
	t = t * s2 + 1.0 / 3.0;
	t = t * s2 + 1.0;
	return (double)e * PM_LN2_HI + (2.0 * s * t + (double)e * PM_LN2_LO);
}

#define PM_LN10_LO -2.17075622338224935e-16 /* ln 10 - PM_LN10 */

/* 10^x = exp(x ln 10) with x ln 10 carried as an unevaluated sum */
PM_FN double pm_pow10(double x)
{
	double hi, lo;
	pm_two_prod(x, PM_LN10, &hi, &lo);
	return pm_exp2(hi, lo + x * PM_LN10_LO);
}

/* x^(1/3) for x > 0, evaluated as exp(log(x) * (1/3)) */
PM_FN double pm_cbrt_pos(double x) { return pm_exp(pm_log(x) * (1.0 / 3.0)); }

/* atan and cos below are branch-free (selects only), so that independent
 * calls interleave on the GPU; each select picks a value computed exactly as
 * a branch would compute it. */
PM_FN double pm_atan(double x)
{
	const double ax = x < 0.0 ? -x : x;
	const int inv = ax > 1.0;
	const double t = inv ? 1.0 / ax : ax;
	const int red = t > PM_TAN_PI_12;
	const double yr = (t * PM_SQRT3 - 1.0) / (t + PM_SQRT3);
	const double y = red ? yr : t;
	const double off = red ? PM_PI_6 : 0.0;
	const double y2 = y * y;
	double p, res;
	/* y - y^3/3 + y^5/5 - ... - y^27/27 */
	p = -1.0 / 27.0;
	p = p * y2 + 1.0 / 25.0;
	p = p * y2 - 1.0 / 23.0;
	p = p * y2 + 1.0 / 21.0;
	p = p * y2 - 1.0 / 19.0;
	p = p * y2 + 1.0 / 17.0;
	p = p * y2 - 1.0 / 15.0;
	p = p * y2 + 1.0 / 13.0;
	p = p * y2 - 1.0 / 11.0;
	p = p * y2 + 1.0 / 9.0;
	p = p * y2 - 1.0 / 7.0;
	p = p * y2 + 1.0 / 5.0;
	p = p * y2 - 1.0 / 3.0;
	p = p * y2 + 1.0;
	res = off + y * p;
	res = inv ? PM_PI_2 - res : res;
	return x < 0.0 ? -res : res;
}

/* cos(t) for |t| <= pi/2: Taylor series of cos on [0, pi/4], of sin of the
 * complement above; both are evaluated and one is selected */
PM_FN double pm_cos(double t)
{
	const double a = t < 0.0 ? -t : t;
	const double b = (PM_PI_2_HI - a) + PM_PI_2_LO; /* the subtraction is exact */
	const double a2 = a * a, b2 = b * b;
	double p, q;
	/* 1 - a^2/2! + ... + a^20/20! */
	p = 1.0 / 2432902008176640000.0;
	p = -p * a2 + 1.0 / 6402373705728000.0;
	p = -p * a2 + 1.0 / 20922789888000.0;
	p = -p * a2 + 1.0 / 87178291200.0;
	p = -p * a2 + 1.0 / 479001600.0;
	p = -p * a2 + 1.0 / 3628800.0;
	p = -p * a2 + 1.0 / 40320.0;
	p = -p * a2 + 1.0 / 720.0;
	p = -p * a2 + 1.0 / 24.0;
	p = -p * a2 + 0.5;
	/* b - b^3/3! + ... + b^21/21! */
	q = 1.0 / 51090942171709440000.0;
	q = -q * b2 + 1.0 / 121645100408832000.0;
	q = -q * b2 + 1.0 / 355687428096000.0;
	q = -q * b2 + 1.0 / 1307674368000.0;
	q = -q * b2 + 1.0 / 6227020800.0;
	q = -q * b2 + 1.0 / 39916800.0;
	q = -q * b2 + 1.0 / 362880.0;
	q = -q * b2 + 1.0 / 5040.0;
	q = -q * b2 + 1.0 / 120.0;
	q = -q * b2 + 1.0 / 6.0;
	return a <= 0.78539816339744830962 ? 1.0 - p * a2 : b - b * (q * b2);
}

#endif
