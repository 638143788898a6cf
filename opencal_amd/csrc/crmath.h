/* Correctly rounded elementary functions for the SCIARA-fv2 device program:
 * 10^x, atan, cos, x^3 and x^(1/3) as pow(x, 1.0/3.0) computes it (the
 * exponent is the double nearest 1/3), the five calls of
 * OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp:20,61,85,153-154.
 *
 * Each function evaluates its result as a double-double (hi + lo) with a
 * relative error below about 2^-99 and returns hi, the double nearest
 * hi + lo.  The result is therefore the correctly rounded value unless the
 * exact value lies within 2^-99 (relative) of a midpoint between two
 * doubles, which happens with probability about 2^-45 per call.
 *
 * Only IEEE-754 operations that are correctly rounded by definition are
 * used: + - * /, fma (for exact products), floor, frexp and ldexp.  The same
 * source therefore gives the same bits on the x86-64 host (gcc,
 * -ffp-contract=off) and on the GPU (hipcc, -ffp-contract=off).
 *
 * Methods (double-double arithmetic after Dekker, "A floating-point technique
 * for extending the available precision", 1971, and Hida, Li and Bailey,
 * "Algorithms for quad-double precision floating point arithmetic", 2001):
 *   exp   X = k ln2/64 + r with a three-part ln2/64, |r| <= ln2/128,
 *         e^X = 2^(k/64) (1 + expm1(r)) with a 64-entry table of 2^(j/64)
 *         and the Taylor series of expm1 to degree 10;
 *   log   x = 2^e m, m in [1, 2), c = 1 + j/64 nearest m, s = (m-c)/(m+c),
 *         log x = e ln2 + log c + 2 atanh(s), series to s^13;
 *   atan  y = |x| or 1/|x| in [0, 1], c = j/64 nearest y,
 *         atan y = atan c + atan((y-c)/(1+yc)), series to u^15;
 *   cos   w = |t| or pi/2 - |t| (three-part pi/2) in [0, pi/4],
 *         cos or sin of c + d, c = j/64, by the addition formulas with the
 *         Taylor series of sin d and cos d - 1.
 * The tables are generated by gen_crmath_tables.py.
 *
 * Domains where the result is correctly rounded: pow10 for results in the
 * normal range (x >= -307.6); pow_third for x > 0; cube for
 * 2^-340 < |x| < 2^340; atan everywhere; cos for |t| <= pi/2 + 1/128 (NaN
 * beyond, which SCIARA never asks for: its angles come from atan).
 * Subnormal results are rounded twice and may be off by one unit. */
#ifndef OPENCAL_CRMATH_H
#define OPENCAL_CRMATH_H

#include <math.h>

#if defined(__HIPCC__)
#define CR_FN __device__ static inline
#define CR_TABLE __device__ static const
#else
#define CR_FN static inline
#define CR_TABLE static const
#endif

typedef struct {
	double hi, lo;
} cr_dd_t;

#include "crmath_tables.h"

CR_FN cr_dd_t cr_mk(double hi, double lo)
{
	cr_dd_t r;
	r.hi = hi;
	r.lo = lo;
	return r;
}

CR_FN cr_dd_t cr_neg(cr_dd_t a) { return cr_mk(-a.hi, -a.lo); }

/* a + b = s + e exactly (Knuth's two-sum) */
CR_FN cr_dd_t cr_two_sum(double a, double b)
{
	const double s = a + b, bb = s - a;
	return cr_mk(s, (a - (s - bb)) + (b - bb));
}

/* a + b = s + e exactly, for |a| >= |b| or a == 0 (Dekker) */
CR_FN cr_dd_t cr_fast_two_sum(double a, double b)
{
	const double s = a + b;
	return cr_mk(s, b - (s - a));
}

/* a b = p + e exactly (the residual of the rounded product, by fma) */
CR_FN cr_dd_t cr_two_prod(double a, double b)
{
	const double p = a * b;
	return cr_mk(p, fma(a, b, -p));
}

/* a + b, also under cancellation */
CR_FN cr_dd_t cr_add(cr_dd_t a, cr_dd_t b)
{
	cr_dd_t s = cr_two_sum(a.hi, b.hi);
	const cr_dd_t t = cr_two_sum(a.lo, b.lo);
	s.lo += t.hi;
	s = cr_fast_two_sum(s.hi, s.lo);
	s.lo += t.lo;
	return cr_fast_two_sum(s.hi, s.lo);
}

/* a + b without cancellation (same signs, or |b| well below |a|) */
CR_FN cr_dd_t cr_add_fast(cr_dd_t a, cr_dd_t b)
{
	cr_dd_t s = cr_two_sum(a.hi, b.hi);
	s.lo += a.lo + b.lo;
	return cr_fast_two_sum(s.hi, s.lo);
}

CR_FN cr_dd_t cr_mul(cr_dd_t a, cr_dd_t b)
{
	cr_dd_t p = cr_two_prod(a.hi, b.hi);
	p.lo = fma(a.hi, b.lo, fma(a.lo, b.hi, p.lo));
	return cr_fast_two_sum(p.hi, p.lo);
}

CR_FN cr_dd_t cr_mul_d(cr_dd_t a, double b)
{
	cr_dd_t p = cr_two_prod(a.hi, b);
	p.lo = fma(a.lo, b, p.lo);
	return cr_fast_two_sum(p.hi, p.lo);
}

/* a / b: three quotient digits, each from the remainder of the previous */
CR_FN cr_dd_t cr_div(cr_dd_t a, cr_dd_t b)
{
	const double q1 = a.hi / b.hi;
	cr_dd_t r = cr_add(a, cr_mul_d(b, -q1));
	const double q2 = r.hi / b.hi;
	r = cr_add(r, cr_mul_d(b, -q2));
	const double q3 = r.hi / b.hi;
	return cr_add_fast(cr_fast_two_sum(q1, q2), cr_mk(q3, 0.0));
}

/* e^X = 2^m (hi + lo) for X.hi in [-746, 709.8] */
CR_FN cr_dd_t cr_exp_core(cr_dd_t X, int* m)
{
	const double kd = floor(X.hi * CR_INV_LN2_64 + 0.5);
	const int N = (int)kd;
	const int j = N & 63;
	*m = (N - j) / 64;
	/* r = X - kd ln2/64: the first difference is exact (Sterbenz; X.hi when
	 * kd = 0), the rest is below 2^-52 |X| */
	const cr_dd_t p1 = cr_two_prod(kd, CR_LN2_64_1);
	const cr_dd_t p2 = cr_two_prod(kd, CR_LN2_64_2);
	const double r1 = X.hi - p1.hi;
	const double s = ((X.lo - p1.lo) - p2.hi) - (p2.lo + kd * CR_LN2_64_3);
	const cr_dd_t r = cr_two_sum(r1, s);
	/* (e^r - 1)/r = sum r^(k-1)/k!: k = 7..10 in double, k = 1..6 in
	 * double-double */
	const double rh = r.hi;
	double t = 1.0 / 3628800.0;
	t = t * rh + 1.0 / 362880.0;
	t = t * rh + 1.0 / 40320.0;
	t = t * rh + 1.0 / 5040.0;
	cr_dd_t p = cr_add_fast(CR_EXP_C[6], cr_two_prod(rh, t));
	p = cr_add_fast(CR_EXP_C[5], cr_mul(r, p));
	p = cr_add_fast(CR_EXP_C[4], cr_mul(r, p));
	p = cr_add_fast(CR_EXP_C[3], cr_mul(r, p));
	p = cr_add_fast(CR_EXP_C[2], cr_mul(r, p));
	p = cr_add_fast(CR_EXP_C[1], cr_mul(r, p));
	const cr_dd_t em1 = cr_mul(r, p);
	const cr_dd_t T = CR_EXP2_J64[j];
	return cr_add_fast(T, cr_mul(T, em1));
}

/* ln x for finite x > 0 */
CR_FN cr_dd_t cr_log_core(double x)
{
	int e;
	double m = frexp(x, &e); /* [0.5, 1) */
	m = m + m;
	e = e - 1;
	const int j = (int)floor((m - 1.0) * 64.0 + 0.5); /* 0..64 */
	const double c = 1.0 + (double)j * 0.015625;
	const double num = m - c; /* exact (Sterbenz) */
	const cr_dd_t den = cr_two_sum(m, c);
	/* s = num / den: a quotient and the exact remainder's correction */
	const double q1 = num / den.hi;
	const cr_dd_t pq = cr_two_prod(q1, den.hi);
	const double rr = ((num - pq.hi) - pq.lo) - q1 * den.lo;
	const cr_dd_t s = cr_fast_two_sum(q1, rr / den.hi);
	/* 2 atanh(s) = 2 s sum s^(2k)/(2k+1), |s| <= 2^-8: k = 3..6 in double */
	const cr_dd_t v = cr_mul(s, s);
	double t = 1.0 / 13.0;
	t = t * v.hi + 1.0 / 11.0;
	t = t * v.hi + 1.0 / 9.0;
	t = t * v.hi + 1.0 / 7.0;
	cr_dd_t q = cr_add_fast(CR_ODD_INV[2], cr_mul_d(v, t));
	q = cr_add_fast(CR_ODD_INV[1], cr_mul(v, q));
	q = cr_add_fast(CR_ODD_INV[0], cr_mul(v, q));
	cr_dd_t lm = cr_mul(s, q);
	lm.hi = 2.0 * lm.hi;
	lm.lo = 2.0 * lm.lo;
	cr_dd_t E = cr_two_prod((double)e, CR_LN2_HI);
	E.lo = fma((double)e, CR_LN2_LO, E.lo);
	E = cr_fast_two_sum(E.hi, E.lo);
	return cr_add(cr_add(E, CR_LOG_C64[j]), lm);
}

/* 10^x, as pow(10, x) */
CR_FN double cr_pow10(double x)
{
	if (!(x == x)) return x + x;
	if (x > 308.5) return HUGE_VAL;
	if (x < -330.0) return 0.0;
	cr_dd_t X = cr_two_prod(x, CR_LN10_HI);
	X.lo = fma(x, CR_LN10_LO, X.lo);
	X = cr_fast_two_sum(X.hi, X.lo);
	int m;
	const cr_dd_t e = cr_exp_core(X, &m);
	return ldexp(e.hi, m);
}

/* pow(x, 1.0 / 3.0): the exponent is the double nearest 1/3 */
CR_FN double cr_pow_third(double x)
{
	if (!(x > 0.0)) return x == 0.0 ? 0.0 : (x - x) / (x - x);
	if (x == HUGE_VAL) return x;
	const cr_dd_t X = cr_mul_d(cr_log_core(x), 1.0 / 3.0);
	int m;
	const cr_dd_t e = cr_exp_core(X, &m);
	return ldexp(e.hi, m);
}

/* pow(x, 3.0): x^2 exactly, then x^3 = p.hi + p.lo + s.lo x rounded once */
CR_FN double cr_cube(double x)
{
	const double ax = fabs(x);
	if (!(ax > 0x1p-340 && ax < 0x1p340)) return x * x * x;
	const cr_dd_t s = cr_two_prod(x, x);
	const cr_dd_t p = cr_two_prod(s.hi, x);
	return p.hi + (p.lo + s.lo * x);
}

CR_FN double cr_atan(double x)
{
	const double ax = fabs(x);
	if (!(ax == ax)) return x + x;
	if (ax >= 0x1p54) return x < 0.0 ? -CR_PI_2_HI : CR_PI_2_HI;
	if (ax < 0x1p-27) return x; /* x (1 - x^2/3 + ...) rounds to x */
	/* y = 1/ax as a double-double: the residual 1 - q ax is exact */
	const int inv = ax > 1.0;
	const double q0 = 1.0 / ax;
	const double r0 = fma(-q0, ax, 1.0);
	const cr_dd_t y = inv ? cr_fast_two_sum(q0, r0 / ax) : cr_mk(ax, 0.0);
	const int j = (int)floor(y.hi * 64.0 + 0.5); /* 0..64 */
	const double c = (double)j * 0.015625;
	const cr_dd_t num = cr_two_sum(y.hi - c, y.lo); /* y.hi - c exact */
	cr_dd_t den = cr_two_prod(y.hi, c);
	den.lo = fma(y.lo, c, den.lo);
	den = cr_add_fast(cr_mk(1.0, 0.0), den);
	const cr_dd_t u = cr_div(num, den); /* |u| <= 2^-7 */
	const cr_dd_t v = cr_mul(u, u);
	/* atan u = u sum (-v)^k/(2k+1): k = 4..7 in double */
	double t = -1.0 / 15.0;
	t = t * v.hi + 1.0 / 13.0;
	t = t * v.hi - 1.0 / 11.0;
	t = t * v.hi + 1.0 / 9.0;
	cr_dd_t q = cr_add_fast(cr_neg(CR_ODD_INV[3]), cr_mul_d(v, t));
	q = cr_add_fast(CR_ODD_INV[2], cr_mul(v, q));
	q = cr_add_fast(cr_neg(CR_ODD_INV[1]), cr_mul(v, q));
	q = cr_add_fast(CR_ODD_INV[0], cr_mul(v, q));
	cr_dd_t res = cr_add(CR_ATAN_J64[j], cr_mul(u, q));
	if (inv) res = cr_add(cr_mk(CR_PI_2_HI, CR_PI_2_MID), cr_neg(res));
	return x < 0.0 ? -res.hi : res.hi;
}

CR_FN double cr_cos(double t)
{
	const double a = fabs(t);
	if (!(a <= CR_PI_2_HI + 0.0078125)) return (t - t) / (t - t);
	if (a < 0x1p-27) return 1.0; /* 1 - a^2/2 rounds to 1 */
	/* cos a = sin(pi/2 - a) above pi/4; pi/2 - a with a three-part pi/2
	 * (the first difference is exact) */
	const int sine = a > CR_PI_4;
	const cr_dd_t w = sine ? cr_add(cr_two_sum(CR_PI_2_HI - a, CR_PI_2_MID), cr_mk(CR_PI_2_LO, 0.0))
	                       : cr_mk(a, 0.0);
	const int j = (int)floor(w.hi * 64.0 + 0.5); /* 0..51 */
	const double c = (double)j * 0.015625;
	const cr_dd_t d = cr_two_sum(w.hi - c, w.lo); /* |d| <= 2^-7, w.hi - c exact */
	const cr_dd_t v = cr_mul(d, d);
	/* sin d = d sum (-v)^k/(2k+1)!: k = 3..5 in double */
	double ts = -1.0 / 39916800.0;
	ts = ts * v.hi + 1.0 / 362880.0;
	ts = ts * v.hi - 1.0 / 5040.0;
	cr_dd_t qs = cr_add_fast(CR_SIN_C[2], cr_mul_d(v, ts));
	qs = cr_add_fast(cr_neg(CR_SIN_C[1]), cr_mul(v, qs));
	qs = cr_add_fast(CR_SIN_C[0], cr_mul(v, qs));
	const cr_dd_t S = cr_mul(d, qs);
	/* cos d - 1 = v sum_{k>=1} (-v)^(k-1) (-1)/(2k)!: k = 4..5 in double */
	double tc = -1.0 / 3628800.0;
	tc = tc * v.hi + 1.0 / 40320.0;
	cr_dd_t qc = cr_add_fast(cr_neg(CR_COS_C[3]), cr_mul_d(v, tc));
	qc = cr_add_fast(CR_COS_C[2], cr_mul(v, qc));
	qc = cr_add_fast(cr_neg(CR_COS_C[1]), cr_mul(v, qc));
	const cr_dd_t CM = cr_mul(v, qc);
	/* cos(c + d) = cos c + (cos c (cos d - 1) - sin c sin d)
	 * sin(c + d) = sin c + (sin c (cos d - 1) + cos c sin d) */
	const cr_dd_t A = sine ? CR_SIN_J64[j] : CR_COS_J64[j];
	const cr_dd_t B = sine ? CR_COS_J64[j] : cr_neg(CR_SIN_J64[j]);
	const cr_dd_t res = cr_add(A, cr_add(cr_mul(A, CM), cr_mul(B, S)));
	return res.hi;
}

#endif
