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

/* ---- first evaluations with Ziv's rounding test ----
 *
 * crz_pow10, crz_atan, crz_cos and crz_pow_third return the same double as
 * cr_pow10, cr_atan, cr_cos and cr_pow_third, faster in nearly every case.
 * Each first evaluates its function as a double-double hi + lo with a
 * relative error below about 2^-73 (10^x, x^(1/3)) and 2^-78 (atan, cos): the same argument reductions and
 * tables as above, the leading terms of each series as exact products
 * (fma) and the rest in double.  When every value within 2^-70 (relative)
 * of hi + lo rounds to hi, hi is the correctly rounded result (Ziv, "Fast
 * evaluation of elementary mathematical functions with correctly rounded
 * last bit", ACM TOMS 17(3), 1991); otherwise, about once in 2^17 calls,
 * the evaluation above decides.  tests/test_crmath.py compares them with
 * the full evaluations over seeded samples of each domain. */

/* hi = RN(hi + lo); 1 when every value within e of hi + lo rounds to hi */
CR_FN int crz_rounds(double hi, double lo, double e) { return hi + (lo + e) == hi + (lo - e); }

#define CRZ_BOUND 0x1p-70
#define CRZ_THIRD_HI 0x1.5555555555555p-2   /* 1/3 = HI + LO */
#define CRZ_THIRD_LO 0x1.5555555555555p-56
#define CRZ_SIXTH_HI 0x1.5555555555555p-3   /* 1/6 = HI + LO */
#define CRZ_SIXTH_LO 0x1.5555555555555p-57

/* a (hi + lo) for a double-double and a constant khi + klo */
CR_FN cr_dd_t crz_mul_k(cr_dd_t a, double khi, double klo)
{
	cr_dd_t p = cr_two_prod(a.hi, khi);
	p.lo = fma(a.hi, klo, fma(a.lo, khi, p.lo));
	return p;
}

/* e^X = 2^m (hi + lo) for |X.hi| <= 700, relative error below 2^-73 */
CR_FN cr_dd_t crz_exp_first(cr_dd_t X, int* mout)
{
	/* X = k ln2/64 + r as in cr_exp_core */
	const double kd = floor(X.hi * CR_INV_LN2_64 + 0.5);
	const int N = (int)kd;
	const int j = N & 63;
	const int m = (N - j) / 64;
	const cr_dd_t p1 = cr_two_prod(kd, CR_LN2_64_1);
	const cr_dd_t p2 = cr_two_prod(kd, CR_LN2_64_2);
	const double r1 = X.hi - p1.hi;
	const double s = ((X.lo - p1.lo) - p2.hi) - (p2.lo + kd * CR_LN2_64_3);
	const cr_dd_t r = cr_two_sum(r1, s);
	/* e^r - 1 = r + r^2/2 + r^3 (1/6 + r/24 + ... + r^5/8!), |r| <= 2^-7.4;
	 * r.hi^2 exactly, r.lo through the square's cross term */
	const cr_dd_t q = cr_two_prod(r.hi, r.hi);
	double t = 1.0 / 40320.0;
	t = t * r.hi + 1.0 / 5040.0;
	t = t * r.hi + 1.0 / 720.0;
	t = t * r.hi + 1.0 / 120.0;
	t = t * r.hi + 1.0 / 24.0;
	t = t * r.hi + 1.0 / 6.0;
	const double tail3 = q.hi * (r.hi * t);
	const cr_dd_t e0 = cr_two_sum(r.hi, 0.5 * q.hi);
	const cr_dd_t em1 = cr_fast_two_sum(e0.hi, e0.lo + (r.lo + (0.5 * q.lo + r.hi * r.lo) + tail3));
	/* 2^(j/64) (1 + em1) */
	const cr_dd_t T = CR_EXP2_J64[j];
	cr_dd_t p = cr_two_prod(T.hi, em1.hi);
	p.lo = fma(T.hi, em1.lo, fma(T.lo, em1.hi, p.lo));
	const cr_dd_t s2 = cr_two_sum(T.hi, p.hi);
	*mout = m;
	return cr_fast_two_sum(s2.hi, s2.lo + (T.lo + p.lo));
}

CR_FN int crz_pow10_first(double x, double* out)
{
	if (!(x >= -300.0 && x <= 300.0)) return 0;
	cr_dd_t X = cr_two_prod(x, CR_LN10_HI);
	X.lo = fma(x, CR_LN10_LO, X.lo);
	X = cr_fast_two_sum(X.hi, X.lo);
	int m;
	const cr_dd_t res = crz_exp_first(X, &m);
	if (!crz_rounds(res.hi, res.lo, res.hi * CRZ_BOUND)) return 0;
	*out = ldexp(res.hi, m);
	return 1;
}

/* pow(x, 1.0 / 3.0) for 2^-1000 <= x <= 2^1000: ln x as cr_log_core reduces
 * it, with 2 atanh(s) = 2 (s + s^3/3 + s^5 (1/5 + ... + s^8/13)), |s| <=
 * 2^-8, s^3/3 from exact products; then e^(ln x RN(1/3)), relative error
 * below 2^-73 */
CR_FN int crz_pow_third_first(double x, double* out)
{
	if (!(x >= 0x1p-1000 && x <= 0x1p1000)) return 0;
	int e;
	double m = frexp(x, &e); /* [0.5, 1) */
	m = m + m;
	e = e - 1;
	const int j = (int)floor((m - 1.0) * 64.0 + 0.5); /* 0..64 */
	const double c = 1.0 + (double)j * 0.015625;
	const double num = m - c; /* exact (Sterbenz) */
	const cr_dd_t den = cr_two_sum(m, c);
	const double rd = 1.0 / den.hi;
	const double sh = num * rd;
	const double sl = (fma(-sh, den.hi, num) - sh * den.lo) * rd;
	const cr_dd_t v = cr_two_prod(sh, sh);
	cr_dd_t w = cr_two_prod(sh, v.hi);
	w.lo = fma(sh, v.lo, w.lo);
	const cr_dd_t c3 = crz_mul_k(w, CRZ_THIRD_HI, CRZ_THIRD_LO);
	double t = 1.0 / 13.0;
	t = t * v.hi + 1.0 / 11.0;
	t = t * v.hi + 1.0 / 9.0;
	t = t * v.hi + 1.0 / 7.0;
	t = t * v.hi + 1.0 / 5.0;
	const double h5 = w.hi * (v.hi * t);
	const cr_dd_t a0 = cr_two_sum(sh, c3.hi);
	const cr_dd_t at = cr_fast_two_sum(a0.hi, a0.lo + ((sl + v.hi * sl) + (c3.lo + h5)));
	cr_dd_t E = cr_two_prod((double)e, CR_LN2_HI);
	E.lo = fma((double)e, CR_LN2_LO, E.lo);
	E = cr_fast_two_sum(E.hi, E.lo);
	const cr_dd_t L = cr_add(cr_add(E, CR_LOG_C64[j]), cr_mk(2.0 * at.hi, 2.0 * at.lo));
	int mm;
	const cr_dd_t res = crz_exp_first(cr_mul_d(L, 1.0 / 3.0), &mm);
	if (!crz_rounds(res.hi, res.lo, res.hi * CRZ_BOUND)) return 0;
	*out = ldexp(res.hi, mm);
	return 1;
}

CR_FN int crz_atan_first(double x, double* out)
{
	const double ax = fabs(x);
	if (!(ax >= 0x1p-27 && ax < 0x1p54)) return 0;
	const int inv = ax > 1.0;
	const double q0 = 1.0 / ax;
	const double r0 = fma(-q0, ax, 1.0);
	const cr_dd_t y = inv ? cr_fast_two_sum(q0, r0 * q0) : cr_mk(ax, 0.0);
	const int j = (int)floor(y.hi * 64.0 + 0.5);
	const double c = (double)j * 0.015625;
	const cr_dd_t num = cr_two_sum(y.hi - c, y.lo);
	cr_dd_t den = cr_two_prod(y.hi, c);
	den.lo = fma(y.lo, c, den.lo);
	den = cr_add_fast(cr_mk(1.0, 0.0), den);
	/* u = num / den = uh + ul from one reciprocal and the remainder */
	const double rd = 1.0 / den.hi;
	const double uh = num.hi * rd;
	const double ul = ((fma(-uh, den.hi, num.hi) + num.lo) - uh * den.lo) * rd;
	/* atan u = u - u^3/3 + u^5 (1/5 - u^2/7 + ... + u^8/13), |u| <= 2^-7:
	 * uh^3/3 from exact products, ul through the cubic's derivative */
	const cr_dd_t v = cr_two_prod(uh, uh);
	cr_dd_t w = cr_two_prod(uh, v.hi);
	w.lo = fma(uh, v.lo, w.lo);
	const cr_dd_t c3 = crz_mul_k(w, -CRZ_THIRD_HI, -CRZ_THIRD_LO);
	double t = 1.0 / 13.0;
	t = t * v.hi - 1.0 / 11.0;
	t = t * v.hi + 1.0 / 9.0;
	t = t * v.hi - 1.0 / 7.0;
	t = t * v.hi + 1.0 / 5.0;
	const double h5 = w.hi * (v.hi * t);
	const cr_dd_t T = CR_ATAN_J64[j];
	const cr_dd_t s1 = cr_two_sum(T.hi, uh);
	const cr_dd_t s2 = cr_two_sum(s1.hi, c3.hi);
	const double lo = (s1.lo + s2.lo) + (T.lo + ((ul - v.hi * ul) + (c3.lo + h5)));
	cr_dd_t res = cr_fast_two_sum(s2.hi, lo);
	if (inv) res = cr_add(cr_mk(CR_PI_2_HI, CR_PI_2_MID), cr_neg(res));
	if (!crz_rounds(res.hi, res.lo, res.hi * CRZ_BOUND)) return 0;
	*out = x < 0.0 ? -res.hi : res.hi;
	return 1;
}

CR_FN int crz_cos_first(double t, double* out)
{
	const double a = fabs(t);
	if (!(a >= 0x1p-27 && a <= CR_PI_2_HI + 0.0078125)) return 0;
	const int sine = a > CR_PI_4;
	const cr_dd_t w = sine ? cr_add(cr_two_sum(CR_PI_2_HI - a, CR_PI_2_MID), cr_mk(CR_PI_2_LO, 0.0))
	                       : cr_mk(a, 0.0);
	const int j = (int)floor(w.hi * 64.0 + 0.5);
	const double c = (double)j * 0.015625;
	const cr_dd_t d = cr_two_sum(w.hi - c, w.lo); /* |d| <= 2^-7 */
	const cr_dd_t v = cr_two_prod(d.hi, d.hi);
	/* sin d = d - d^3/6 + d^5 (1/5! - d^2/7! + d^4/9!): d.hi^3/6 from
	 * exact products, d.lo through the derivative 1 - d^2/2 */
	cr_dd_t d3 = cr_two_prod(d.hi, v.hi);
	d3.lo = fma(d.hi, v.lo, d3.lo);
	const cr_dd_t c3 = crz_mul_k(d3, -CRZ_SIXTH_HI, -CRZ_SIXTH_LO);
	double ts = 1.0 / 362880.0;
	ts = ts * v.hi - 1.0 / 5040.0;
	ts = ts * v.hi + 1.0 / 120.0;
	const double h5 = d3.hi * (v.hi * ts);
	const cr_dd_t s0 = cr_two_sum(d.hi, c3.hi);
	const cr_dd_t S = cr_fast_two_sum(s0.hi, s0.lo + ((d.lo - 0.5 * v.hi * d.lo) + (c3.lo + h5)));
	/* cos d - 1 = -d^2/2 + d^4 (1/4! - d^2/6! + d^4/8!) */
	double tc = 1.0 / 40320.0;
	tc = tc * v.hi - 1.0 / 720.0;
	tc = tc * v.hi + 1.0 / 24.0;
	const double h4 = v.hi * (v.hi * tc);
	const cr_dd_t CM = cr_fast_two_sum(-0.5 * v.hi, (-0.5 * v.lo - d.hi * d.lo) + h4);
	/* cos(c + d) = cos c + (cos c (cos d - 1) - sin c sin d)
	 * sin(c + d) = sin c + (sin c (cos d - 1) + cos c sin d) */
	const cr_dd_t A = sine ? CR_SIN_J64[j] : CR_COS_J64[j];
	const cr_dd_t B = sine ? CR_COS_J64[j] : cr_neg(CR_SIN_J64[j]);
	cr_dd_t P1 = cr_two_prod(A.hi, CM.hi);
	P1.lo = fma(A.hi, CM.lo, fma(A.lo, CM.hi, P1.lo));
	cr_dd_t P2 = cr_two_prod(B.hi, S.hi);
	P2.lo = fma(B.hi, S.lo, fma(B.lo, S.hi, P2.lo));
	const cr_dd_t s1 = cr_two_sum(A.hi, P2.hi);
	const cr_dd_t s2 = cr_two_sum(s1.hi, P1.hi);
	const cr_dd_t res = cr_fast_two_sum(s2.hi, (s1.lo + s2.lo) + (A.lo + (P1.lo + P2.lo)));
	if (!crz_rounds(res.hi, res.lo, res.hi * CRZ_BOUND)) return 0;
	*out = res.hi;
	return 1;
}

CR_FN double crz_pow10(double x)
{
	double r;
	return crz_pow10_first(x, &r) ? r : cr_pow10(x);
}

CR_FN double crz_atan(double x)
{
	double r;
	return crz_atan_first(x, &r) ? r : cr_atan(x);
}

CR_FN double crz_cos(double t)
{
	double r;
	return crz_cos_first(t, &r) ? r : cr_cos(t);
}

CR_FN double crz_pow_third(double x)
{
	double r;
	return crz_pow_third_first(x, &r) ? r : cr_pow_third(x);
}

#endif
