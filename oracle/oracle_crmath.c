/* ORACLE — test infrastructure, not product code.
 *
 * Checks of the correctly rounded functions of opencal_amd/csrc/crmath.h
 * (the device's 10^x, atan, cos, x^3 and pow(x, 1/3), standing in for the
 * glibc calls of OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp:20,
 * 61,85,153-154):
 *   orc_cr_eval    the functions elementwise (the same source the GPU runs);
 *   orc_quad_eval  the same functions in binary128 (libquadmath: 113-bit
 *                  significand), rounded once to double: the correctly
 *                  rounded value except within about 2^-57 (relative) of a
 *                  rounding midpoint;
 *   orc_libm_eval  the host libm calls exactly as the reference writes them;
 *   orc_cr_check   counts, over n seeded points of a domain, where crmath
 *                  differs from the binary128 result and where it differs
 *                  from the host libm, with the largest difference in ulp.
 * fn: 0 pow(10, x), 1 atan, 2 cos, 3 pow(x, 3.0), 4 pow(x, 1.0/3.0). */
#include "oracle.h"
#include "../opencal_amd/csrc/crmath.h"
#include <math.h>
#include <quadmath.h>
#include <stdint.h>
#include <string.h>

static double cr_fn(int fn, double x)
{
	switch (fn) {
	case 0: return cr_pow10(x);
	case 1: return cr_atan(x);
	case 2: return cr_cos(x);
	case 3: return cr_cube(x);
	case 4: return cr_pow_third(x);
	}
	return NAN;
}

static double quad_fn(int fn, double x)
{
	__float128 q = (__float128)x;
	switch (fn) {
	case 0: return (double)powq((__float128)10.0, q);
	case 1: return (double)atanq(q);
	case 2: return (double)cosq(q);
	case 3: return (double)(q * q * q);   /* q*q exact in 113 bits */
	case 4: return (double)powq(q, (__float128)(1.0 / 3.0));
	}
	return NAN;
}

static double libm_fn(int fn, double x)
{
	switch (fn) {
	case 0: return pow(10, x);
	case 1: return atan(x);
	case 2: return cos(x);
	case 3: return pow(x, 3.0);
	case 4: return pow(x, 1.0 / 3.0);
	}
	return NAN;
}

int orc_cr_eval(int fn, const double* x, int n, double* y)
{
	int k;
	if (fn < 0 || fn > 4) return -1;
	for (k = 0; k < n; k++) y[k] = cr_fn(fn, x[k]);
	return 0;
}

int orc_quad_eval(int fn, const double* x, int n, double* y)
{
	int k;
	if (fn < 0 || fn > 4) return -1;
	for (k = 0; k < n; k++) y[k] = quad_fn(fn, x[k]);
	return 0;
}

int orc_libm_eval(int fn, const double* x, int n, double* y)
{
	int k;
	if (fn < 0 || fn > 4) return -1;
	for (k = 0; k < n; k++) y[k] = libm_fn(fn, x[k]);
	return 0;
}

static uint64_t splitmix(uint64_t* s)
{
	uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static double ulp_diff(double a, double b)
{
	if (a == b) return 0.0;
	if (!isfinite(a) || !isfinite(b)) return INFINITY;
	return fabs(a - b) / (nextafter(fabs(b), INFINITY) - fabs(b));
}

/* Point k of n: uniform in [lo, hi] (log_scale 0) or log-uniform over
 * [lo, hi] with lo > 0 (log_scale 1), from SplitMix64 seeded by seed + k. */
static double point(double lo, double hi, int log_scale, uint64_t seed, long k)
{
	uint64_t s = seed + (uint64_t)k * 0x632BE59BD9B4E019ull;
	const double u = (double)(splitmix(&s) >> 11) * 0x1p-53;
	return log_scale ? exp(log(lo) + u * (log(hi) - log(lo))) : lo + u * (hi - lo);
}

/* out[0] points where crmath != binary128, out[1] points where crmath !=
 * libm, out[2] the largest |crmath - libm| in ulp of the libm value, out[3]
 * the largest |crmath - binary128| in ulp, out[4] points where libm !=
 * binary128. */
int orc_cr_check(int fn, double lo, double hi, int log_scale, long n, unsigned long long seed, double* out)
{
	long k, bad_q = 0, bad_l = 0, bad_lq = 0;
	double max_l = 0.0, max_q = 0.0;
	if (fn < 0 || fn > 4) return -1;
#pragma omp parallel for reduction(+ : bad_q, bad_l, bad_lq) reduction(max : max_l, max_q) schedule(static)
	for (k = 0; k < n; k++) {
		const double x = point(lo, hi, log_scale, seed, k);
		const double c = cr_fn(fn, x), q = quad_fn(fn, x), l = libm_fn(fn, x);
		if (c != q) {
			bad_q++;
			if (ulp_diff(c, q) > max_q) max_q = ulp_diff(c, q);
		}
		if (c != l) {
			bad_l++;
			if (ulp_diff(c, l) > max_l) max_l = ulp_diff(c, l);
		}
		if (l != q) bad_lq++;
	}
	out[0] = (double)bad_q;
	out[1] = (double)bad_l;
	out[2] = max_l;
	out[3] = max_q;
	out[4] = (double)bad_lq;
	return 0;
}

/* The first evaluations with Ziv's rounding test (crz_pow10, crz_atan,
 * crz_cos, crz_pow_third of crmath.h, which the GPU runs) against the full
 * evaluations (cr_pow10, cr_atan, cr_cos, cr_pow_third), over n seeded
 * points of a domain: out[0] points where the two results differ (0
 * expected), out[1] points where the first evaluation left the result to
 * the full one.  fn: 0 pow(10, x), 1 atan, 2 cos, 4 pow(x, 1.0/3.0). */
int orc_crz_check(int fn, double lo, double hi, int log_scale, long n, unsigned long long seed, double* out)
{
	long k, bad = 0, slow = 0;
	if (fn < 0 || fn > 4 || fn == 3) return -1;
#pragma omp parallel for reduction(+ : bad, slow) schedule(static)
	for (k = 0; k < n; k++) {
		const double x = point(lo, hi, log_scale, seed, k);
		double f = 0.0;
		int ok = 0;
		double z;
		switch (fn) {
		case 0: ok = crz_pow10_first(x, &f); z = crz_pow10(x); break;
		case 1: ok = crz_atan_first(x, &f); z = crz_atan(x); break;
		case 2: ok = crz_cos_first(x, &f); z = crz_cos(x); break;
		default: ok = crz_pow_third_first(x, &f); z = crz_pow_third(x); break;
		}
		const double c = cr_fn(fn, x);
		if (!ok) slow++;
		/* bit patterns: a NaN result must be the same NaN */
		if (memcmp(&z, &c, sizeof z) != 0 || (ok && memcmp(&f, &c, sizeof f) != 0)) bad++;
	}
	out[0] = (double)bad;
	out[1] = (double)slow;
	return 0;
}
