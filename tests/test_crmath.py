"""The correctly rounded functions of opencal_amd/csrc/crmath.h (the GPU's
10^x, atan, cos, x^3 and pow(x, 1/3) of SCIARA-fv2, standing in for glibc's
pow, atan and cos at OpenCAL-OMP/examples/calomp_sciara-fv2/source/
Sciara.cpp:20,61,85,153-154), on the host from the same source.

Pins: (1) equality with binary128 arithmetic (libquadmath, 113-bit
significand) rounded once, on seeded samples of each domain: the functions
return the correctly rounded double; (2) every point where they differ from
the host glibc is a point where glibc differs from binary128, i.e. glibc's
own misrounding; (3) the committed 10^7-point table
(tests/golden/crmath_pin_1e7.txt, made by make_crmath_pin.py) has no
binary128 mismatch in any row; (4) special values and exact cases."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import native

HERE = os.path.dirname(os.path.abspath(__file__))

DOMAINS = [
    (0, -1.2, 2.0, 0), (0, -300.0, 300.0, 0),
    (1, 0.0, 100.0, 0), (1, 1e-8, 1e8, 1), (1, -3.0, 3.0, 0),
    (2, 0.0, 1.5707963267948966, 0), (2, 1.5, 1.5707963267948966, 0),
    (3, 1000.0, 1400.0, 0), (3, 1e-100, 1e100, 1),
    (4, 1.0, 1.001, 0), (4, 1.0, 2.0, 0), (4, 1e-300, 1e300, 1),
]


def _eval(name, fn, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty_like(x)
    assert getattr(native.oracle(), name)(fn, native.dptr(x), x.size, native.dptr(y)) == 0
    return y


def cr(fn, x):
    return _eval("orc_cr_eval", fn, x)


def quad(fn, x):
    return _eval("orc_quad_eval", fn, x)


def libm(fn, x):
    return _eval("orc_libm_eval", fn, x)


@pytest.mark.parametrize("dom", DOMAINS, ids=["%d:%g..%g" % d[:3] for d in DOMAINS])
def test_correctly_rounded_on_sample(dom):
    fn, lo, hi, lg = dom
    lib = native.oracle()
    lib.orc_cr_check.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_long, C.c_ulonglong,
                                 C.POINTER(C.c_double)]
    out = (C.c_double * 5)()
    assert lib.orc_cr_check(fn, lo, hi, lg, 200_000, 7 + fn, out) == 0
    bad_quad, bad_libm, max_ulp_libm, _, libm_vs_quad = list(out)
    assert bad_quad == 0
    # crmath equals binary128 everywhere, so it differs from glibc exactly
    # where glibc differs from binary128, and then by one ulp
    assert bad_libm == libm_vs_quad and max_ulp_libm <= 1.0


def test_committed_1e7_table_has_no_misrounding():
    rows = [ln for ln in open(os.path.join(HERE, "golden", "crmath_pin_1e7.txt")) if not ln.startswith("#")]
    assert len(rows) == len(DOMAINS)
    for ln in rows:
        cols = [c.strip() for c in ln.split("|")]
        assert int(cols[1]) == 0, ln
        bad_libm, max_ulp = cols[2].split()
        assert int(bad_libm) == int(cols[3]) and float(max_ulp) <= 1.0, ln


def test_exact_and_special_values():
    assert cr(0, np.arange(0.0, 23.0)).tolist() == [10.0 ** k for k in range(23)]   # exact powers
    assert cr(0, np.array([-1.0]))[0] == 0.1
    assert cr(0, np.array([400.0, -400.0])).tolist() == [math.inf, 0.0]
    assert math.isnan(cr(0, np.array([math.nan]))[0])
    assert cr(1, np.array([0.0, 1.0, -1.0, math.inf, -math.inf, 1e-30])).tolist() == \
        [0.0, math.pi / 4, -math.pi / 4, math.pi / 2, -math.pi / 2, 1e-30]
    assert cr(2, np.array([0.0, 1e-30, math.pi / 3, math.pi / 2])).tolist() == \
        quad(2, np.array([0.0, 1e-30, math.pi / 3, math.pi / 2])).tolist()
    assert cr(2, np.array([0.0]))[0] == 1.0
    assert math.isnan(cr(2, np.array([2.0]))[0])            # outside the documented domain
    assert cr(3, np.array([0.0, 2.0, -3.0, 1e200])).tolist() == [0.0, 8.0, -27.0, math.inf]
    assert cr(4, np.array([0.0, 1.0, math.inf])).tolist() == [0.0, 1.0, math.inf]
    assert math.isnan(cr(4, np.array([-1.0]))[0])
    # the exponent is the double just below 1/3: 8^y rounds to 2, but
    # 1000^y rounds to the double below 10, as glibc's pow also gives
    x = np.array([8.0, 27.0, 1000.0, 1e-300, 1e300, 0.001])
    assert cr(4, x).tolist() == quad(4, x).tolist()
    assert cr(4, np.array([8.0, 1000.0])).tolist() == [2.0, np.nextafter(10.0, 0.0)]


def test_values_where_glibc_misrounds_are_binary128():
    """A seeded sample holding glibc misroundings of pow(10, x): crmath
    returns the binary128 value there, one ulp from glibc."""
    x = np.linspace(-1.2, 2.0, 400_001)
    c, q, m = cr(0, x), quad(0, x), libm(0, x)
    assert np.array_equal(c, q)
    diff = c != m
    assert diff.sum() > 0
    assert (np.abs(c[diff] - m[diff]) <= np.spacing(np.abs(m[diff]))).all()


# (fn, lo, hi, log scale, whole domain inside the first evaluation's range)
CRZ_DOMAINS = [
    (0, -1.2, 2.0, 0, 1), (0, -300.0, 300.0, 0, 1), (0, -330.0, 308.5, 0, 0),
    (1, 1e-8, 1e8, 1, 1), (1, -3.0, 3.0, 0, 1), (1, 1e-40, 1e40, 1, 0), (1, -500.0, 500.0, 0, 1),
    (2, 0.0, 1.5786, 0, 1), (2, 1.5, 1.5786, 0, 1), (2, 1e-12, 0.8, 1, 0), (2, -1.6, 1.6, 0, 0),
    (4, 1.0, 1.001, 0, 1), (4, 0.5, 2.0, 0, 1), (4, 1e-300, 1e300, 1, 1), (4, 1e-320, 1e-290, 1, 0),
]


@pytest.mark.parametrize("dom", CRZ_DOMAINS, ids=["%d:%g..%g" % d[:3] for d in CRZ_DOMAINS])
def test_first_evaluation_equals_full(dom):
    """crz_pow10 / crz_atan / crz_cos / crz_pow_third (a cheaper first
    evaluation, Ziv's rounding test, the full evaluation when the test
    fails; what the GPU runs) return the bits of cr_pow10 / cr_atan / cr_cos
    / cr_pow_third at every point,
    NaN and the cases outside the first evaluation's range included; inside
    that range the full evaluation is needed at about 2^-16 of the points.
    (A sweep of 1.3 x 10^9 points over these domains found no difference.)"""
    fn, lo, hi, lg, inside = dom
    lib = native.oracle()
    lib.orc_crz_check.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_long, C.c_ulonglong,
                                  C.POINTER(C.c_double)]
    out = (C.c_double * 2)()
    n = 4_000_000
    assert lib.orc_crz_check(fn, lo, hi, lg, n, 1234 + fn, out) == 0
    assert out[0] == 0
    if inside:
        assert out[1] < n * 2.0 ** -13   # rounding-test failures only
