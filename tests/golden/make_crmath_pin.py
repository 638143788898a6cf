"""Writes crmath_pin_1e7.txt: for each function of opencal_amd/csrc/crmath.h
and each input domain, 10^7 seeded points evaluated three ways on the host:
crmath (the device's source), binary128 (libquadmath, rounded once to double)
and the host glibc exactly as the reference calls it
(OpenCAL-OMP/examples/calomp_sciara-fv2/source/Sciara.cpp:20,61,85,153-154).
The table counts the points where crmath differs from binary128 and from
glibc, and where glibc differs from binary128.

    make -C oracle && python tests/golden/make_crmath_pin.py

tests/test_crmath.py reads the committed table and runs a smaller sample of
the same check."""
import ctypes as C
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

NAMES = ["pow(10,x)", "atan(x)", "cos(x)", "pow(x,3.0)", "pow(x,1.0/3.0)"]
# (fn, lo, hi, log-uniform): the ranges SCIARA-fv2 feeds each call (power-law
# exponents, slope ratios, slope angles, temperatures, the cooling factor),
# and wide ranges beside them
DOMAINS = [
    (0, -1.2, 2.0, 0), (0, -300.0, 300.0, 0),
    (1, 0.0, 100.0, 0), (1, 1e-8, 1e8, 1), (1, -3.0, 3.0, 0),
    (2, 0.0, 1.5707963267948966, 0), (2, 1.5, 1.5707963267948966, 0),
    (3, 1000.0, 1400.0, 0), (3, 1e-100, 1e100, 1),
    (4, 1.0, 1.001, 0), (4, 1.0, 2.0, 0), (4, 1e-300, 1e300, 1),
]
N = 10_000_000
SEED = 20061001


def main() -> None:
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_build", "liboracle.so"))
    lib.orc_cr_check.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, C.c_long, C.c_ulonglong,
                                 C.POINTER(C.c_double)]
    lines = [
        "# crmath.h against binary128 and the host glibc, %d seeded points per row" % N,
        "# host: %s, glibc %s; generated by tests/golden/make_crmath_pin.py" % (platform.machine(),
                                                                                platform.libc_ver()[1]),
        "# columns: function lo hi scale | crmath!=binary128 | crmath!=glibc max_ulp | glibc!=binary128",
    ]
    for fn, lo, hi, lg in DOMAINS:
        out = (C.c_double * 5)()
        t = time.time()
        assert lib.orc_cr_check(fn, lo, hi, lg, N, SEED + fn, out) == 0
        lines.append("%-15s %-22r %-22r %-3s | %d | %d %.1f | %d" % (
            NAMES[fn], lo, hi, "log" if lg else "lin", int(out[0]), int(out[1]), out[2], int(out[4])))
        print(lines[-1], "(%.1f s)" % (time.time() - t), file=sys.stderr)
    with open(os.path.join(HERE, "crmath_pin_1e7.txt"), "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
