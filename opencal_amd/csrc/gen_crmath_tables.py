"""Generates crmath_tables.h: the constants and tables of crmath.h, each a
double-double (or triple-double) split of a value computed here to 60
significant digits with Python's decimal module.

    python opencal_amd/csrc/gen_crmath_tables.py > opencal_amd/csrc/crmath_tables.h

Splitting: hi = the double nearest the value, lo = the double nearest
(value - hi), and for a triple mid/lo likewise on the remainder.  Every
number is printed as a C99 hexadecimal literal, so the compilers read back the
exact bits.
"""
from decimal import Decimal, getcontext

getcontext().prec = 60
D = Decimal


def atan_small(x: Decimal) -> Decimal:
    """Taylor series of atan, for |x| < 0.2."""
    s, term, k = D(0), x, 0
    x2 = x * x
    while True:
        t = term / (2 * k + 1)
        if abs(t) < D(10) ** -70:
            return s
        s += -t if k % 2 else t
        term *= x2
        k += 1


def atan(x: Decimal) -> Decimal:
    # atan x = 2 atan(x / (1 + sqrt(1 + x^2))), applied until the argument is small
    halvings = 0
    while abs(x) > D("0.1"):
        x = x / (1 + (1 + x * x).sqrt())
        halvings += 1
    return atan_small(x) * (2 ** halvings)


def pi() -> Decimal:
    # Machin: pi/4 = 4 atan(1/5) - atan(1/239)
    return 4 * (4 * atan_small(D(1) / 5) - atan_small(D(1) / 239))


def sin_cos(x: Decimal):
    s, c, term, k = D(0), D(0), D(1), 0
    while True:   # term = x^k / k!
        if abs(term) < D(10) ** -70:
            return s, c
        r = k % 4
        if r == 0:
            c += term
        elif r == 1:
            s += term
        elif r == 2:
            c -= term
        else:
            s -= term
        k += 1
        term = term * x / k


def split(v: Decimal, parts: int = 2):
    out = []
    for _ in range(parts):
        h = float(v)
        out.append(h)
        v = v - D(h)
    return out


def hexlit(x: float) -> str:
    return "0.0" if x == 0.0 else x.hex()


def dd(v: Decimal) -> str:
    h, l = split(v)
    return "{%s, %s}" % (hexlit(h), hexlit(l))


def table(name: str, vals) -> str:
    rows = ",\n".join("\t" + dd(v) for v in vals)
    return "CR_TABLE cr_dd_t %s[%d] = {\n%s\n};\n" % (name, len(vals), rows)


def main() -> None:
    ln2 = D(2).ln()
    ln10 = D(10).ln()
    p = pi()
    out = ["/* Generated by gen_crmath_tables.py: do not edit.  Double-double",
           " * (hi, lo) splits of values computed to 60 digits. */",
           "#ifndef OPENCAL_CRMATH_TABLES_H", "#define OPENCAL_CRMATH_TABLES_H", ""]
    # scalar constants
    l2h, l2l = split(ln2)
    out.append("#define CR_LN2_HI %s" % hexlit(l2h))
    out.append("#define CR_LN2_LO %s" % hexlit(l2l))
    a, b, c = split(ln2 / 64, 3)
    out.append("#define CR_LN2_64_1 %s  /* ln2/64 = 1 + 2 + 3 */" % hexlit(a))
    out.append("#define CR_LN2_64_2 %s" % hexlit(b))
    out.append("#define CR_LN2_64_3 %s" % hexlit(c))
    out.append("#define CR_INV_LN2_64 %s  /* 64/ln2, rounded */" % hexlit(float(64 / ln2)))
    h, l = split(ln10)
    out.append("#define CR_LN10_HI %s" % hexlit(h))
    out.append("#define CR_LN10_LO %s" % hexlit(l))
    a, b, c = split(p / 2, 3)
    out.append("#define CR_PI_2_HI %s  /* pi/2 = HI + MID + LO */" % hexlit(a))
    out.append("#define CR_PI_2_MID %s" % hexlit(b))
    out.append("#define CR_PI_2_LO %s" % hexlit(c))
    out.append("#define CR_PI_4 %s  /* pi/4, rounded */" % hexlit(float(p / 4)))
    out.append("")
    # series coefficients that need a low part (the rest are read as doubles)
    fact = [D(1)]
    for k in range(1, 16):
        fact.append(fact[-1] * k)
    coeffs = {
        "CR_EXP_C": [D(1) / fact[k] for k in range(0, 7)],            # 1/k!, k = 0..6
        "CR_ODD_INV": [D(1) / (2 * k + 1) for k in range(0, 5)],      # 1/(2k+1), k = 0..4
        "CR_SIN_C": [D(1) / fact[2 * k + 1] for k in range(0, 4)],    # 1/(2k+1)!, k = 0..3
        "CR_COS_C": [D(1) / fact[2 * k] for k in range(0, 5)],        # 1/(2k)!, k = 0..4
    }
    for name, vals in coeffs.items():
        out.append(table(name, vals))
    # tables
    out.append("/* 2^(j/64), j = 0..63 */")
    out.append(table("CR_EXP2_J64", [(D(j) / 64 * ln2).exp() for j in range(64)]))
    out.append("/* ln(1 + j/64), j = 0..64 */")
    out.append(table("CR_LOG_C64", [(1 + D(j) / 64).ln() for j in range(65)]))
    out.append("/* atan(j/64), j = 0..64 */")
    out.append(table("CR_ATAN_J64", [atan(D(j) / 64) for j in range(65)]))
    sc = [sin_cos(D(j) / 64) for j in range(52)]
    out.append("/* sin(j/64), j = 0..51 */")
    out.append(table("CR_SIN_J64", [s for s, _ in sc]))
    out.append("/* cos(j/64), j = 0..51 */")
    out.append(table("CR_COS_J64", [c for _, c in sc]))
    out.append("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
