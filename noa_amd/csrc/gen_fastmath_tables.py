"""Generates fastmath_tables.hh: the lookup tables of the table-driven FP64
log and exp in fastmath.hh (Tang's method).

    python noa_amd/csrc/gen_fastmath_tables.py

All values are computed with 60-digit decimal arithmetic and rounded once.

log: x = 2^k z with z in [0.5, 1) (the significand with exponent field
0x3fe); z is cut into 256 subintervals by the top 8 mantissa bits of x; for
subinterval i with centre c_i (c = 0.5 for the first, which holds z = x / 2
for x just above 1, and c = 1 for the last, which holds x just below 1),
  invc[i]  = 1 / c_i rounded to double,
  logc_hi  = -log(invc[i]) rounded to a multiple of 2^-43, minus 1022 LN2_HI
             (exact: a multiple of 2^-43 below 2^10), so that E LN2_HI +
             logc_hi = k LN2_HI - log(invc) is exact with E = k + 1022, the
             biased exponent field of x (no integer subtraction per call),
  logc_lo  = the remainder of -log(invc[i]), minus 1022 LN2_LO, rounded to
             double (exact for the first and last subintervals: LN2_LO has
             42 significant bits, so near x = 1 the table parts cancel
             exactly and log x keeps full relative accuracy).
  Then r = z invc - 1 lies in [-2^-9, 2^-8] (the widest ranges are those of
  the two subintervals at 1, where r = 2z - 1 or z - 1 exactly).
log1p(r) = r + r^2 P(r): P is the degree-4 Chebyshev economization of the
  series (log1p(r) - r) / r^2 = sum_n (-1)^(n+1) r^n / (n + 2) over that r
  range (exact rational arithmetic: the series to degree 40 in the shifted
  and scaled variable, its Chebyshev expansion cut after T_4, converted back
  to powers of r), coefficients rounded once; the truncation error is below
  0.07 * 2^-53 relative to log1p(r) (checked here against 60-digit ln).
exp: T[j] = 2^(j/128), j < 128, as a double-double (hi, lo).
shorter exp: the degree-2 Chebyshev economization Q of (e^r - 1 - r) / r^2
  over the same r range (truncation below 1.5 ulp, checked here).
"""
from __future__ import annotations

import os
from decimal import Decimal, getcontext

from fractions import Fraction
from math import comb

getcontext().prec = 60
HERE = os.path.dirname(os.path.abspath(__file__))
N_LOG = 256
LOG_SHIFT = 44  # 52 - log2(N_LOG)
N_EXP = 128
LOG_DEG = 4
HALF = 0x3fe0000000000000  # bits of 0.5
E_BIAS = 1022  # biased exponent field of x = 2^k z, z in [0.5, 1): E = k + 1022


def d2f(d: Decimal) -> float:
    """Decimal -> nearest double."""
    return float(d)


def bits_to_float(b: int) -> float:
    import struct
    return struct.unpack("<d", struct.pack("<Q", b))[0]


def ln(d: Decimal) -> Decimal:
    return d.ln()


def round_to_multiple(d: Decimal, q: Decimal) -> Decimal:
    return (d / q).to_integral_value() * q


def ln2_parts():
    """LN2_HI: ln 2 rounded to a multiple of 2^-43; LN2_LO: the remainder
    rounded to 42 significant bits (a multiple of 2^-86), so that E LN2_LO
    is exact for the biased exponents E < 2^11. Exact rationals."""
    ln2 = Fraction(Decimal(2).ln())
    hi = Fraction(round(ln2 * 2 ** 43), 2 ** 43)
    lo = Fraction(round((ln2 - hi) * 2 ** 86), 2 ** 86)
    assert Fraction(float(hi)) == hi and Fraction(float(lo)) == lo
    return hi, lo


def log_tables():
    invc, lhi, llo = [], [], []
    ln2_hi, ln2_lo = ln2_parts()
    for i in range(N_LOG):
        a = Decimal(bits_to_float(HALF + (i << LOG_SHIFT)))
        b = Decimal(bits_to_float(HALF + ((i + 1) << LOG_SHIFT))) if i + 1 < N_LOG else Decimal(1)
        c = (a + b) / 2
        if i == 0:
            c = Decimal(1) / 2  # x just above 1: r = 2 z - 1 exactly
        elif i == N_LOG - 1:
            c = Decimal(1)  # x just below 1: r = z - 1 exactly
        ic = d2f(Decimal(1) / c)
        if i == 0:
            h, lo = -ln2_hi, -ln2_lo  # -log 2 as exactly -(LN2_HI + LN2_LO)
        else:
            lc = Fraction(-ln(Decimal(ic)))
            h = Fraction(round(lc * 2 ** 43), 2 ** 43)
            lo = lc - h
        hf = h - E_BIAS * ln2_hi
        assert Fraction(float(hf)) == hf  # the folded high part stays exact
        lof = lo - E_BIAS * ln2_lo
        if i in (0, N_LOG - 1):
            assert Fraction(float(lof)) == lof  # exact where log x -> 0
        invc.append(ic)
        lhi.append(float(hf))
        llo.append(float(lof))
    return invc, lhi, llo


def _to_chebyshev(p):
    """Power coefficients in s -> Chebyshev coefficients (Horner in the T basis)."""
    n = len(p) - 1
    c = [Fraction(0)] * (n + 2)
    for k in range(n, -1, -1):
        nc = [Fraction(0)] * (n + 2)
        for j, v in enumerate(c):
            if v == 0:
                continue
            if j == 0:
                nc[1] += v
            else:  # s T_j = (T_(j+1) + T_(j-1)) / 2
                nc[j + 1] += v / 2
                nc[j - 1] += v / 2
        nc[0] += p[k]
        c = nc
    return c


def _chebyshev_to_power(c):
    T = [[Fraction(1)], [Fraction(0), Fraction(1)]]
    for j in range(2, len(c)):
        a = [Fraction(0)] + [2 * v for v in T[j - 1]]
        b = T[j - 2] + [Fraction(0)] * (len(a) - len(T[j - 2]))
        T.append([x - y for x, y in zip(a, b)])
    p = [Fraction(0)] * len(c)
    for j in range(len(c)):
        for k, v in enumerate(T[j]):
            p[k] += c[j] * v
    return p


def log_poly(lo: Fraction, hi: Fraction, deg: int = LOG_DEG, terms: int = 40):
    """Coefficients (constant first) of P(r) ~ (log1p(r) - r) / r^2 on [lo, hi]."""
    m, h = (lo + hi) / 2, (hi - lo) / 2
    ps = [Fraction(0)] * (terms + 1)  # the series in s, r = m + h s
    for n in range(terms + 1):
        cn = Fraction((-1) ** (n + 1), n + 2)
        for k in range(n + 1):
            ps[k] += cn * comb(n, k) * m ** (n - k) * h ** k
    p_s = _chebyshev_to_power(_to_chebyshev(ps)[:deg + 1])
    pr = [Fraction(0)] * (deg + 1)  # back to r: s = (r - m) / h
    for k, v in enumerate(p_s):
        for j in range(k + 1):
            pr[j] += v * comb(k, j) * (1 / h) ** k * (-m) ** (k - j)
    coef = [float(v) for v in pr]
    worst = Decimal(0)
    for i in range(2001):
        r = lo + (hi - lo) * Fraction(i, 2000)
        if r == 0:
            continue
        q = Fraction(0)
        for v in reversed(coef):
            q = q * r + Fraction(v)
        approx = r + r * r * q
        exact = (1 + Decimal(r.numerator) / Decimal(r.denominator)).ln()
        worst = max(worst, abs((Decimal(approx.numerator) / Decimal(approx.denominator) - exact) / exact))
    assert worst < Decimal(2) ** -56, worst
    return coef


def h_poly(deg: int = 6, hi: Fraction = Fraction(1, 4), terms: int = 60):
    """FP32 coefficients (constant first) of P(u) ~ h(u) = (log1p(u) - u) / u^2
    on [0, hi] for the mixed-precision pair integrand: the series
    sum_n (-1)^(n+1) u^n / (n + 2) to `terms` terms in the shifted and scaled
    variable, its Chebyshev expansion cut after T_deg, back to powers of u;
    coefficients rounded to float. Checked here: < 1e-8 relative over [0, hi]
    in exact arithmetic (FP32 evaluation adds its own ~6e-8)."""
    import struct
    m, h = hi / 2, hi / 2
    ps = [Fraction(0)] * (terms + 1)
    for n in range(terms + 1):
        cn = Fraction((-1) ** (n + 1), n + 2)
        for k in range(n + 1):
            ps[k] += cn * comb(n, k) * m ** (n - k) * h ** k
    p_s = _chebyshev_to_power(_to_chebyshev(ps)[:deg + 1])
    pr = [Fraction(0)] * (deg + 1)
    for k, v in enumerate(p_s):
        for j in range(k + 1):
            pr[j] += v * comb(k, j) * (1 / h) ** k * (-m) ** (k - j)
    coef = [struct.unpack("<f", struct.pack("<f", float(v)))[0] for v in pr]
    worst = Decimal(0)
    for i in range(1, 2001):
        u = hi * Fraction(i, 2000)
        ud = Decimal(u.numerator) / Decimal(u.denominator)
        exact = ((1 + ud).ln() - ud) / (ud * ud)
        v = Decimal(0)
        for c in reversed(coef):
            v = v * ud + Decimal(c)
        worst = max(worst, abs((v - exact) / exact))
    assert worst < Decimal("1e-8"), worst
    return coef


def exp4_poly(deg: int = 2, terms: int = 30):
    """Coefficients (constant first) of Q(r) ~ (e^r - 1 - r) / r^2 on [-h, h],
    h = ln2 / (2 N_EXP) with a margin for the rounding of k: the degree-2
    Chebyshev economization of sum_n r^n / (n + 2)! (the series of Q), for the
    shorter exp (exp_fin4: e^r - 1 = r (1 + r Q(r))). Truncation error checked
    below 1.5 ulp of e^r (2^-52 relative units)."""
    from math import factorial
    ln2 = Fraction(Decimal(2).ln())
    h = ln2 / (2 * N_EXP) * (1 + Fraction(1, 1 << 20))
    ps = [Fraction(1, factorial(n + 2)) * h ** n for n in range(terms + 1)]  # the series in s, r = h s
    p_s = _chebyshev_to_power(_to_chebyshev(ps)[:deg + 1])
    coef = [float(v * (1 / h) ** k) for k, v in enumerate(p_s)]
    worst = Decimal(0)
    for i in range(-1000, 1001):
        r = h * Fraction(i, 1000)
        q = Fraction(0)
        for v in reversed(coef):
            q = q * r + Fraction(v)
        approx = 1 + r + r * r * q
        exact = (Decimal(r.numerator) / Decimal(r.denominator)).exp()
        worst = max(worst, abs((Decimal(approx.numerator) / Decimal(approx.denominator) - exact) / exact))
    assert worst < Decimal(1.5) * Decimal(2) ** -52, worst
    return coef


def exp_tables():
    th, tl = [], []
    for j in range(N_EXP):
        v = (Decimal(2) ** (Decimal(j) / N_EXP))
        h = d2f(v)
        th.append(h)
        tl.append(d2f(v - Decimal(h)))
    return th, tl


def constants():
    ln2 = Decimal(2).ln()
    ln2_hi, ln2_lo = ln2_parts()
    ln2n = ln2 / N_EXP
    ln2n_hi = d2f(ln2n)
    ln2n_lo = d2f(ln2n - Decimal(ln2n_hi))
    return {
        "LN2_HI": float(ln2_hi), "LN2_LO": float(ln2_lo),
        "INV_LN2_N": d2f(N_EXP / ln2), "LN2_N_HI": ln2n_hi, "LN2_N_LO": ln2n_lo,
    }


def fmt(vals) -> str:
    rows = []
    for k in range(0, len(vals), 4):
        rows.append("    " + ", ".join(float.hex(v) for v in vals[k:k + 4]) + ",")
    return "\n".join(rows)


def main() -> None:
    invc, lhi, llo = log_tables()
    th, tl = exp_tables()
    c = constants()
    out = [
        "// Generated by gen_fastmath_tables.py; do not edit.",
        "// Lookup tables and constants of the table-driven FP64 log and exp.",
        "// NOA_FM_TABLE is defined by fastmath.hh (device constant memory under hipcc).",
        "#pragma once",
        "",
        "namespace noa_hip {",
        "namespace fm {",
        "",
        f"constexpr int LOG_N = {N_LOG};",
        f"constexpr int LOG_SHIFT = {LOG_SHIFT};",
        f"constexpr int EXP_N = {N_EXP};",
    ]
    for k, v in c.items():
        out.append(f"constexpr double {k} = {float.hex(v)};")
    poly = log_poly(-Fraction(1, 2 * N_LOG), Fraction(1, N_LOG))
    out.append("// log1p(r) = r + r^2 (LOG_P0 + r (LOG_P1 + ...)), r in [-2^-9, 2^-8]")
    for k, v in enumerate(poly):
        out.append(f"constexpr double LOG_P{k} = {float.hex(v)};")
    out.append("// shorter exp: e^r - 1 = r (1 + r (EXP4_P0 + r (EXP4_P1 + r EXP4_P2))), |r| <= ln2 / 256")
    for k, v in enumerate(exp4_poly()):
        out.append(f"constexpr double EXP4_P{k} = {float.hex(v)};")
    out.append("// FP32: (log1p(u) - u) / u^2 = H_P0 + u (H_P1 + ...), u in [0, 1/4]")
    for k, v in enumerate(h_poly()):
        out.append(f"constexpr float H_P{k} = {float.hex(v)}f;")
    for name, vals in (("LOG_INVC", invc), ("LOG_LOGC_HI", lhi), ("LOG_LOGC_LO", llo),
                       ("EXP_T_HI", th), ("EXP_T_LO", tl)):
        out += ["", f"NOA_FM_TABLE double {name}[{len(vals)}] = {{", fmt(vals), "};"]
    out += ["", "}  // namespace fm", "}  // namespace noa_hip", ""]
    with open(os.path.join(HERE, "fastmath_tables.hh"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
