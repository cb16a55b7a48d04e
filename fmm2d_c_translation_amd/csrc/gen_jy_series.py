"""Writes jy_series.h: coefficients of the ascending series of J0 and of the
Y0 remainder (Abramowitz & Stegun 9.1.10, 9.1.13) as polynomials in
w = x^2/4, exactly rounded from rational arithmetic:
    J0(x) = P(w),  P(w) = sum_k JA[k] w^k,  JA[k] = (-1)^k / (k!)^2
    Y0(x) = (2/pi) [(ln(x/2) + gamma) J0(x) + Q(w)],
            Q(w) = sum_k YC[k] w^k,  YC[k] = -(-1)^k H_k / (k!)^2
with H_k the k-th harmonic number; J1 = -J0' and Y1 = -Y0' follow from
P'(w) and Q'(w) (dw/dx = x/2).  JY_NTERMS<i> is the number of terms that
leaves the truncated tails of P, P', Q and Q' below 1e-18 for x <= JY_XMAX<i>;
jy_nterms picks the shortest series for a bound on x.
Usage:
    python gen_jy_series.py > jy_series.h"""
from fractions import Fraction
from math import factorial

K = 24
H = [Fraction(0)]
for k in range(1, K + 1):
    H.append(H[-1] + Fraction(1, k))
JA = [Fraction((-1) ** k, factorial(k) ** 2) for k in range(K)]
YC = [-Fraction((-1) ** k) * H[k] / factorial(k) ** 2 for k in range(K)]

XMAX = [0.35, 0.6, 1.2, 4.0]
NT = []
for xm in XMAX:
    w = Fraction(xm) ** 2 / 4
    for n in range(2, K):
        tails = []
        for c in (JA, YC):
            tails.append(sum(abs(c[k]) * w ** k for k in range(n, K)))
            tails.append(sum(k * abs(c[k]) * w ** (k - 1) for k in range(n, K)))
        if max(tails) < Fraction(1, 10 ** 18):
            NT.append(n)
            break
    else:
        raise SystemExit("series too short")


def row(name, c):
    return f"constexpr double {name}[{K}] = {{" + ", ".join(f"{float(v):.17e}" for v in c) + "};"


print("// jy_series.h -- generated by gen_jy_series.py: ascending-series")
print("// coefficients of J0 and of the Y0 remainder in w = x^2/4.  Do not edit.")
print("#pragma once")
print(f"#define JY_K {K}")
print(row("JY_JA", JA))
print(row("JY_YC", YC))
for i, (xm, n) in enumerate(zip(XMAX, NT)):
    print(f"constexpr double JY_XMAX{i} = {xm!r};")
    print(f"constexpr int JY_NTERMS{i} = {n};")
print("// series length for arguments up to xmax (0: beyond the series)")
print("inline int jy_nterms(double xmax) {")
print("  return xmax <= 0.0 ? 0 : xmax <= JY_XMAX0 ? JY_NTERMS0 : xmax <= JY_XMAX1 ? JY_NTERMS1")
print("       : xmax <= JY_XMAX2 ? JY_NTERMS2 : xmax <= JY_XMAX3 ? JY_NTERMS3 : 0;")
print("}")
