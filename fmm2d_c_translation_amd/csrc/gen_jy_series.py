"""Writes jy_series.h: coefficients of the ascending series of J0 and of the
Y0 remainder (Abramowitz & Stegun 9.1.10, 9.1.13) as polynomials in
w = x^2/4, exactly rounded from rational arithmetic:
    J0(x) = P(w),  P(w) = sum_k JA[k] w^k,  JA[k] = (-1)^k / (k!)^2
    Y0(x) = (2/pi) [(ln(x/2) + gamma) J0(x) + Q(w)],
            Q(w) = sum_k YC[k] w^k,  YC[k] = -(-1)^k H_k / (k!)^2
with H_k the k-th harmonic number; J1 = -J0' and Y1 = -Y0' follow from
P'(w) and Q'(w) (dw/dx = x/2).  JY_NTERMS<i> is the number of terms that
leaves the truncated tails of P, P', Q and Q' below 1e-18 for x <= JY_XMAX<i>;
jy_nterms picks the shortest series for a bound on x, and the length N
names the tier.  JY_TIER[i] holds, per tier, economised forms of P and Q
on the tier's range whose values and derivatives meet the same 1e-18
bound with fewer terms (x <= 0.6: 8 instead of 9; x <= 1.2: 10 instead of
11; x <= 4: 14 instead of 18); the device evaluates those.
Usage:
    python gen_jy_series.py > jy_series.h"""
from fractions import Fraction
from math import comb, factorial

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


# Economised polynomials of each tier: P and Q over [0, W], W = XMAX^2 / 4,
# each of the fewest terms for which both the polynomial and its derivative
# stay within 1e-18 of P, P' (Q, Q'), so that one Horner chain with its
# derivative chain still gives all four.  The series to K terms (the tail
# bounded), then its top power replaced, power by power, by the shifted
# Chebyshev polynomial T*_n(w / W) = T_n(2 w / W - 1), whose leading
# coefficient is 2^(2n-1) / W^n: each step s T*_n changes the polynomial by
# at most |s| on [0, W] and its derivative by at most |s| 2 n^2 / W (Markov's
# inequality) -- Chebyshev economisation (Lanczos, Applied Analysis, 1956,
# ch. VII).  All in rational arithmetic, rounded once.
def cheb_shifted(n):
    T = [[Fraction(1)], [Fraction(0), Fraction(1)]]
    for m in range(2, n + 1):
        a = [Fraction(0)] + [2 * x for x in T[m - 1]]
        b = T[m - 2] + [Fraction(0)] * (len(a) - len(T[m - 2]))
        T.append([x - y for x, y in zip(a, b)])
    out = [Fraction(0)] * (n + 1)  # T_n(2t - 1) in powers of t
    for j, cj in enumerate(T[n]):
        for i in range(j + 1):
            out[i] += cj * comb(j, i) * 2 ** i * (-1) ** (j - i)
    return out


def economise(c, W, tol, tail, dtail):
    """Fewest-term economised form of sum_k c[k] w^k on [0, W] whose value
    and derivative stay within tol; tail and dtail bound the series' value
    and derivative beyond the len(c) terms given."""
    for d in range(1, len(c)):
        e = list(c)
        ev = ed = Fraction(0)
        for n in range(len(e) - 1, d, -1):
            if e[n] == 0:
                continue
            ts = cheb_shifted(n)
            s = e[n] * W ** n / ts[n]
            ev += abs(s)
            ed += abs(s) * 2 * n * n / W
            for i in range(n + 1):
                e[i] -= s * ts[i] / W ** i
        if ev + tail < tol and ed + dtail < tol:
            return e[:d + 1]
    raise SystemExit("economisation failed")


KE = 30  # series terms kept before economising
HE = [Fraction(0)]
for k in range(1, KE + 12):
    HE.append(HE[-1] + Fraction(1, k))
JAE = [Fraction((-1) ** k, factorial(k) ** 2) for k in range(KE + 11)]
YCE = [-Fraction((-1) ** k) * HE[k] / factorial(k) ** 2 for k in range(KE + 11)]
ECON = []
for xm in XMAX:
    W = Fraction(xm) ** 2 / 4
    tier = []
    for c in (JAE, YCE):
        # terms KE.. fall faster than geometrically (ratio below W / k^2 <
        # 1/100 here): twice the next ten bound the rest
        tail = 2 * sum(abs(c[k]) * W ** k for k in range(KE, KE + 10))
        dtail = 2 * sum(k * abs(c[k]) * W ** (k - 1) for k in range(KE, KE + 10))
        tier.append(economise(c[:KE], W, Fraction(1, 10 ** 18), tail, dtail))
    ECON.append(tier)


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
print("// economised P and Q of each tier (w in [0, JY_XMAX<i>^2 / 4]; values and")
print("// derivatives within 1e-18): term counts, then coefficients from w^0 up")
EMAX = max(len(c) for t in ECON for c in t)
print(f"#define JY_EMAX {EMAX}")
print("struct JyTier {")
print("  int n[2];")
print(f"  double c[2][{EMAX}];")
print("};")
print("constexpr JyTier JY_TIER[4] = {")
for t in ECON:
    ns = ", ".join(str(len(c)) for c in t)
    rows = ", ".join("{" + ", ".join(f"{float(v):.17e}" for v in c) + "}" for c in t)
    print(f"    {{{{{ns}}}, {{{rows}}}}},")
print("};")
print("// series length for arguments up to xmax (0: beyond the series)")
print("inline int jy_nterms(double xmax) {")
print("  return xmax <= 0.0 ? 0 : xmax <= JY_XMAX0 ? JY_NTERMS0 : xmax <= JY_XMAX1 ? JY_NTERMS1")
print("       : xmax <= JY_XMAX2 ? JY_NTERMS2 : xmax <= JY_XMAX3 ? JY_NTERMS3 : 0;")
print("}")
