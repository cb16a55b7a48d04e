"""Writes hank_asym.h: H0 and H1 of a real argument x > 4 in the modulus
form of Hankel's expansion (Abramowitz & Stegun 9.2.5-9.2.6),
    H_nu(x) = sqrt(2 / (pi x)) (P_nu(x) + i Q_nu(x)) e^{i (x - (2 nu + 1) pi / 4)},
with P_nu and Q_nu -- smooth functions of u = 4 / x on (0, 1], tending to
1 and 0 as u -> 0 -- given as polynomials in s = 2 u - 1 in [-1, 1]: the
Chebyshev interpolants of P_nu, Q_nu (computed from mpmath's hankel1 at 40
digits) truncated where the remaining Chebyshev coefficients sum below
1e-18, then written in powers of s (the sum of their magnitudes stays near
|P_nu(0)|, so Horner in s loses nothing) and rounded once to doubles.
Usage:
    python gen_hank_asym.py > hank_asym.h"""
import mpmath as mp

mp.mp.dps = 40
NODES = 48


def pq(nu, u):
    if u == 0:
        return mp.mpf(1), mp.mpf(0)
    x = 4 / u
    v = mp.hankel1(nu, x) * mp.expj(-(x - (2 * nu + 1) * mp.pi / 4)) * mp.sqrt(mp.pi * x / 2)
    return v.real, v.imag


def cheb(f):
    """Chebyshev coefficients of f on u in [0, 1] from NODES first-kind nodes."""
    t = [mp.cos(mp.pi * (k + mp.mpf(1) / 2) / NODES) for k in range(NODES)]
    v = [f((tk + 1) / 2) for tk in t]
    c = [2 * sum(v[k] * mp.cos(mp.pi * j * (k + mp.mpf(1) / 2) / NODES) for k in range(NODES)) / NODES
         for j in range(NODES)]
    c[0] /= 2
    return c


def to_powers(c):
    T = [[mp.mpf(1)], [mp.mpf(0), mp.mpf(1)]]
    for m in range(2, len(c)):
        a = [mp.mpf(0)] + [2 * x for x in T[m - 1]]
        b = T[m - 2] + [mp.mpf(0)] * (len(a) - len(T[m - 2]))
        T.append([x - y for x, y in zip(a, b)])
    out = [mp.mpf(0)] * len(c)
    for j, cj in enumerate(c):
        for i, t in enumerate(T[j]):
            out[i] += cj * t
    return out


rows = []
for nu in (0, 1):
    for part in (0, 1):
        c = cheb(lambda u: pq(nu, u)[part])
        n = len(c)
        while n > 1 and sum(abs(x) for x in c[n - 1:]) < mp.mpf(10) ** -18:
            n -= 1
        n += 1  # the first coefficient kept above the cut
        rows.append((("P" if part == 0 else "Q") + str(nu), to_powers(c[:n])))

print("// hank_asym.h -- generated by gen_hank_asym.py: the modulus form of H0 and")
print("// H1 for real x > 4, P_nu and Q_nu as polynomials in s = 8/x - 1.  Do not edit.")
print("#pragma once")
for name, a in rows:
    print(f"constexpr int HA_N{name} = {len(a)};")
    print(f"constexpr double HA_{name}[{len(a)}] = {{" + ", ".join(f"{float(x):.17e}" for x in a) + "};")
