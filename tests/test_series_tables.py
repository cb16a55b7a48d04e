"""The real Bessel series tables of the Helmholtz kernels (csrc/jy_series.h,
written by csrc/gen_jy_series.py), checked on the CPU:

  1. each tier's economised P and Q (the J0 series and the Y0 remainder in
     w = x^2/4, A&S 9.1.10 and 9.1.13), with their coefficients as the
     doubles stored, against the exact series in rational arithmetic: value
     and derivative within 1e-18 plus the coefficients' own rounding, over
     the tier's whole range;
  2. H0 and H1 assembled from them as the device assembles them
     (special.h, hank_series_real_w) against scipy's Bessel functions, an
     independent implementation: a dropped term, a wrong sign or a wrong
     tier range fails here;
  3. the modulus-form tables for real x > 4 (csrc/hank_asym.h, written by
     csrc/gen_hank_asym.py) assembled into H0 and H1 as the device does,
     against scipy."""
import os
import re
from fractions import Fraction
from math import factorial

import numpy as np
import pytest
from scipy import special as sp

from conftest import ROOT

HDR = os.path.join(ROOT, "fmm2d_c_translation_amd", "csrc", "jy_series.h")
EULER = 0.57721566490153286061


def _tables():
    txt = open(HDR).read()
    xmax = [float(v) for v in re.findall(r"JY_XMAX\d = ([0-9.e+-]+);", txt)]
    body = txt[txt.index("JY_TIER[4] = {"):]
    tiers = []
    for m in re.finditer(r"\{\{(\d+), (\d+)\}, \{\{([^}]*)\}, \{([^}]*)\}\}\}", body):
        n = (int(m.group(1)), int(m.group(2)))
        cs = [[float(v) for v in m.group(g).split(",")][:n[i]] for i, g in enumerate((3, 4))]
        tiers.append(cs)
    assert len(xmax) == 4 and len(tiers) == 4
    return xmax, tiers


def _series(k):
    H = sum(Fraction(1, j) for j in range(1, k + 1))
    return Fraction((-1) ** k, factorial(k) ** 2), -Fraction((-1) ** k) * H / factorial(k) ** 2


@pytest.mark.parametrize("tier", range(4))
def test_economised_tiers_match_the_series(tier):
    xmax, tiers = _tables()
    W = Fraction(xmax[tier]) ** 2 / 4
    exact = [_series(k) for k in range(40)]
    for which, c in enumerate(tiers[tier]):
        cf = [Fraction(v) for v in c]
        # rounding of the stored coefficients (half an ulp each, at w = W)
        slack = sum(abs(Fraction(np.spacing(v))) / 2 * W ** k * (1 + k / W)
                    for k, v in enumerate(c))
        for t in np.linspace(0.0, 1.0, 41):
            w = W * Fraction(t)
            ev = sum(e[which] * w ** k for k, e in enumerate(exact))
            ed = sum(k * e[which] * w ** (k - 1) for k, e in enumerate(exact) if k)
            v = sum(a * w ** k for k, a in enumerate(cf))
            d = sum(k * a * w ** (k - 1) for k, a in enumerate(cf) if k)
            assert abs(v - ev) < Fraction(1, 10 ** 18) + slack, (tier, which, float(t))
            assert abs(d - ed) < Fraction(1, 10 ** 18) + slack, (tier, which, float(t))


def _horner_d(c, w):
    v, dv = c[-1] * w + c[-2], c[-1]
    for a in reversed(c[:-2]):
        dv = dv * w + v
        v = v * w + a
    return v, dv


@pytest.mark.parametrize("tier", range(4))
def test_tier_hankel_values_match_scipy(tier):
    xmax, tiers = _tables()
    x = np.concatenate([np.geomspace(1e-6, xmax[tier], 400), np.linspace(0.01, xmax[tier], 400)])
    w = x * x * 0.25
    p, dp = _horner_d(tiers[tier][0], w)
    q, dq = _horner_d(tiers[tier][1], w)
    hx = 0.5 * x
    lg = np.log(hx) + EULER
    j1 = -dp * hx
    y0 = (lg * p + q) * (2.0 / np.pi)
    y1 = -(2.0 / np.pi) * (p / x - lg * j1 + dq * hx)
    for got, ref, name in ((p, sp.j0(x), "J0"), (j1, sp.j1(x), "J1"), (y0, sp.y0(x), "Y0"),
                           (y1, sp.y1(x), "Y1")):
        err = np.abs(got - ref) / np.maximum(1.0, np.abs(ref))
        assert err.max() < 4e-15, (tier, name, float(x[np.argmax(err)]), err.max())


HA = os.path.join(ROOT, "fmm2d_c_translation_amd", "csrc", "hank_asym.h")


def test_modulus_form_tables_match_scipy():
    """H0 and H1 for real x > 4 from hank_asym.h, assembled as the device
    assembles them (special.h, hank_modulus_real), against scipy's Hankel
    functions (an independent implementation), from x = 4 to 1e4."""
    txt = open(HA).read()
    tab = {m.group(1): [float(v) for v in m.group(2).split(",")]
           for m in re.finditer(r"HA_(\w+)\[\d+\] = \{([^}]*)\}", txt)}
    assert set(tab) == {"P0", "Q0", "P1", "Q1"}
    x = np.concatenate([np.geomspace(4.0 + 1e-9, 1e4, 3000), [4.0 + 1e-12, 8.0, 20.0]])
    u = 4.0 / x
    s = 2.0 * u - 1.0

    def horner(c):
        v = np.full_like(s, c[-1])
        for a in reversed(c[:-1]):
            v = v * s + a
        return v

    p0, q0, p1, q1 = (horner(tab[k]) for k in ("P0", "Q0", "P1", "Q1"))
    a = np.sqrt(u * (0.5 / np.pi)) * np.sqrt(0.5)
    cr, ci = (np.cos(x) + np.sin(x)) * a, (np.sin(x) - np.cos(x)) * a
    h0 = (p0 * cr - q0 * ci) + 1j * (p0 * ci + q0 * cr)
    h1 = (p1 * ci + q1 * cr) + 1j * (q1 * ci - p1 * cr)
    for got, ref, name in ((h0, sp.hankel1(0, x), "H0"), (h1, sp.hankel1(1, x), "H1")):
        err = np.abs(got - ref) / np.abs(ref)
        assert err.max() < 3e-15, (name, float(x[np.argmax(err)]), err.max())
