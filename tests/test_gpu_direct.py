"""GPU direct sums behind c2d_direct*_ (cauchykernels2d.c:31-442), l2d_direct*_
(lapkernels2d.c:19-494), r2d_direct*_ (rlapkernels2d.c:16-491) and
h2d_direct*_ (helmkernels2d.c:21-520): every
variant against the compiled reference library where it is built, and
potential/gradient against closed forms (log|z|, 1/z; log r and its
derivatives; (i/4) H0(kr) and (i/4) k H1(kr) (x-y).v / r with scipy's
hankel1), with a coincident source-target pair to exercise the skip test."""
import ctypes as C

import numpy as np
import pytest
from scipy import special as sp

from conftest import rel_l2

pytestmark = pytest.mark.gpu
P = lambda a: a.ctypes.data_as(C.c_void_p)
VARIANTS = ["cp", "cg", "ch", "dp", "dg", "dh", "cdp", "cdg", "cdh"]


def _inputs(nd, ns, nt, seed):
    rng = np.random.default_rng(seed)
    s = np.asfortranarray(rng.random((2, ns)))
    t = np.asfortranarray(rng.random((2, nt)))
    t[:, 0] = s[:, 0]  # coincident pair: skipped by every variant
    c = np.asfortranarray(rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns)))
    q = np.asfortranarray(rng.standard_normal((nd, ns)) + 1j * rng.standard_normal((nd, ns)))
    th = rng.random((nd, ns)) * 2 * np.pi
    v = np.asfortranarray(np.stack([np.cos(th), np.sin(th)], axis=1))
    return s, t, c, q, v


def _call(lib, kind, var, nd, zk, s, t, c, q, v, ncg, nch):
    ns, nt = s.shape[1], t.shape[1]
    pg = {"p": 1, "g": 2, "h": 3}[var[-1]]
    dt = float if kind == "r2d" else complex
    if kind == "r2d":
        c, q = np.asfortranarray(c.real), np.asfortranarray(q.real)
    pot = np.zeros((nd, nt), dt, order="F")
    grad = np.zeros((nd, ncg, nt), dt, order="F")
    hess = np.zeros((nd, nch, nt), dt, order="F")
    args = [C.byref(C.c_int32(nd))]
    if kind == "h2d":
        args.append((C.c_double * 2)(zk.real, zk.imag))
    args += [P(s), C.byref(C.c_int32(ns))]
    if "c" in var[:-1]:
        args.append(P(c))
    if "d" in var[:-1]:
        args.append(P(q))
        if kind != "c2d":
            args.append(P(v))
    args += [P(t), C.byref(C.c_int32(nt)), P(pot)]
    if pg >= 2:
        args.append(P(grad))
    if pg >= 3:
        args.append(P(hess))
    args.append(C.byref(C.c_double(1e-14)))
    getattr(lib, f"{kind}_direct{var}_")(*args)
    return pot, grad, hess, pg


def _closed(kind, var, zk, s, t, c, q, v):
    """pot, grad (closed forms) of the sum over sources, coincident pair dropped."""
    dx = t[0][:, None] - s[0][None]
    dy = t[1][:, None] - s[1][None]
    rr = dx * dx + dy * dy
    keep = rr > 0
    rr = np.where(keep, rr, 1.0)
    r = np.sqrt(rr)
    useC, useQ = "c" in var[:-1], "d" in var[:-1]
    cc = c[0] if useC else np.zeros(s.shape[1])
    qq = q[0] if useQ else np.zeros(s.shape[1])
    v1, v2 = v[0, 0], v[0, 1]
    if kind == "c2d":
        z = np.where(keep, dx + 1j * dy, 1.0)
        pot = (np.log(r) * keep) @ cc + (keep / z) @ qq
        g = (keep / z) @ cc - (keep / z**2) @ qq
        return pot, g[None]
    if kind in ("l2d", "r2d"):
        if kind == "r2d":
            cc, qq = cc.real, qq.real
        pot = (np.log(r) * keep) @ cc + ((-dx / rr) * keep) @ (qq * v1) + ((-dy / rr) * keep) @ (qq * v2)
        gx = ((dx / rr) * keep) @ cc + ((-(rr - 2 * dx**2) / rr**2) * keep) @ (qq * v1) \
            + ((2 * dx * dy / rr**2) * keep) @ (qq * v2)
        gy = ((dy / rr) * keep) @ cc + ((2 * dx * dy / rr**2) * keep) @ (qq * v1) \
            + ((-(rr - 2 * dy**2) / rr**2) * keep) @ (qq * v2)
        return pot, np.stack([gx, gy])
    kr = zk * r
    h0, h1 = sp.hankel1(0, kr) * keep, sp.hankel1(1, kr) * keep
    dot = dx * v1[None] + dy * v2[None]
    pot = 0.25j * (h0 @ cc + (zk * h1 * dot / r) @ qq)
    gxc = -0.25j * (zk * h1 * dx / r) @ cc
    gyc = -0.25j * (zk * h1 * dy / r) @ cc
    hp = (sp.hankel1(0, kr) - sp.hankel1(1, kr) / kr) * keep  # H1'(kr)
    gxd = 0.25j * zk * ((hp * zk * dot * dx / rr + h1 * (v1[None] / r - dot * dx / (r * rr))) @ qq)
    gyd = 0.25j * zk * ((hp * zk * dot * dy / rr + h1 * (v2[None] / r - dot * dy / (r * rr))) @ qq)
    return pot, np.stack([gxc + gxd, gyc + gyd])


@pytest.mark.parametrize("kind", ["c2d", "l2d", "r2d", "h2d"])
@pytest.mark.parametrize("var", VARIANTS)
def test_direct_sums(eng, ref, kind, var):
    print(f"direct {kind} {var}", flush=True)
    nd, zk = 2, 9.0 + 0.5j
    s, t, c, q, v = _inputs(nd, 700, 300, 3)
    ncg, nch = (1, 1) if kind == "c2d" else (2, 3)
    pot, grad, hess, pg = _call(eng, kind, var, nd, zk, s, t, c, q, v, ncg, nch)
    assert np.all(np.isfinite(pot))
    if ref is not None:
        rp, rg, rh, _ = _call(ref, kind, var, nd, zk, s, t, c, q, v, ncg, nch)
        assert rel_l2(pot, rp) < 1e-12
        if pg >= 2:
            assert rel_l2(grad, rg) < 1e-12
        if pg >= 3:
            assert rel_l2(hess, rh) < 1e-12
    # closed forms for channel 0
    cp, cg = _closed(kind, var, zk, s, t, c[:1], q[:1], v[:1])
    assert rel_l2(pot[0], cp) < 1e-12
    if pg >= 2:
        assert rel_l2(grad[0], cg) < 1e-11


@pytest.mark.parametrize("var", ["ch", "dh", "cdh"])
def test_r2d_hessian_finite_difference(eng, var):
    """The r2d Hessian against central differences of the r2d gradient
    (step 1e-6, targets kept 1e-2 away from every source: truncation error
    ~ (h/r)^2 = 1e-8, rounding ~ 1e-16 |grad| / h ~ 1e-11 relative)."""
    print(f"r2d hessian fd {var}", flush=True)
    nd = 1
    s, t, c, q, v = _inputs(nd, 300, 200, 5)
    dmin = np.sqrt(((t[:, :, None] - s[:, None, :]) ** 2).sum(0)).min(1)
    t = np.asfortranarray(t[:, dmin > 1e-2])
    _, _, hess, _ = _call(eng, "r2d", var, nd, 0, s, t, c, q, v, 2, 3)
    h = 1e-6
    gvar = var[:-1] + "g"
    def g_at(dx, dy):
        tt = np.asfortranarray(t + np.array([[dx], [dy]]))
        return _call(eng, "r2d", gvar, nd, 0, s, tt, c, q, v, 2, 3)[1][0]
    gxp, gxm = g_at(h, 0), g_at(-h, 0)
    gyp, gym = g_at(0, h), g_at(0, -h)
    hxx = (gxp[0] - gxm[0]) / (2 * h)
    hxy = (gyp[0] - gym[0]) / (2 * h)
    hyy = (gyp[1] - gym[1]) / (2 * h)
    assert rel_l2(hess[0, 0], hxx) < 1e-6
    assert rel_l2(hess[0, 1], hxy) < 1e-6
    assert rel_l2(hess[0, 2], hyy) < 1e-6
