"""Mixed precision: FP32 near field (Laplace) and M2L (Laplace and Helmholtz)
when eps >= 1e-6, selected per thread with fmm2dhip_set_precision.

The bar is the one the FP64 solve is held to: relative l2 error within eps
against FP64 direct sums.  The direct sums come from this library's GPU
c2d/l2d/h2d_direct entry points, which tests/test_gpu_direct.py checks
against the compiled reference and against closed forms (log|z|, 1/z,
scipy's hankel1).  The Cauchy potential carries a box-centre-dependent
imaginary gauge (laprouts2d.c:193-224), so it is compared with the FP64 FMM
on the same tree instead.  Below eps = 1e-6 the mixed mode must reproduce
the FP64 solve bit for bit; Hessians (pg = 3) keep the FP64 near field."""
import ctypes as C

import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from conftest import rel_l2

pytestmark = pytest.mark.gpu
P = lambda a: a.ctypes.data_as(C.c_void_p)


def _inputs(seed, n, nt, kind="unif"):
    rng = np.random.default_rng(seed)
    if kind == "unif":
        src = rng.random((2, n))
    else:  # 90% on a thin circle, 10% uniform
        th = rng.random(n) * 2 * np.pi
        r = 0.4 + 1e-3 * rng.standard_normal(n)
        src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
        k = n // 10
        src[:, :k] = rng.random((2, k))
    tg = rng.random((2, nt))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dp = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = rng.random(n) * 2 * np.pi
    dv = np.stack([np.cos(a), np.sin(a)])[None]
    return np.asfortranarray(src), np.asfortranarray(tg), ch, dp, dv


def _direct(kind, var, zk, s, t, c, q, v, ncg):
    """pot, grad of the library's GPU direct sum (self pairs skipped)."""
    lib = F.lib()
    ns, nt = s.shape[1], t.shape[1]
    pot = np.zeros((1, nt), complex, order="F")
    grad = np.zeros((1, ncg, nt), complex, order="F")
    args = [C.byref(C.c_int32(1))]
    if kind == "h2d":
        args.append((C.c_double * 2)(zk.real, zk.imag))
    args += [P(s), C.byref(C.c_int32(ns))]
    cc = np.asfortranarray(np.asarray(c, complex).reshape(1, ns)) if c is not None else None
    qq = np.asfortranarray(np.asarray(q, complex).reshape(1, ns)) if q is not None else None
    vv = np.asfortranarray(v) if v is not None else None
    if cc is not None:
        args.append(P(cc))
    if qq is not None:
        args.append(P(qq))
        if kind != "c2d":
            args.append(P(vv))
    args += [P(t), C.byref(C.c_int32(nt)), P(pot), P(grad), C.byref(C.c_double(1e-14))]
    getattr(lib, f"{kind}_direct{var}g_")(*args)
    return pot[0], grad[0]


def _both(fn):
    with F.precision(F.FP64):
        a = fn()
    with F.precision(F.MIXED):
        b = fn()
    assert a.ier == 0 and b.ier == 0, (a.ier, b.ier)
    return a, b


@pytest.mark.parametrize("eps", [1e-3, 1e-6])
@pytest.mark.parametrize("cd", ["c", "d", "cd"])
@pytest.mark.parametrize("kind", ["unif", "circle"])
def test_cfmm2d_mixed_within_eps(eps, cd, kind):
    print(f"cfmm2d mixed eps={eps} cd={cd} kind={kind}", flush=True)
    s, t, ch, dp, _ = _inputs(21, 20000, 5000, kind)
    c = ch if "c" in cd else None
    q = dp if "d" in cd else None
    a, b = _both(lambda: F.cfmm2d(eps, s, charges=c, dipstr=q, pg=2, targets=t, pgt=2))
    for pts, pk, gk in ((s, "pot", "grad"), (t, "targ_pot", "targ_grad")):
        _, g = _direct("c2d", cd, 0, s, pts, c, q, None, 1)
        assert rel_l2(a[gk], g[0]) < eps, ("fp64", gk, rel_l2(a[gk], g[0]))
        e = rel_l2(b[gk], g[0])
        print(f"  {gk}: fp64 {rel_l2(a[gk], g[0]):.2e} mixed {e:.2e}", flush=True)
        assert e < eps, (gk, e)
        assert rel_l2(b[pk], a[pk]) < eps, (pk, rel_l2(b[pk], a[pk]))


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_lfmm2d_mixed_within_eps(eps):
    print(f"lfmm2d mixed eps={eps}", flush=True)
    s, t, ch, dp, dv = _inputs(22, 20000, 4000)
    a, b = _both(lambda: F.lfmm2d(eps, s, charges=ch, dipstr=dp, dipvec=dv, pg=2, targets=t,
                                  pgt=2))
    for pts, pk, gk in ((s, "pot", "grad"), (t, "targ_pot", "targ_grad")):
        p, g = _direct("l2d", "cd", 0, s, pts, ch, dp, dv, 2)
        for key, ref in ((pk, p), (gk, g)):
            e = rel_l2(b[key], ref)
            print(f"  {key}: fp64 {rel_l2(a[key], ref):.2e} mixed {e:.2e}", flush=True)
            assert e < eps, (key, e)


@pytest.mark.parametrize("zk", [40.0, 5.0 + 0.3j])
def test_hfmm2d_mixed_within_eps(zk):
    print(f"hfmm2d mixed zk={zk}", flush=True)
    eps = 1e-6
    s, t, ch, dp, dv = _inputs(23, 20000, 4000)
    a, b = _both(lambda: F.hfmm2d(eps, zk, s, charges=ch, dipstr=dp, dipvec=dv, pg=2, targets=t,
                                  pgt=2))
    for pts, pk, gk in ((s, "pot", "grad"), (t, "targ_pot", "targ_grad")):
        p, g = _direct("h2d", "cd", complex(zk), s, pts, ch, dp, dv, 2)
        for key, ref in ((pk, p), (gk, g)):
            e = rel_l2(b[key], ref)
            print(f"  {key}: fp64 {rel_l2(a[key], ref):.2e} mixed {e:.2e}", flush=True)
            assert e < eps, (key, e)


def test_mixed_is_fp64_below_1e6_and_for_hessians():
    s, t, ch, dp, dv = _inputs(24, 6000, 1000)
    a, b = _both(lambda: F.cfmm2d(1e-9, s, charges=ch, dipstr=dp, pg=2, targets=t, pgt=2))
    for k in ("pot", "grad", "targ_pot", "targ_grad"):
        assert np.array_equal(a[k], b[k]), k
    a, b = _both(lambda: F.hfmm2d(1e-8, 20.0, s, charges=ch, pg=2))
    for k in ("pot", "grad"):
        assert np.array_equal(a[k], b[k]), k
    # Hessians keep the FP64 near field; M2L is FP32, so agreement is to eps
    a, b = _both(lambda: F.cfmm2d(1e-6, s, charges=ch, pg=3))
    assert rel_l2(b["hess"], a["hess"]) < 1e-6
