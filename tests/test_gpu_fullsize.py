"""Parity at the benchmark configurations' full sizes.

The FMM result at these sizes is checked through properties that do not
need a CPU FMM of the same size: relative l2 error against FP64 direct sums
at a fixed random subsample of 10^4 evaluation points (the GPU direct-sum
entry points, checked against the compiled reference and closed forms in
test_gpu_direct.py), which must stay within eps -- the reference's own
accuracy criterion (test/helmholtz/test_hfmm2d.f:106).  The Laplace
configuration of 1e5 points is also compared with the compiled reference
itself where it travelled with the tree.

  C1: cfmm2d_ndiv, 1e5 uniform points, complex charges, pot + grad at the
      sources, ndiv 40, eps 1e-6 (FP64).
  C2: hfmm2d, 1e6 sources + 1e6 targets, zk = 40, charges + dipoles,
      eps 1e-9 and 1e-12 (FP64) and 1e-6 (mixed).
  C3: cfmm2d, 4e6 points (90% on a thin circle), real charges, eps ladder
      1e-3 .. 1e-12, FP64 and (eps >= 1e-6) mixed; Re(pot) is gauge-free for
      real charges."""
import ctypes as C

import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import abi
from conftest import rel_l2

pytestmark = pytest.mark.gpu
P = lambda a: a.ctypes.data_as(C.c_void_p)
NSUB = 10_000


def _direct(kind, var, zk, s, t, c=None, q=None, v=None, ncg=1):
    lib = F.lib()
    ns, nt = s.shape[1], t.shape[1]
    pot = np.zeros((1, nt), complex, order="F")
    grad = np.zeros((1, ncg, nt), complex, order="F")
    args = [C.byref(C.c_int32(1))]
    if kind == "h2d":
        args.append((C.c_double * 2)(complex(zk).real, complex(zk).imag))
    args += [P(np.asfortranarray(s)), C.byref(C.c_int32(ns))]
    if c is not None:
        args.append(P(np.asfortranarray(np.asarray(c, complex).reshape(1, ns))))
    if q is not None:
        args.append(P(np.asfortranarray(np.asarray(q, complex).reshape(1, ns))))
        if kind != "c2d":
            args.append(P(np.asfortranarray(v)))
    args += [P(np.asfortranarray(t)), C.byref(C.c_int32(nt)), P(pot), P(grad),
             C.byref(C.c_double(1e-14))]
    getattr(lib, f"{kind}_direct{var}g_")(*args)
    return pot[0], grad[0]


def test_c1_matches_reference(ref):
    print("C1 cfmm2d_ndiv 1e5", flush=True)
    rng = np.random.default_rng(1)
    n = 100_000
    src = rng.random((2, n))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    kw = dict(eps=1e-6, sources=src, charges=ch, pg=2, ndiv=40)
    a = abi.cfmm2d_call(F.lib(), "cfmm2d_ndiv_", **kw)
    assert a.ier == 0
    sub = np.sort(rng.choice(n, NSUB, replace=False))
    _, g = _direct("c2d", "c", 0, src, src[:, sub], c=ch)
    assert rel_l2(a["grad"][sub], g[0]) < 1e-6
    if ref is not None:
        b = abi.cfmm2d_call(ref, "cfmm2d_ndiv_", **kw)
        assert rel_l2(a["pot"], b["pot"]) < 1e-11
        assert rel_l2(a["grad"], b["grad"]) < 1e-11


@pytest.mark.parametrize("eps,mode", [(1e-9, F.FP64), (1e-6, F.MIXED), (1e-12, F.FP64)])
def test_c2_within_eps_of_direct_sum(eps, mode):
    print(f"C2 hfmm2d 1e6+1e6 eps={eps} mode={mode}", flush=True)
    rng = np.random.default_rng(2)
    n = 1_000_000
    src, trg = rng.random((2, n)), rng.random((2, n))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dp = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    th = rng.random(n) * 2 * np.pi
    dv = np.stack([np.cos(th), np.sin(th)])[None]
    with F.precision(mode):
        r = abi.hfmm2d_call(F.lib(), "hfmm2d_", eps, 40.0, src, charges=ch, dipstr=dp, dipvec=dv,
                            pg=2, targets=trg, pgt=2)
    assert r.ier == 0
    for pts, pk, gk in ((src, "pot", "grad"), (trg, "targ_pot", "targ_grad")):
        sub = np.sort(rng.choice(n, NSUB, replace=False))
        p, g = _direct("h2d", "cd", 40.0, src, pts[:, sub], c=ch, q=dp, v=dv, ncg=2)
        ep, eg = rel_l2(r[pk][sub], p), rel_l2(r[gk][:, sub], g)
        print(f"  {pk}: {ep:.2e}  {gk}: {eg:.2e}", flush=True)
        assert ep < eps and eg < eps


@pytest.fixture(scope="module")
def c3():
    """4e6 points, 90% within ~1e-3 of a circle of radius 0.4, 10% uniform;
    real charges; direct sums at a fixed subsample (computed once)."""
    rng = np.random.default_rng(3)
    n = 4_000_000
    th = rng.random(n) * 2 * np.pi
    r = 0.4 + 1e-3 * rng.standard_normal(n)
    src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
    k = n // 10
    src[:, :k] = rng.random((2, k))
    ch = rng.standard_normal(n)
    sub = np.sort(rng.choice(n, NSUB, replace=False))
    p, g = _direct("c2d", "c", 0, src, src[:, sub], c=ch)
    return src, ch, sub, p.real, g[0]


@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-9, 1e-12])
def test_c3_precision_ladder(c3, eps):
    src, ch, sub, dpot, dgrad = c3
    modes = [F.FP64] + ([F.MIXED] if eps >= 1e-6 else [])
    for mode in modes:
        print(f"C3 cfmm2d 4e6 circle eps={eps} mode={mode}", flush=True)
        with F.precision(mode):
            r = abi.cfmm2d_call(F.lib(), "cfmm2d_", eps, src, charges=ch, pg=2)
        assert r.ier == 0
        ep = rel_l2(r["pot"][sub].real, dpot)
        eg = rel_l2(r["grad"][sub], dgrad)
        print(f"  Re pot {ep:.2e}  grad {eg:.2e}", flush=True)
        assert ep < eps and eg < eps
