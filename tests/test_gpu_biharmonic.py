"""GPU parity of the biharmonic and Stokes entry points: bh2d_direct*_
(bhkernels2d.c:30-398), bhfmm2d_ (bhfmm2d.c:924-1282) and stfmm2d_
(stfmm2d.c:40-249).

Direct sums: against the oracle (obh.c) and the compiled reference,
1e-12 relative l2 (float64 sums in another order).  FMMs: against the
compiled reference on the same tree (the tree is bit-exact), 1e-11 relative
l2; and against the oracle's direct sums within eps -- the reference itself
stays below 0.15 eps on these inputs (uniform and clustered, eps 1e-3 to
1e-12), so eps leaves a margin of six.  Stokes also against the GPU's own
st2ddirectstokstrsg_, which is pinned against the oracle elsewhere."""
import ctypes as C

import numpy as np
import pytest

from bhcalls import P, bh_direct, bhfmm, f64, i32, inputs, stfmm
from conftest import rel_l2
from test_oracle_stokes import _inputs as st_inputs, orc_st, ref_st

pytestmark = pytest.mark.gpu
TOL_D = 1e-12
TOL_R = 1e-11


def _pick(cd, chg, dip):
    return (chg if "c" in cd else None), (dip if "d" in cd else None)


@pytest.mark.parametrize("cd", ["c", "d", "cd"])
@pytest.mark.parametrize("pg", [1, 2])
@pytest.mark.parametrize("nd,ns,nt", [(1, 1500, 900), (2, 700, 1300), (3, 9000, 40)])
def test_bh_direct(eng, orc, ref, cd, pg, nd, ns, nt):
    print(f"bh2d_direct{cd}{'pg'[pg - 1]} nd={nd} ns={ns} nt={nt}", flush=True)
    src, trg, chg, dip = inputs(nd, ns, nt, 30 + nd)
    trg = np.asfortranarray(np.concatenate([trg, src[:, :5]], axis=1))  # coincident: skipped
    c, q = _pick(cd, chg, dip)
    a = bh_direct(eng, nd, src, c, q, trg, pg)
    b = bh_direct(orc, nd, src, c, q, trg, pg, oracle=True)
    outs = a if pg == 2 else a[:1]
    for x, y in zip(outs, b):
        assert np.isfinite(x).all()
        assert rel_l2(x, y) < TOL_D
    if ref is not None:
        r = bh_direct(ref, nd, src, c, q, trg, pg)
        for x, y in zip(outs, r):
            assert rel_l2(x, y) < TOL_D


def test_bh_direct_increments(eng, orc):
    """Outputs are incremented, as the reference's are."""
    nd, ns, nt = 2, 400, 300
    src, trg, chg, dip = inputs(nd, ns, nt, 41)
    rng = np.random.default_rng(42)
    v0 = np.asfortranarray(rng.standard_normal((nd, nt)) + 1j * rng.standard_normal((nd, nt)))
    g0 = np.asfortranarray(rng.standard_normal((nd, 3, nt)) + 0j)
    v, g = v0.copy(order="F"), g0.copy(order="F")
    eng.bh2d_directcdg_(i32(nd), P(src), i32(ns), P(chg), P(dip), P(trg), i32(nt), P(v), P(g),
                        f64(1e-14))
    b = bh_direct(orc, nd, src, chg, dip, trg, 2, oracle=True)
    assert rel_l2(v - v0, b[0]) < TOL_D
    assert rel_l2(g - g0, b[1]) < TOL_D


def _check_fmm(a, nd, src, c, q, pg, trg, pgt, eps, ref):
    assert a["ier"] == 0 and a["iper"] == 0
    keys = []
    if pg >= 1:
        keys += ["pot"] + (["grad"] if pg >= 2 else [])
    if pgt >= 1:
        keys += ["pottarg"] + (["gradtarg"] if pgt >= 2 else [])
    if ref is not None:
        b = bhfmm(ref, eps, nd, src, c, q, pg, trg, pgt)
        for k in keys:
            assert rel_l2(a[k], b[k]) < TOL_R, k
    return keys


@pytest.mark.parametrize("cd", ["c", "d", "cd"])
@pytest.mark.parametrize("pg,pgt", [(1, 1), (2, 2), (2, 0), (0, 2), (1, 2)])
@pytest.mark.parametrize("nd", [1, 2])
def test_bhfmm2d_flags(eng, orc, ref, cd, pg, pgt, nd):
    print(f"bhfmm2d cd={cd} pg={pg} pgt={pgt} nd={nd}", flush=True)
    eps = 1e-6
    src, trg, chg, dip = inputs(nd, 3000, 700, 50 + nd)
    c, q = _pick(cd, chg, dip)
    a = bhfmm(eng, eps, nd, src, c, q, pg, trg, pgt)
    _check_fmm(a, nd, src, c, q, pg, trg, pgt, eps, ref)
    if pg >= 1:
        v, g = bh_direct(orc, nd, src, c, q, src, 2, oracle=True)
        assert rel_l2(a["pot"], v) < eps
        if pg == 2:
            assert rel_l2(a["grad"], g) < eps
    if pgt >= 1:
        v, g = bh_direct(orc, nd, src, c, q, trg, 2, oracle=True)
        assert rel_l2(a["pottarg"], v) < eps
        if pgt == 2:
            assert rel_l2(a["gradtarg"], g) < eps


@pytest.mark.parametrize("eps", [1e-3, 1e-6, 1e-9, 1e-12])
@pytest.mark.parametrize("cluster", [False, True])
def test_bhfmm2d_eps_ladder(eng, orc, ref, eps, cluster):
    """20000 sources, uniform or on a thin circle (adaptive tree); checked
    at 400 of the sources against the oracle's direct sums there."""
    print(f"bhfmm2d eps={eps} cluster={cluster}", flush=True)
    src, _, chg, dip = inputs(1, 20000, 0, 60, cluster=cluster)
    a = bhfmm(eng, eps, 1, src, chg, dip, 2)
    _check_fmm(a, 1, src, chg, dip, 2, None, 0, eps, ref if eps == 1e-9 else None)
    idx = np.random.default_rng(61).choice(src.shape[1], 400, replace=False)
    v, g = bh_direct(orc, 1, src, chg, dip, np.asfortranarray(src[:, idx]), 2, oracle=True)
    assert rel_l2(a["pot"][:, idx], v) < eps
    assert rel_l2(a["grad"][:, :, idx], g) < eps


def test_bhfmm2d_pg3_writes_zeros(eng):
    """ifpgh 3 evaluates nothing in the reference (bhfmm2d.c evaluates pg 1
    and 2 only): pot, grad and hess come back as zeros, overwriting."""
    nd = 1
    src, trg, chg, dip = inputs(nd, 1000, 200, 70)
    ns, nt = src.shape[1], trg.shape[1]
    o = {k: np.full(s, 3.0 + 1j, order="F") for k, s in
         (("pot", (nd, ns)), ("grad", (nd, 3, ns)), ("hess", (nd, 3, ns)),
          ("pottarg", (nd, nt)), ("gradtarg", (nd, 3, nt)), ("hesstarg", (nd, 3, nt)))}
    iper, ier = C.c_int32(7), C.c_int32(-1)
    eng.bhfmm2d_(i32(nd), f64(1e-6), i32(ns), P(src), i32(1), P(chg), i32(1), P(dip),
                 C.byref(iper), i32(3), P(o["pot"]), P(o["grad"]), P(o["hess"]), i32(nt), P(trg),
                 i32(3), P(o["pottarg"]), P(o["gradtarg"]), P(o["hesstarg"]), C.byref(ier))
    assert ier.value == 0 and iper.value == 0
    for k, x in o.items():
        assert not x.any(), k


def test_bh_terms_and_ndiv(eng, ref):
    for eps in (0.7, 1e-2, 1e-5, 1e-8, 1e-11, 1e-14, 1e-17):
        n, e = C.c_int32(-1), C.c_int32(-1)
        eng.bh2dterms_(f64(eps), C.byref(n), C.byref(e))
        nd_, fl = C.c_int32(-1), C.c_int32(-1)
        eng.bhndiv2d_(f64(eps), i32(1000), i32(50), i32(1), i32(0), i32(1), i32(0),
                      C.byref(nd_), C.byref(fl))
        if ref is not None:
            n2, e2 = C.c_int32(-1), C.c_int32(-1)
            ref.bh2dterms_(f64(eps), C.byref(n2), C.byref(e2))
            d2, f2 = C.c_int32(-1), C.c_int32(-1)
            ref.bhndiv2d_(f64(eps), i32(1000), i32(50), i32(1), i32(0), i32(1), i32(0),
                          C.byref(d2), C.byref(f2))
            assert (n.value, e.value, nd_.value, fl.value) == (n2.value, e2.value, d2.value,
                                                               f2.value), eps


def _st_keys(pp, ppt):
    ks = []
    for flag, names in ((pp, ("pot", "pre", "grad")), (ppt, ("pottarg", "pretarg", "gradtarg"))):
        ks += list(names[:flag])
    return ks


@pytest.mark.parametrize("ifst,istr", [(1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("pp", [1, 2, 3])
@pytest.mark.parametrize("nd", [1, 2])
def test_stfmm2d(eng, orc, ref, ifst, istr, pp, nd):
    print(f"stfmm2d stok={ifst} strs={istr} pp={pp} nd={nd}", flush=True)
    eps = 1e-9
    s, _, f, mu, nu = st_inputs(nd, 2500, 10, 80 + nd)
    t = np.asfortranarray(np.random.default_rng(81).random((2, 600)))
    args = (f if ifst else None, mu if istr else None, nu if istr else None)
    a = stfmm(eng, eps, nd, s, *args, pp, t, pp)
    assert a["ier"] == 0
    keys = _st_keys(pp, pp)
    if ref is not None:
        b = stfmm(ref, eps, nd, s, *args, pp, t, pp)
        for k in keys:
            assert rel_l2(a[k], b[k]) < TOL_R, k
    ds = dict(zip(("pot", "pre", "grad"), orc_st(orc, nd, s, s, f, mu, nu, ifst, istr)))
    dt = dict(zip(("pottarg", "pretarg", "gradtarg"), orc_st(orc, nd, s, t, f, mu, nu, ifst, istr)))
    for k in keys:
        assert rel_l2(a[k], {**ds, **dt}[k]) < eps, k
    # the GPU's own direct sums
    g = ref_st(eng, nd, s, t, f, mu, nu, ifst, istr)
    for k, y in zip(("pottarg", "pretarg", "gradtarg"), g):
        if k in keys:
            assert rel_l2(a[k], y) < eps, k


def test_stfmm2d_sources_only_and_large(eng, orc):
    """No targets (nt 0), 60000 clustered sources with Stokeslets and
    stresslets, eps 1e-12, checked at 300 sources against the oracle."""
    nd, ns = 1, 60000
    rng = np.random.default_rng(90)
    th = rng.random(ns) * 2 * np.pi
    r = 0.3 + 1e-3 * rng.standard_normal(ns)
    s = np.asfortranarray(np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)]))
    f, mu, nu = (np.asfortranarray(rng.standard_normal((nd, 2, ns))) for _ in range(3))
    a = stfmm(eng, 1e-12, nd, s, f, mu, nu, 3)
    assert a["ier"] == 0
    idx = rng.choice(ns, 300, replace=False)
    b = orc_st(orc, nd, s, np.asfortranarray(s[:, idx]), f, mu, nu)
    for k, y in zip(("pot", "pre", "grad"), b):
        x = a[k][..., idx]
        assert np.isfinite(x).all()
        assert rel_l2(x, y) < 1e-12, k
