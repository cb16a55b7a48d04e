"""The reference's MATLAB/numbl front end end to end on the GPU: the mwrap
gateway (matlab/fmm2d.c) and numbl shim (matlab/numbl/mex_shim.cpp) linked
against libfmm2d_hip.so, driven through mex_dispatch by the Python restatement
of cfmm2d.m / lfmm2d.m / rfmm2d.m / hfmm2d.m and the *2ddir.m direct sums.

  * The reference's own MATLAB tests (matlab/cfmm2dTest.m, hfmm2dTest.m):
    ns = 4000 sources, nt = 3999 targets, eps = 1e-5, zk = 1.1 + 0.1i, nd 1
    and 3, the first 10 points against the gateway's direct sums, error
    < eps.
  * Against the oracle's *_ndiv on the same inputs with the ndiv and
    idivflag the gateway's hndiv2d chose (same tree, 1e-11 relative l2).
  * nd = 2 with Hessians, and the Stokes direct sums through st2ddir.
"""
import numpy as np
import pytest

from fmm2d_c_translation_amd import abi, mex
from conftest import rel_l2
from test_oracle_stokes import orc_st

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not mex.available(), reason="MEX gateway not built")]
TOL = 1e-11
NTEST = 10


def _srcinfo(kind, cd, nd=1, ns=4000, seed=0):
    rng = np.random.default_rng(seed)
    s = {"sources": rng.random((2, ns))}
    sh = (ns,) if nd == 1 else (nd, ns)
    if nd > 1:
        s["nd"] = nd
    cplx = lambda: rng.random(sh) + 1j * rng.random(sh)
    if "c" in cd:  # cfmm2dTest.m:5 complex(rand): real values (the Cauchy pot's
        # real part is then free of the box-dependent log branch)
        s["charges"] = rng.random(sh) if kind in "rc" else cplx()
    if "d" in cd:
        s["dipstr"] = rng.random(sh) if kind == "r" else cplx()
        if kind != "c":
            s["dipvec"] = rng.random((2, ns) if nd == 1 else (nd, 2, ns))
    return s


def _fmm(kind, eps, zk, s, pg, targ=None, pgt=0):
    f = getattr(mex, f"{kind}fmm2d")
    return f(eps, zk, s, pg, targ, pgt) if kind == "h" else f(eps, s, pg, targ, pgt)


def _dir(kind, zk, s, targ, pgt):
    f = getattr(mex, f"{kind}2ddir")
    return f(zk, s, targ, pgt) if kind == "h" else f(s, targ, pgt)


@pytest.mark.parametrize("nd", [1, 3])
@pytest.mark.parametrize("kind", ["c", "l", "r", "h"])
@pytest.mark.parametrize("cd", ["c", "d", "cd"])
def test_reference_matlab_suite(kind, cd, nd):
    """{c,h}fmm2dTest.m: sources to sources and to targets, pg / pgt = 1..3,
    relative error over the first 10 points below eps; the Cauchy pot is
    compared by its real part (its imaginary part is box-dependent)."""
    print(f"mex {kind}fmm2d cd={cd} nd={nd}", flush=True)
    eps, zk = 1e-5, 1.1 + 0.1j
    s = _srcinfo(kind, cd, nd=nd)
    rng = np.random.default_rng(1)
    targ = rng.random((2, 3999))
    stmp, ttmp = s["sources"][:, :NTEST], targ[:, :NTEST]
    for p in (1, 2, 3):
        U1 = _fmm(kind, eps, zk, s, p)
        U2 = _dir(kind, zk, s, stmp, p)
        V1 = _fmm(kind, eps, zk, s, 0, targ, p)
        V2 = _dir(kind, zk, s, ttmp, p)
        assert U1.ier == 0 and V1.ier == 0
        pairs = [("pot", "pottarg"), ("grad", "gradtarg"), ("hess", "hesstarg")][:p]
        for (k, kt) in pairs:
            for got, ref in ((U1[k], U2[kt]), (V1[kt], V2[kt])):
                d = got[..., :NTEST] - ref
                if kind == "c" and k == "pot":
                    d = d.real
                e = np.linalg.norm(d) / np.linalg.norm(ref)
                assert e < eps, (kind, cd, nd, p, k, e)


@pytest.mark.parametrize("kind,cd,nd", [("c", "cd", 1), ("l", "cd", 2), ("r", "cd", 1),
                                        ("h", "cd", 1), ("h", "d", 2)])
def test_against_oracle(eng, orc, ref, kind, cd, nd):
    """The gateway's call of {c,l,h}fmm2d_ndiv_ equals the oracle's on the
    same inputs, with ndiv/idivflag from the gateway's hndiv2d; rfmm2d_ndiv_
    (no oracle entry of its own: its parity with the oracle's log-kernel FMM
    is covered with the C entry points) against the compiled reference where
    it is built, and against the library called directly."""
    print(f"mex {kind}fmm2d vs oracle nd={nd}", flush=True)
    eps, zk = 1e-9, 12.0 + 0.5j
    s = _srcinfo(kind, cd, nd=nd, ns=6000, seed=3)
    targ = np.random.default_rng(4).random((2, 2500))
    U = _fmm(kind, eps, zk, s, 3, targ, 3)
    assert U.ier == 0
    nd_, idf = mex.gateway().call("hndiv2d", mex.SPECS["hndiv2d"], eps, 6000, 2500,
                                  int("c" in cd), int("d" in cd), 3, 3, 20, 0)
    kw = dict(sources=s["sources"], charges=s.get("charges"), dipstr=s.get("dipstr"), pg=3,
              targets=targ, pgt=3, nd=nd, ndiv=int(nd_[0, 0]), idivflag=int(idf[0, 0]),
              ifnear=1)
    if kind != "c":
        kw["dipvec"] = s.get("dipvec")
    if kind == "h":
        checks = [abi.hfmm2d_call(orc, "orc_hfmm2d_ndiv_", eps, zk, **kw)]
    elif kind == "r":
        checks = [abi.rfmm2d_call(eng, "rfmm2d_ndiv_", eps, **kw)]
        if ref is not None:
            checks.append(abi.rfmm2d_call(ref, "rfmm2d_ndiv_", eps, **kw))
    else:
        checks = [getattr(abi, f"{kind}fmm2d_call")(orc, f"orc_{kind}fmm2d_ndiv_", eps, **kw)]
    for o in checks:
        _same(U, o)


def _same(U, o):
    assert o.ier == 0
    for k, ko in (("pot", "pot"), ("grad", "grad"), ("hess", "hess"), ("pottarg", "targ_pot"),
                  ("gradtarg", "targ_grad"), ("hesstarg", "targ_hess")):
        a, b = np.asarray(U[k]), np.asarray(o[ko])
        assert a.size == b.size, (k, a.shape, b.shape)
        assert rel_l2(a.ravel(order="F"), b.ravel(order="F")) < TOL, k


def test_st2ddir_through_gateway(orc):
    print("mex st2ddir", flush=True)
    rng = np.random.default_rng(5)
    ns, nt, nd = 1500, 700, 2
    s = {"sources": rng.random((2, ns)), "nd": nd, "stoklet": rng.standard_normal((nd, 2, ns)),
         "strslet": rng.standard_normal((nd, 2, ns)), "strsvec": rng.standard_normal((nd, 2, ns))}
    targ = rng.random((2, nt)) + 1.2
    U = mex.st2ddir(s, targ, 3)
    f = lambda k: np.asfortranarray(s[k])
    pot, pre, grad = orc_st(orc, nd, np.asfortranarray(s["sources"]), targ, f("stoklet"),
                            f("strslet"), f("strsvec"), th=1e-15)
    assert rel_l2(U.pottarg, pot) < 1e-12
    assert rel_l2(U.pretarg, pre) < 1e-12
    assert rel_l2(U.gradtarg, grad) < 1e-12
    s1 = {"sources": s["sources"], "stoklet": s["stoklet"][0]}
    V = mex.st2ddir(s1, targ, 3)
    st1 = np.asfortranarray(s["stoklet"][:1])
    p1, q1, g1 = orc_st(orc, 1, np.asfortranarray(s["sources"]), targ, st1, st1, st1,
                        ifst=1, istr=0, th=1e-15)
    assert rel_l2(V.pottarg, p1[0]) < 1e-12 and rel_l2(V.gradtarg, g1[0]) < 1e-12


@pytest.mark.parametrize("nd", [1, 2])
def test_stfmm2d_reference_suite(nd):
    """matlab/stfmm2dTest.m through the gateway: 4000 sources with
    Stokeslets and stresslets, 3999 targets, eps 1e-5, ifppreg =
    ifppregtarg = 3; pot, pre and grad at the first ten sources and targets
    against st2ddir there, relative l2 over all six below eps."""
    print(f"mex stfmm2dTest nd={nd}", flush=True)
    rng = np.random.default_rng(40 + nd)
    ns, nt, eps, ntest = 4000, 3999, 1e-5, 10
    sh = (2, ns) if nd == 1 else (nd, 2, ns)
    s = {"sources": rng.random((2, ns)), "stoklet": rng.random(sh), "strslet": rng.random(sh),
         "strsvec": rng.random(sh)}
    if nd > 1:
        s["nd"] = nd
    targ = rng.random((2, nt))
    U1 = mex.stfmm2d(eps, s, 3, targ, 3)
    assert U1.ier == 0
    U2 = mex.st2ddir(s, s["sources"][:, :ntest], 3)
    U3 = mex.st2ddir(s, targ[:, :ntest], 3)
    first = lambda a: np.asarray(a)[..., :ntest].ravel(order="F")
    pairs = [(U1.pot, U2.pottarg), (U1.pre, U2.pretarg), (U1.grad, U2.gradtarg),
             (U1.pottarg, U3.pottarg), (U1.pretarg, U3.pretarg), (U1.gradtarg, U3.gradtarg)]
    err = sum(np.linalg.norm(first(a) - np.asarray(b).ravel(order="F")) ** 2 for a, b in pairs)
    ra = sum(np.linalg.norm(np.asarray(b)) ** 2 for _, b in pairs)
    assert np.sqrt(err / ra) < eps


def test_stfmm2d_gateway_matches_library(eng):
    """The gateway's stfmm2d output equals a direct stfmm2d_ call of the
    library on the same inputs (sources only, Stokeslets only)."""
    from bhcalls import stfmm
    print("mex stfmm2d vs library", flush=True)
    rng = np.random.default_rng(44)
    ns = 3000
    s = {"sources": rng.random((2, ns)), "stoklet": rng.standard_normal((2, ns))}
    U = mex.stfmm2d(1e-9, s, 2)
    a = stfmm(eng, 1e-9, 1, np.asfortranarray(s["sources"]),
              np.asfortranarray(s["stoklet"].reshape(1, 2, ns)), None, None, 2)
    assert U.ier == 0 and a["ier"] == 0
    assert np.array_equal(np.asarray(U.pot), a["pot"][0])
    assert np.array_equal(np.asarray(U.pre).ravel(), a["pre"].ravel())
    assert U.grad is None and U.pottarg is None
