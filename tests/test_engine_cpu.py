"""CPU-side checks of the engine library: it loads, exports every symbol
its header declares, and its host-side tree build (fed keys computed on the
CPU through the test hook) reproduces the reference tree bit for bit."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle
import fmm2d_c_translation_amd as F
from conftest import ROOT

HDR = os.path.join(ROOT, "include", "fmm2d_hip.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^(?:void|int|long long|fmm2dhip_server)\s*\**\s*(\w+)\s*\(",
                                 txt, flags=re.M)))


def test_header_symbols_exported(eng):
    names = declared()
    assert "cfmm2d_" in names and "lfmm2d_ndiv_" in names
    assert "fmm2dhip_server_create" in names and "fmm2dhip_server_submit" in names
    for n in names:
        assert hasattr(eng, n), n


def test_precision_mode_is_thread_state(eng):
    prev = eng.fmm2dhip_set_precision(F.MIXED)
    assert eng.fmm2dhip_get_precision() == F.MIXED
    assert eng.fmm2dhip_set_precision(7) == F.MIXED  # invalid: unchanged
    assert eng.fmm2dhip_get_precision() == F.MIXED
    with F.precision(F.FP64):
        assert eng.fmm2dhip_get_precision() == F.FP64
    assert eng.fmm2dhip_get_precision() == F.MIXED
    eng.fmm2dhip_set_precision(prev)


def test_library_fails_loudly_when_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(F, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(F, "_lib", None)
    with pytest.raises(F.ExtensionMissing):
        F.lib()


def _hook_tree(src, tg, ndiv, idiv):
    lib = F.lib()
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    src = np.asfortranarray(src)
    ns = src.shape[1]
    tgt = np.zeros((2, 1), order="F") if tg is None else np.asfortranarray(tg)
    nt = 0 if tg is None else tg.shape[1]
    nl, nb = C.c_int(0), C.c_int(0)
    assert lib.fmm2dhip_test_tree(P(src), ns, P(tgt), nt, idiv, ndiv, C.byref(nl), C.byref(nb),
                                  *([None] * 9)) == 0
    L, n = nl.value, nb.value
    it = np.zeros(17 * n + 2 * (L + 1), np.int32)
    ce = np.zeros((n, 2)); bs = np.zeros(L + 1)
    isrc = np.zeros(max(ns, 1), np.int32); isrcse = np.zeros((n, 2), np.int32)
    itarg = np.zeros(max(nt, 1), np.int32); itargse = np.zeros((n, 2), np.int32)
    nlist = np.zeros((n, 4), np.int32); lists = np.zeros((n, 65), np.int32)
    assert lib.fmm2dhip_test_tree(P(src), ns, P(tgt), nt, idiv, ndiv, C.byref(nl), C.byref(nb),
                                  P(it), P(ce), P(bs), P(isrc), P(isrcse), P(itarg), P(itargse),
                                  P(nlist), P(lists)) == 0
    t = oracle.Tree(L, n, it, ce, bs, isrc[:ns], isrcse, itarg[:nt], itargse)
    t.nlist, t.lists = nlist, lists
    return t


GOLD = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "lap_*.npz")))


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_engine_tree_matches_golden(path):
    g = dict(np.load(path, allow_pickle=False))
    t = _hook_tree(g["src"], g.get("targ"), 15, 0)
    assert t.nlevels == int(g["tree_nlevels"]) and t.nboxes == int(g["tree_nboxes"])
    for k in ("itree", "centers", "boxsize", "isrc", "isrcse", "nlist", "lists"):
        assert np.array_equal(getattr(t, k), g["tree_" + k]), k


@pytest.mark.parametrize("seed,n,ndiv,idiv,kind", [
    (1, 6000, 8, 0, "unif"), (2, 6000, 8, 1, "unif"), (3, 6000, 8, 2, "unif"),
    (4, 30000, 10, 0, "circle"), (5, 10000, 40, 0, "blob"), (6, 2500, 3, 0, "circle"),
    (7, 4096, 20, 0, "grid")])
def test_engine_tree_matches_oracle(seed, n, ndiv, idiv, kind):
    rng = np.random.default_rng(seed)
    if kind == "unif":
        src = rng.random((2, n))
    elif kind == "circle":
        th = rng.random(n) * 2 * np.pi
        r = 0.4 + 1e-3 * rng.standard_normal(n)
        src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
        src[:, : n // 10] = rng.random((2, n // 10))
    elif kind == "blob":
        src = 0.5 + 0.05 * rng.standard_normal((2, n))
    else:
        g = np.linspace(0, 1, int(np.sqrt(n)))
        X, Y = np.meshgrid(g, g)
        src = np.stack([X.ravel(), Y.ravel()])
    tg = rng.random((2, n // 3)) if kind == "unif" else None
    a = _hook_tree(src, tg, ndiv, idiv)
    b = oracle.oracle_tree(src, tg, ndiv, idiv)
    assert (a.nlevels, a.nboxes) == (b.nlevels, b.nboxes)
    for k in ("itree", "centers", "boxsize", "isrc", "isrcse", "itarg", "itargse", "nlist", "lists"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_engine_h2dterms_matches_reference(eng, ref):
    """The engine's expansion-order rule (host code, evaluated over the
    orders it needs) against the compiled reference's h2dterms_ (h2dterms.c:22-78)."""
    if ref is None:
        pytest.skip("reference library not built here")
    rng = np.random.default_rng(11)
    cases = [(bs, zk, eps) for bs in (1.0, 0.25, 1 / 64) for zk in (40.0, 3.0 + 0.5j, 100.0)
             for eps in (1e-3, 1e-6, 1e-10, 1e-14)]
    cases += [(2.0 ** -int(rng.integers(0, 8)) * rng.uniform(0.5, 1.5),
               complex(rng.uniform(0.1, 120), rng.choice([0.0, rng.uniform(0, 5)])),
               float(rng.choice([1e-4, 1e-9, 1e-12]))) for _ in range(40)]
    for bs, zk, eps in cases:
        out = []
        for lib in (eng, ref):
            n, ier = C.c_int32(0), C.c_int32(0)
            lib.h2dterms_(C.byref(C.c_double(bs)), (C.c_double * 2)(complex(zk).real, complex(zk).imag),
                          C.byref(C.c_double(eps)), C.byref(n), C.byref(ier))
            out.append((n.value, ier.value))
        assert out[0] == out[1], (bs, zk, eps, out)
