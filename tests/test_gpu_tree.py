"""The device tree (fmm2d_c_translation_amd/csrc/dtree.hip) against the
oracle's reference-exact tree: same boxes, centres (bit for bit), point
counts, leaf point sets in original order, and lists 1-4 as sets of boxes
(tree_routs2d.c:369-510).  Compared in a numbering-free form (treecanon)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from treecanon import canon_dev, canon_ref, make_points

pytestmark = pytest.mark.gpu


def dev_tree(eng, src, tg, ndiv, idiv):
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    src = np.ascontiguousarray(src.T)
    ns = src.shape[0]
    tgt = np.zeros((1, 2)) if tg is None else np.ascontiguousarray(tg.T)
    nt = 0 if tg is None else tgt.shape[0]
    nl, nb, nli = C.c_int(0), C.c_int(0), C.c_int(0)
    cap_b, cap_l = 1, 1
    for _ in range(2):
        bi = np.zeros((cap_b, 12), np.int32)
        ce = np.zeros((cap_b, 2))
        li = np.zeros(cap_l, np.int32)
        isrc = np.zeros(ns, np.int32)
        itarg = np.zeros(max(nt, 1), np.int32)
        rc = eng.fmm2dhip_test_devtree(P(src), ns, P(tgt), nt, idiv, ndiv, cap_b, cap_l,
                                       C.byref(nl), C.byref(nb), C.byref(nli), P(bi), P(ce),
                                       P(li), P(isrc), P(itarg))
        if rc == 0:
            return nl.value, canon_dev(nb.value, bi, ce, li, isrc, itarg[:nt])
        assert rc == 102, rc
        cap_b, cap_l = nb.value, max(nli.value, 1)
    raise AssertionError("capacity retry failed")


CASES = [("unif", 1, 3000, 8, 0), ("unif", 2, 3000, 8, 1), ("unif", 3, 3000, 8, 2),
         ("circle", 4, 8000, 10, 0), ("blob", 5, 6000, 20, 0), ("circle", 6, 2500, 3, 0),
         ("grid", 7, 4096, 20, 0), ("twoblob", 8, 4000, 5, 0), ("line", 9, 3000, 4, 0),
         ("twoblob", 10, 3000, 1, 2), ("unif", 11, 1, 1, 0), ("unif", 12, 40, 100, 0),
         ("unif", 13, 200000, 20, 0), ("circle", 14, 100000, 15, 2),
         # deeper than the first point sort covers: the keys are sorted again on every bit
         ("deep", 15, 3000, 10, 0), ("deeper", 16, 3000, 10, 2)]


@pytest.mark.parametrize("kind,seed,n,ndiv,idiv", CASES)
def test_device_tree_matches_reference_tree(eng, kind, seed, n, ndiv, idiv):
    print(f"devtree {kind} n={n} ndiv={ndiv} idiv={idiv}", flush=True)
    src = make_points(kind, seed, n)
    tg = make_points("unif", seed + 100, max(n // 3, 1)) if (idiv != 0 or seed % 2) else None
    L, d = dev_tree(eng, src, tg, ndiv, idiv)
    t = oracle.oracle_tree(src, tg, ndiv, idiv)
    r = canon_ref(t)
    assert L == t.nlevels
    assert set(d) == set(r) and len(d) == len(r)
    names = ("nchild", "nsrc", "ntrg", "centre", "lists", "leaf sources", "leaf targets")
    for cell, want in r.items():
        got = d[cell]
        for k, name in enumerate(names):
            assert got[k] == want[k], (cell, name, got[k], want[k])
