"""The device tree's construction rule, checked on the CPU against the
reference tree.

The device builds the adaptive tree (split while a box holds more than ndiv
points, c_translation/src/pts_tree2d.c:530-610) and then, in place of the
reference's flag passes (pts_tree2d.c:767-960), the closure under the 2:1
rule "every in-domain same-level neighbour of a split box exists", finest
level first.  This test restates that construction in Python on the keys
(fmm2d_c_translation_amd/csrc/dtree.hip) and checks that its box set and
split set equal the oracle's reference-exact tree on clustered, ragged and
degenerate point sets.  The HIP code itself is compared with the same trees
in test_gpu_tree.py.
"""
import numpy as np
import pytest

import oracle
from treecanon import make_points, ref_cells

DEPTH = 29


def point_keys(p, cx, cy, bs0):
    """Morton path keys by the reference's centre recurrence
    (tree_routs2d.c:81-84), as fmm2d_c_translation_amd/csrc/common.h."""
    x, y = p
    n = x.size
    key = np.zeros(n, np.uint64)
    cx = np.full(n, cx)
    cy = np.full(n, cy)
    bs = bs0
    for l in range(1, DEPTH + 1):
        qx = (x - cx >= 0).astype(np.uint64)
        qy = (y - cy >= 0).astype(np.uint64)
        bs = bs / 2
        cx = cx + np.where(qx == 1, 1.0, -1.0) * bs / 2
        cy = cy + np.where(qy == 1, 1.0, -1.0) * bs / 2
        key |= (qx | (qy << np.uint64(1))) << np.uint64(2 * (DEPTH - l))
    return np.sort(key)


def cell_of(key):
    ix = iy = 0
    j = 0
    while key:
        ix |= (key & 1) << j
        iy |= ((key >> 1) & 1) << j
        key >>= 2
        j += 1
    return ix, iy


def key_of(ix, iy):
    k = 0
    j = 0
    while ix or iy:
        k |= (ix & 1) << (2 * j) | (iy & 1) << (2 * j + 1)
        ix >>= 1
        iy >>= 1
        j += 1
    return k


def closure_tree(src, tg, ndiv, idiv):
    pts = src if tg is None else np.concatenate([src, tg], axis=1)
    xmin, xmax = pts[0].min(), pts[0].max()
    ymin, ymax = pts[1].min(), pts[1].max()
    bs0 = max(xmax - xmin, ymax - ymin)
    cx, cy = (xmin + xmax) / 2, (ymin + ymax) / 2
    ks = point_keys(src, cx, cy, bs0)
    kt = point_keys(tg, cx, cy, bs0) if tg is not None else np.zeros(0, np.uint64)

    def count(k, key, l):
        lo = np.uint64(key << (2 * (DEPTH - l)))
        hi = np.uint64((key + 1) << (2 * (DEPTH - l))) if (key + 1) < 4 ** l else None
        a = np.searchsorted(k, lo)
        b = k.size if hi is None else np.searchsorted(k, hi)
        return b - a

    split = []
    level = [0]
    l = 0
    while True:
        r = []
        for key in level:
            cs, ct = count(ks, key, l), count(kt, key, l)
            c = cs if idiv == 0 else ct if idiv == 1 else max(cs, ct)
            if c > ndiv:
                r.append(key)
        split.append(set(r))
        if not r:
            break
        level = [4 * k + q for k in sorted(r) for q in range(4)]
        l += 1
    L = l
    for k in range(L - 1, 1, -1):
        side = 1 << k
        for key in split[k]:
            ix, iy = cell_of(key)
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    jx, jy = ix + dx, iy + dy
                    if 0 <= jx < side and 0 <= jy < side:
                        split[k - 1].add(key_of(jx, jy) >> 2)
    boxes = {(0, 0, 0)}
    for l in range(L):
        for key in split[l]:
            for q in range(4):
                boxes.add((l + 1,) + cell_of(4 * key + q))
    splitc = {(l,) + cell_of(k) for l in range(L) for k in split[l]}
    return L, boxes, splitc


CASES = [("unif", 1, 3000, 8, 0), ("unif", 2, 3000, 8, 1), ("unif", 3, 3000, 8, 2),
         ("circle", 4, 8000, 10, 0), ("blob", 5, 6000, 20, 0), ("circle", 6, 2500, 3, 0),
         ("grid", 7, 4096, 20, 0), ("twoblob", 8, 4000, 5, 0), ("line", 9, 3000, 4, 0),
         ("twoblob", 10, 3000, 1, 2)]


@pytest.mark.parametrize("kind,seed,n,ndiv,idiv", CASES)
def test_closure_equals_reference_tree(kind, seed, n, ndiv, idiv):
    src = make_points(kind, seed, n)
    tg = make_points("unif", seed + 100, n // 3) if idiv != 0 else None
    L, boxes, splitc = closure_tree(src, tg, ndiv, idiv)
    t = oracle.oracle_tree(src, tg, ndiv, idiv)
    cells = ref_cells(t)
    assert L == t.nlevels
    assert set(cells) == boxes and len(cells) == len(boxes)
    assert {cells[b] for b in range(t.nboxes) if t.nchild[b]} == splitc
