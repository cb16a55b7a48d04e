"""Trees in a numbering-free form, for comparing the device tree (level-major,
Morton order inside a level) with the reference's (its own numbering after
level restriction, c_translation/src/pts_tree2d.c:966-1081).

A box is named by its cell (level, ix, iy); ix, iy index the 2^level x
2^level grid of the root box, digit bit 0 = x half, bit 1 = y half
(children in the order (-x,-y), (+x,-y), (-x,+y), (+x,+y),
tree_routs2d.c:77-92).
"""
import numpy as np

def make_points(kind, seed, n):
    rng = np.random.default_rng(seed)
    if kind == "unif":
        return rng.random((2, n))
    if kind == "circle":
        th = rng.random(n) * 2 * np.pi
        r = 0.4 + 1e-3 * rng.standard_normal(n)
        src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
        src[:, : n // 10] = rng.random((2, n // 10))
        return src
    if kind == "blob":
        return 0.5 + 0.05 * rng.standard_normal((2, n))
    if kind == "twoblob":
        a = 0.2 + 1e-3 * rng.standard_normal((2, n // 2))
        b = 0.8 + 0.02 * rng.standard_normal((2, n - n // 2))
        return np.concatenate([a, b], axis=1)
    if kind == "line":
        t = rng.random(n)
        return np.stack([t, 0.3 + 1e-4 * t])
    if kind in ("deep", "deeper"):  # uniform plus a tight cluster: trees 17-25 levels deep
        src = rng.random((2, n))
        sig = 1e-5 if kind == "deep" else 1e-7
        k = n // 6
        src[:, :k] = np.array([[0.3], [0.7]]) + sig * rng.standard_normal((2, k))
        return src
    if kind == "grid":
        g = np.linspace(0, 1, int(np.sqrt(n)))
        X, Y = np.meshgrid(g, g)
        return np.stack([X.ravel(), Y.ravel()])
    raise ValueError(kind)


def ref_cells(t):
    """Cell of every box of a reference-layout tree (oracle.Tree, 1-based)."""
    cell = [None] * t.nboxes
    cell[0] = (0, 0, 0)
    for b in range(t.nboxes):  # parents precede their children
        if t.nchild[b]:
            l, ix, iy = cell[b]
            for q in range(4):
                cell[t.ichild[b, q] - 1] = (l + 1, 2 * ix + (q & 1), 2 * iy + (q >> 1))
    return cell


def canon_ref(t):
    """{cell: (nchild, nsrc, ntrg, (cx, cy), (l1, l2, l3, l4) as sorted cell
    tuples, leaf sources, leaf targets)} of a reference-layout tree."""
    cell = ref_cells(t)
    out = {}
    widths = [(0, 13), (13, 27), (40, 20), (60, 5)]
    for b in range(t.nboxes):
        s0, s1 = t.isrcse[b]
        t0, t1 = t.itargse[b]
        lists = []
        for k, (o, _) in enumerate(widths):
            ids = t.lists[b, o:o + t.nlist[b, k]]
            lists.append(tuple(sorted(cell[i - 1] for i in ids)))
        leaf = t.nchild[b] == 0
        srcs = tuple(t.isrc[s0 - 1:s1] - 1) if leaf else ()
        trgs = tuple(t.itarg[t0 - 1:t1] - 1) if leaf else ()
        out[cell[b]] = (int(t.nchild[b]), int(s1 - s0 + 1), int(t1 - t0 + 1),
                        (float(t.centers[b, 0]), float(t.centers[b, 1])), tuple(lists),
                        srcs, trgs)
    return out


def canon_dev(nb, boxinfo, centers, lists, isrc, itarg):
    """Same form for the device tree returned by fmm2dhip_test_devtree."""
    cell = [tuple(int(v) for v in boxinfo[b, :3]) for b in range(nb)]
    out = {}
    p = 0
    for b in range(nb):
        lev, ix, iy, nch, s0, s1, t0, t1 = (int(v) for v in boxinfo[b, :8])
        ls = []
        for k in range(4):
            n = int(boxinfo[b, 8 + k])
            ls.append(tuple(sorted(cell[i] for i in lists[p:p + n])))
            p += n
        leaf = nch == 0
        srcs = tuple(isrc[s0:s1]) if leaf else ()
        trgs = tuple(itarg[t0:t1]) if leaf else ()
        out[cell[b]] = (nch, s1 - s0, t1 - t0, (float(centers[b, 0]), float(centers[b, 1])),
                        tuple(ls), srcs, trgs)
    return out
