"""The oracle's biharmonic direct sums (oracle/obh.c, restating
bhkernels2d.c:30-398) pinned without the reference: g1 + g2 = d vel / dz and
g3 = d vel / d conj(z), by central differences (vel_x = g1 + g2 + g3,
vel_y = i (g1 + g2 - g3)), and the coincident-pair skip; and against the
compiled reference's bh2d_direct*_ where it is built."""
import numpy as np
import pytest

from bhcalls import bh_direct, inputs
from conftest import rel_l2


@pytest.mark.parametrize("cd", ["c", "d", "cd"])
def test_gradient_is_the_derivative(orc, cd):
    nd = 2
    src, trg, chg, dip = inputs(nd, 30, 5, 1)
    trg = np.asfortranarray(trg + 1.5)
    c = chg if "c" in cd else None
    q = dip if "d" in cd else None
    v, g = bh_direct(orc, nd, src, c, q, trg, 2, oracle=True)
    h = 1e-5
    dv = []
    for a in range(2):
        e = np.zeros((2, 1))
        e[a] = h
        vp, _ = bh_direct(orc, nd, src, c, q, np.asfortranarray(trg + e), 1, oracle=True)
        vm, _ = bh_direct(orc, nd, src, c, q, np.asfortranarray(trg - e), 1, oracle=True)
        dv.append((vp - vm) / (2 * h))
    s = g[:, 0] + g[:, 1]
    assert rel_l2(dv[0], s + g[:, 2]) < 1e-8
    assert rel_l2(dv[1], 1j * (s - g[:, 2])) < 1e-8


def test_skip_coincident(orc):
    src, trg, chg, dip = inputs(1, 20, 3, 2)
    trg = np.asfortranarray(np.concatenate([trg, src[:, :2]], axis=1))
    v, g = bh_direct(orc, 1, src, chg, dip, trg, 2, th=1e-12, oracle=True)
    assert np.isfinite(v).all() and np.isfinite(g).all()


@pytest.mark.parametrize("cd", ["c", "d", "cd"])
@pytest.mark.parametrize("pg", [1, 2])
def test_against_compiled_reference(orc, ref, cd, pg):
    if ref is None:
        pytest.skip("compiled reference not built")
    nd = 3
    src, trg, chg, dip = inputs(nd, 200, 150, 3)
    trg = np.asfortranarray(np.concatenate([trg, src[:, :4]], axis=1))
    c = chg if "c" in cd else None
    q = dip if "d" in cd else None
    a = bh_direct(orc, nd, src, c, q, trg, pg, oracle=True)
    b = bh_direct(ref, nd, src, c, q, trg, pg)
    assert rel_l2(a[0], b[0]) < 1e-13
    if pg >= 2:
        assert rel_l2(a[1], b[1]) < 1e-13
