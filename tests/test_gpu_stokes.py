"""GPU Stokes direct sums behind st2ddirectstokg_ and st2ddirectstokstrsg_
(stokkernels2d.c:11-173) against the oracle (ostokes.c) and the compiled
reference where it is built: Stokeslets, stresslets and both, nd 1-3,
coincident pairs (skipped), the source-chunked launch (few targets, many
sources) and incremented outputs.  Tolerance 1e-12 relative l2 (float64
sums in a different order)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2
from test_oracle_stokes import _inputs, orc_st, ref_st

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.mark.parametrize("nd,ns,nt", [(1, 2000, 1500), (2, 700, 900), (3, 20000, 40)])
@pytest.mark.parametrize("ifst,istr", [(1, 0), (0, 1), (1, 1)])
def test_stokes_direct(eng, orc, ref, nd, ns, nt, ifst, istr):
    print(f"st2d nd={nd} ns={ns} nt={nt} stok={ifst} strs={istr}", flush=True)
    s, t, f, mu, nu = _inputs(nd, ns, nt, 10 + nd)
    t = np.asfortranarray(np.concatenate([t - 1.5, s[:, :3]], axis=1))
    a = ref_st(eng, nd, s, t, f, mu, nu, ifst, istr)
    b = orc_st(orc, nd, s, t, f, mu, nu, ifst, istr)
    for x, y in zip(a, b):
        assert np.isfinite(x).all()
        assert rel_l2(x, y) < TOL
    if ref is not None:
        c = ref_st(ref, nd, s, t, f, mu, nu, ifst, istr)
        for x, y in zip(a, c):
            assert rel_l2(x, y) < TOL
    if ifst == 1 and istr == 0:
        g = ref_st(eng, nd, s, t, f, mu, nu, stok_only=True)
        for x, y in zip(g, b):
            assert rel_l2(x, y) < TOL


def test_stokes_increments_and_flags(eng, orc):
    """Outputs are incremented; flags other than 1 add nothing."""
    nd, ns, nt = 2, 500, 300
    s, t, f, mu, nu = _inputs(nd, ns, nt, 21)
    rng = np.random.default_rng(22)
    pot = np.asfortranarray(rng.standard_normal((nd, 2, nt)))
    pre = np.asfortranarray(rng.standard_normal((nd, nt)))
    grad = np.asfortranarray(rng.standard_normal((nd, 2, 2, nt)))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    i32 = lambda v: C.byref(C.c_int32(v))
    outs = [pot.copy(order="F"), pre.copy(order="F"), grad.copy(order="F")]
    eng.st2ddirectstokstrsg_(i32(nd), P(s), i32(1), P(f), i32(1), P(mu), P(nu), i32(ns), P(t),
                             i32(nt), *[P(o) for o in outs], C.byref(C.c_double(1e-14)))
    b = orc_st(orc, nd, s, t, f, mu, nu)
    for o, init, y in zip(outs, (pot, pre, grad), b):
        assert rel_l2(o - init, y) < TOL
    outs = [pot.copy(order="F"), pre.copy(order="F"), grad.copy(order="F")]
    eng.st2ddirectstokstrsg_(i32(nd), P(s), i32(2), P(f), i32(0), P(mu), P(nu), i32(ns), P(t),
                             i32(nt), *[P(o) for o in outs], C.byref(C.c_double(1e-14)))
    for o, init in zip(outs, (pot, pre, grad)):
        assert np.array_equal(o, init)
