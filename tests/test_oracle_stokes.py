"""The oracle's Stokes direct sums (oracle/ostokes.c, restating
stokkernels2d.c:11-173) pinned without the reference: the velocity gradient
against central differences of the velocity, zero divergence (trace of the
gradient), the Stokes equation lap(u) = grad(p) by finite differences, the
Stokeslet against the textbook 2D Green's function
(-log r I + x x^T / r^2) f / 2, and the skip test; and against the compiled
reference where it is built (st2ddirectstokg_, st2ddirectstokstrsg_)."""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2

P = lambda a: a.ctypes.data_as(C.c_void_p)
I = lambda v: C.c_int(v)


def _inputs(nd, ns, nt, seed):
    rng = np.random.default_rng(seed)
    s = np.asfortranarray(rng.random((2, ns)))
    t = np.asfortranarray(rng.random((2, nt)) * 0.6 + 0.2 + 1.5)  # away from the sources
    f = np.asfortranarray(rng.standard_normal((nd, 2, ns)))
    mu = np.asfortranarray(rng.standard_normal((nd, 2, ns)))
    nu = np.asfortranarray(rng.standard_normal((nd, 2, ns)))
    return s, t, f, mu, nu


def orc_st(orc, nd, s, t, f, mu, nu, ifst=1, istr=1, th=1e-14):
    nt = t.shape[1]
    pot = np.zeros((nd, 2, nt), order="F")
    pre = np.zeros((nd, nt), order="F")
    grad = np.zeros((nd, 2, 2, nt), order="F")
    orc.oracle_st2d_direct(I(nd), P(s), I(s.shape[1]), I(ifst), P(f), I(istr), P(mu), P(nu),
                           P(np.asfortranarray(t)), I(nt), P(pot), P(pre), P(grad),
                           C.c_double(th))
    return pot, pre, grad


def ref_st(lib, nd, s, t, f, mu, nu, ifst=1, istr=1, th=1e-14, stok_only=False):
    """Calls the reference-named entry points (the compiled reference or the
    GPU library): st2ddirectstokg_ when stok_only, else st2ddirectstokstrsg_."""
    nt = t.shape[1]
    pot = np.zeros((nd, 2, nt), order="F")
    pre = np.zeros((nd, nt), order="F")
    grad = np.zeros((nd, 2, 2, nt), order="F")
    i32 = lambda v: C.byref(C.c_int32(v))
    t = np.asfortranarray(t)
    if stok_only:
        lib.st2ddirectstokg_(i32(nd), P(s), P(f), i32(s.shape[1]), P(t), i32(nt), P(pot), P(pre),
                             P(grad), C.byref(C.c_double(th)))
    else:
        lib.st2ddirectstokstrsg_(i32(nd), P(s), i32(ifst), P(f), i32(istr), P(mu), P(nu),
                                 i32(s.shape[1]), P(t), i32(nt), P(pot), P(pre), P(grad),
                                 C.byref(C.c_double(th)))
    return pot, pre, grad


@pytest.mark.parametrize("ifst,istr", [(1, 0), (0, 1), (1, 1)])
def test_gradient_divergence_and_stokes_equation(orc, ifst, istr):
    nd = 2
    s, t, f, mu, nu = _inputs(nd, 40, 6, 1)
    u, p, g = orc_st(orc, nd, s, t, f, mu, nu, ifst, istr)
    h = 1e-4
    for a in range(2):
        e = np.zeros((2, 1))
        e[a] = h
        up = orc_st(orc, nd, s, t + e, f, mu, nu, ifst, istr)
        um = orc_st(orc, nd, s, t - e, f, mu, nu, ifst, istr)
        du = (up[0] - um[0]) / (2 * h)  # d u_b / d x_a, (nd, 2 (b), nt)
        assert rel_l2(g[:, a, :, :], du) < 1e-7, a
    # divergence-free velocity: g(1,1) + g(2,2) = 0
    assert np.abs(g[:, 0, 0] + g[:, 1, 1]).max() < 1e-10 * np.abs(g).max()
    # lap(u) = grad(p), both by five-point differences
    h = 1e-3
    lap = -4 * u
    dp = np.zeros_like(u)
    for a in range(2):
        e = np.zeros((2, 1))
        e[a] = h
        up = orc_st(orc, nd, s, t + e, f, mu, nu, ifst, istr)
        um = orc_st(orc, nd, s, t - e, f, mu, nu, ifst, istr)
        lap = lap + up[0] + um[0]
        dp[:, a] = (up[1] - um[1]) / (2 * h)
    lap /= h * h
    assert rel_l2(lap, dp) < 1e-4


def test_stokeslet_green_function_and_skip(orc):
    rng = np.random.default_rng(2)
    s = np.asfortranarray(rng.random((2, 3)))
    t = np.asfortranarray(np.concatenate([rng.random((2, 4)), s[:, :1]], axis=1))
    f = np.asfortranarray(rng.standard_normal((1, 2, 3)))
    u, p, _ = orc_st(orc, 1, s, t, f, f, f, 1, 0, th=1e-12)
    ue = np.zeros((2, t.shape[1]))
    pe = np.zeros(t.shape[1])
    for i in range(t.shape[1]):
        for j in range(3):
            x = t[:, i] - s[:, j]
            r2 = x @ x
            if r2 < 1e-24:
                continue
            G = -0.5 * np.log(np.sqrt(r2)) * np.eye(2) + np.outer(x, x) / (2 * r2)
            ue[:, i] += G @ f[0, :, j]
            pe[i] += x @ f[0, :, j] / r2
    assert rel_l2(u[0], ue) < 1e-14
    assert rel_l2(p[0], pe) < 1e-14
    assert np.isfinite(u).all()


@pytest.mark.parametrize("ifst,istr,stok_only", [(1, 0, True), (1, 0, False), (0, 1, False),
                                                 (1, 1, False), (0, 0, False)])
def test_against_compiled_reference(orc, ref, ifst, istr, stok_only):
    if ref is None:
        pytest.skip("compiled reference not built")
    nd = 3
    s, t, f, mu, nu = _inputs(nd, 300, 200, 3)
    t = np.asfortranarray(np.concatenate([t - 1.5, s[:, :5]], axis=1))  # coincident pairs
    a = orc_st(orc, nd, s, t, f, mu, nu, ifst, istr)
    b = ref_st(ref, nd, s, t, f, mu, nu, ifst, istr, stok_only=stok_only)
    for x, y in zip(a, b):
        assert rel_l2(x, y) < 1e-13 if np.any(y) else not np.any(x)
