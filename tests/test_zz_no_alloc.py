"""fmm2dhip_h2d_directcdg_noalloc: the Helmholtz charge + dipole direct sum
(h2d_directcdg_, helmkernels2d.c) with its device memory taken from the
library's fixed 64 MiB scratch array instead of hipMalloc.

Two calls in a row, the second smaller, each against h2d_directcdg_ on the
same inputs: the same kernels with the same source chunking, so the outputs
agree bit for bit.  A request beyond the scratch returns 107 and leaves the
outputs untouched."""
import ctypes as C

import numpy as np
import pytest

import fmm2d_c_translation_amd as F

pytestmark = pytest.mark.gpu

P = lambda a: a.ctypes.data_as(C.c_void_p)
i32 = lambda v: C.byref(C.c_int32(int(v)))


def _problem(rng, ns, nt):
    src = np.asfortranarray(rng.random((2, ns)))
    trg = np.asfortranarray(rng.random((2, nt)))
    ch = np.asfortranarray((rng.standard_normal(ns) + 1j * rng.standard_normal(ns)).reshape(1, ns))
    dp = np.asfortranarray((rng.standard_normal(ns) + 1j * rng.standard_normal(ns)).reshape(1, ns))
    th = rng.random(ns) * 2 * np.pi
    dv = np.asfortranarray(np.stack([np.cos(th), np.sin(th)])[None])
    # outputs are incremented: start both calls from the same non-zero values
    pot0 = np.asfortranarray((rng.standard_normal(nt) + 0.5j).reshape(1, nt))
    grad0 = np.asfortranarray(rng.standard_normal((1, 2, nt)) + 0.25j)
    return src, trg, ch, dp, dv, pot0, grad0


def _call(fn, zk, prob, noalloc):
    src, trg, ch, dp, dv, pot0, grad0 = prob
    pot, grad = pot0.copy(order="F"), grad0.copy(order="F")
    args = [i32(1), (C.c_double * 2)(zk.real, zk.imag), P(src), i32(src.shape[1]), P(ch), P(dp),
            P(dv), P(trg), i32(trg.shape[1]), P(pot), P(grad), C.byref(C.c_double(1e-14))]
    if noalloc:
        fn.restype = C.c_int
        rc = fn(*args, None)
    else:
        rc = 0
        fn(*args)
    return rc, pot, grad


def test_noalloc_matches_h2d_directcdg_twice():
    lib = F.lib()
    rng = np.random.default_rng(21)
    zk = complex(30.0, 0.0)
    for ns, nt in ((150_000, 3000), (20_000, 700)):  # the second call smaller
        print(f"noalloc h2d direct ns={ns} nt={nt}", flush=True)
        prob = _problem(rng, ns, nt)
        rc, pa, ga = _call(lib.fmm2dhip_h2d_directcdg_noalloc, zk, prob, True)
        assert rc == 0, rc
        _, pb, gb = _call(lib.h2d_directcdg_, zk, prob, False)
        assert np.array_equal(pa, pb) and np.array_equal(ga, gb)
        assert np.abs(pa - prob[5]).max() > 0  # something was added


def test_noalloc_rejects_requests_beyond_its_scratch():
    lib = F.lib()
    rng = np.random.default_rng(22)
    prob = _problem(rng, 1_200_000, 1000)  # ~77 MB of staged sources alone
    print("noalloc h2d direct beyond the scratch", flush=True)
    rc, pa, ga = _call(lib.fmm2dhip_h2d_directcdg_noalloc, 5.0 + 0j, prob, True)
    assert rc == 107
    assert np.array_equal(pa, prob[5]) and np.array_equal(ga, prob[6])
    # a call that fits still works afterwards
    small = _problem(rng, 5000, 200)
    rc, pa, ga = _call(lib.fmm2dhip_h2d_directcdg_noalloc, 5.0 + 0j, small, True)
    _, pb, gb = _call(lib.h2d_directcdg_, 5.0 + 0j, small, False)
    assert rc == 0 and np.array_equal(pa, pb) and np.array_equal(ga, gb)
