"""Checkers shared by the GPU tests: FP64 direct sums at a subsample of
evaluation points (the library's GPU c2d/l2d/h2d_direct*g_ entry points,
themselves pinned against the compiled reference and closed forms in
test_gpu_direct.py), and the oracle's FMM on the same problem."""
import ctypes as C

import numpy as np

from fmm2d_c_translation_amd import abi, serve

P = lambda a: a.ctypes.data_as(C.c_void_p)
i32 = lambda v: C.byref(C.c_int32(int(v)))


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def direct_at(lib, spec, pts):
    """pot and grad of every source of spec at pts (2, m) by FP64 direct
    summation: Cauchy grad is d/dz (m,), the others (2, m)."""
    kern = spec["kernel"]
    src = np.asfortranarray(spec["sources"], dtype=np.float64)
    ns, m = src.shape[1], pts.shape[1]
    t = np.asfortranarray(pts, dtype=np.float64)
    ch = spec.get("charges")
    dp = spec.get("dipstr")
    c = np.asfortranarray(np.asarray(ch if ch is not None else np.zeros(ns), complex).reshape(1, ns))
    q = np.asfortranarray(np.asarray(dp if dp is not None else np.zeros(ns), complex).reshape(1, ns))
    dv = spec.get("dipvec")
    v = np.asfortranarray(dv if dv is not None else np.zeros((1, 2, ns)), dtype=np.float64)
    pot = np.zeros((1, m), complex, order="F")
    ncg = 1 if kern == serve.CAUCHY else 2
    grad = np.zeros((1, ncg, m), complex, order="F")
    th = C.byref(C.c_double(1e-14))
    if kern == serve.CAUCHY:
        lib.c2d_directcdg_(i32(1), P(src), i32(ns), P(c), P(q), P(t), i32(m), P(pot), P(grad), th)
        return pot[0], grad[0, 0]
    if kern == serve.HELMHOLTZ:
        zk = complex(spec["zk"])
        lib.h2d_directcdg_(i32(1), (C.c_double * 2)(zk.real, zk.imag), P(src), i32(ns), P(c),
                           P(q), P(v), P(t), i32(m), P(pot), P(grad), th)
    else:
        lib.l2d_directcdg_(i32(1), P(src), i32(ns), P(c), P(q), P(v), P(t), i32(m), P(pot),
                           P(grad), th)
    return pot[0], grad[0]


def errors_vs_direct(lib, spec, res, nsub=1000, seed=0):
    """Relative l2 errors of res (pot/grad at the sources) against direct sums
    at nsub sampled sources.  The Cauchy pot carries a box-dependent gauge in
    its imaginary part (laprouts2d.c:193-224), so for that kernel only the
    gradient is compared."""
    n = spec["sources"].shape[1]
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, min(nsub, n), replace=False))
    pot, grad = direct_at(lib, spec, np.asarray(spec["sources"])[:, idx])
    out = {}
    if spec["kernel"] != serve.CAUCHY:
        out["pot"] = rel_l2(res["pot"][idx], pot)
    out["grad"] = rel_l2(res["grad"][..., idx], grad)
    return out


def oracle_solve(orc, spec):
    """The oracle's FMM (drop-in signature, leaf size by the eps rule) on the
    problem of a spec."""
    kw = dict(sources=spec["sources"], charges=spec.get("charges"), dipstr=spec.get("dipstr"),
              pg=spec.get("pg", 1), targets=spec.get("targets"), pgt=spec.get("pgt", 0))
    k = spec["kernel"]
    if k == serve.CAUCHY:
        return abi.cfmm2d_call(orc, "orc_cfmm2d_", spec["eps"], **kw)
    if k == serve.LOG:
        return abi.lfmm2d_call(orc, "orc_lfmm2d_", spec["eps"], dipvec=spec.get("dipvec"), **kw)
    if k == serve.HELMHOLTZ:
        return abi.hfmm2d_call(orc, "orc_hfmm2d_", spec["eps"], spec["zk"],
                               dipvec=spec.get("dipvec"), **kw)
    raise ValueError(k)


def alone(lib, spec):
    """The spec's problem through its drop-in entry point, in its precision."""
    import fmm2d_c_translation_amd as F
    kw = dict(sources=spec["sources"], charges=spec.get("charges"), dipstr=spec.get("dipstr"),
              pg=spec.get("pg", 1), targets=spec.get("targets"), pgt=spec.get("pgt", 0))
    with F.precision(spec.get("precision", serve.FP64)):
        k = spec["kernel"]
        if k == serve.CAUCHY:
            return abi.cfmm2d_call(lib, "cfmm2d_", spec["eps"], **kw)
        if k == serve.REAL:
            return abi.rfmm2d_call(lib, "rfmm2d_", spec["eps"], dipvec=spec.get("dipvec"), **kw)
        if k == serve.LOG:
            return abi.lfmm2d_call(lib, "lfmm2d_", spec["eps"], dipvec=spec.get("dipvec"), **kw)
        return abi.hfmm2d_call(lib, "hfmm2d_", spec["eps"], spec["zk"],
                               dipvec=spec.get("dipvec"), **kw)


def same(a, b, tag):
    assert a["ier"] == 0 and b["ier"] == 0, (tag, a["ier"], b["ier"])
    for k in ("pot", "grad", "hess", "targ_pot", "targ_grad", "targ_hess"):
        if k in b:
            assert k in a, (tag, k)
            assert np.array_equal(a[k], b[k]), (tag, k)


class Hip:
    """Device buffers through the HIP runtime the library itself links
    (/opt/rocm/lib/libamdhip64.so.7), for tests that hand device pointers to
    the library without a second runtime in the process."""

    def __init__(self):
        import os
        path = "/opt/rocm/lib/libamdhip64.so.7"
        if not os.path.exists(path):
            path = "libamdhip64.so"
        self.h = C.CDLL(path)
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a):
        """A device copy of numpy array a (its bytes as laid out); returns the pointer."""
        a = np.ascontiguousarray(a) if not a.flags.f_contiguous else a
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 16)) == 0
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p.value

    def zeros(self, nbytes):
        return self.put(np.zeros(max(nbytes, 16), np.uint8))

    def get(self, ptr, like):
        """Copy len(like.nbytes) bytes at ptr back into a new array shaped like like."""
        out = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []
