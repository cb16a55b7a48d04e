"""Batched and served solves through the library's C ABI (``fmm2dhip_batch``,
``fmm2dhip_server_*``; include/fmm2d_hip.h).

``Problem`` holds one problem's host arrays in the drop-in calls' layouts
(``abi.cfmm2d_call`` and friends) together with the ``fmm2dhip_problem``
struct that points at them; ``Problem.result()`` reads the outputs back in
the same form as the single-problem calls.  ``batch`` solves a list of
problems on the library's worker pool; ``batch_fused`` solves compatible
problems together as one tree (``fmm2dhip_batch_fused``); ``Plan`` captures
repeated solves on fixed points as graphs; ``Server`` is the request loop with
HBM-budget admission, continuous batching and latency statistics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi

CAUCHY, LOG, HELMHOLTZ, REAL = 0, 1, 2, 3
FP64, MIXED = 0, 1


class fmm2dhip_problem(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("nd", C.c_int), ("eps", C.c_double),
        ("zk_re", C.c_double), ("zk_im", C.c_double),
        ("ns", C.c_int), ("sources", C.c_void_p),
        ("ifcharge", C.c_int), ("charge", C.c_void_p),
        ("ifdipole", C.c_int), ("dipstr", C.c_void_p), ("dipvec", C.c_void_p),
        ("ifpgh", C.c_int), ("pot", C.c_void_p), ("grad", C.c_void_p), ("hess", C.c_void_p),
        ("nt", C.c_int), ("targ", C.c_void_p),
        ("ifpghtarg", C.c_int), ("pottarg", C.c_void_p), ("gradtarg", C.c_void_p),
        ("hesstarg", C.c_void_p),
        ("ndiv", C.c_int), ("precision", C.c_int), ("device_ptrs", C.c_int),
        ("ier", C.c_int), ("seconds", C.c_double), ("latency", C.c_double),
    ]


class Problem:
    """One host-memory problem.  kernel: CAUCHY (cfmm2d), LOG (lfmm2d),
    HELMHOLTZ (hfmm2d) or REAL (rfmm2d: float64 strengths and outputs);
    arrays as for the corresponding single call."""

    def __init__(self, kernel, eps, sources, charges=None, dipstr=None, dipvec=None, pg=1,
                 targets=None, pgt=0, nd=1, zk=0.0, ndiv=0, precision=FP64):
        self.kernel, self.nd, self.pg, self.pgt = kernel, nd, pg, pgt
        src = abi._farr(sources, np.float64, (2, -1))
        ns = src.shape[1]
        tg = (abi._farr(targets, np.float64, (2, -1)) if targets is not None
              else np.zeros((2, 1), order="F"))
        nt = tg.shape[1] if targets is not None else 0
        ifc, ifd = int(charges is not None), int(dipstr is not None)
        if kernel == REAL:
            ch = abi._farr(charges, np.float64, (nd, ns)) if ifc else np.zeros((nd, 1), order="F")
            dp = abi._farr(dipstr, np.float64, (nd, ns)) if ifd else np.zeros((nd, 1), order="F")
        else:
            ch = abi._strengths(nd, ns, charges, ifc)
            dp = abi._strengths(nd, ns, dipstr, ifd)
        vec = kernel != CAUCHY
        dv = (abi._farr(dipvec, np.float64, (nd, 2, ns)) if ifd and vec
              else np.zeros((nd, 2, 1), order="F"))
        ncg, nch = (1, 1) if kernel == CAUCHY else (2, 3)
        self.ns, self.nt = ns, nt
        self.out = abi._outs(nd, ns, pg, ncg, nch)
        self.out_t = abi._outs(nd, nt, pgt, ncg, nch)
        if kernel == REAL:
            self.out = tuple(a.real.copy(order="F") for a in self.out)
            self.out_t = tuple(a.real.copy(order="F") for a in self.out_t)
        self._keep = (src, tg, ch, dp, dv)
        p = fmm2dhip_problem()
        p.kernel, p.nd, p.eps = kernel, nd, float(eps)
        p.zk_re, p.zk_im = complex(zk).real, complex(zk).imag
        p.ns, p.sources = ns, src.ctypes.data
        p.ifcharge, p.charge = ifc, ch.ctypes.data
        p.ifdipole, p.dipstr, p.dipvec = ifd, dp.ctypes.data, dv.ctypes.data
        p.ifpgh = pg
        p.pot, p.grad, p.hess = (a.ctypes.data for a in self.out)
        p.nt, p.targ, p.ifpghtarg = nt, tg.ctypes.data, pgt
        p.pottarg, p.gradtarg, p.hesstarg = (a.ctypes.data for a in self.out_t)
        p.ndiv, p.precision, p.device_ptrs = int(ndiv), int(precision), 0
        p.ier = -1
        self.c = p

    @property
    def ier(self):
        return self.c.ier

    def result(self):
        res = abi.FmmResult(ier=self.c.ier, seconds=self.c.seconds, latency=self.c.latency)
        abi._collect(res, "", self.nd, self.ns, self.pg, *self.out)
        abi._collect(res, "targ_", self.nd, self.nt, self.pgt, *self.out_t)
        return res


def batch(lib, problems, nworkers=8):
    """Solve every problem on the library's worker pool; returns the number
    of problems whose ier is not 0."""
    arr = (fmm2dhip_problem * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = p.c
    lib.fmm2dhip_batch.restype = C.c_int
    nbad = lib.fmm2dhip_batch(len(problems), arr, int(nworkers))
    for i, p in enumerate(problems):
        p.c = arr[i]
    return nbad


def batch_fused(lib, problems, nworkers=8):
    """Solve the problems with fmm2dhip_batch_fused: groups of compatible
    problems as one forest each (one launch sequence per group), the rest on
    the worker pool; returns the number of problems whose ier is not 0."""
    arr = (fmm2dhip_problem * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = p.c
    lib.fmm2dhip_batch_fused.restype = C.c_int
    nbad = lib.fmm2dhip_batch_fused(len(problems), arr, int(nworkers))
    for i, p in enumerate(problems):
        p.c = arr[i]
    return nbad


class Plan:
    """fmm2dhip_plan_* over an array of device-memory problem structs (the
    structs and the arrays they point at are kept alive by the caller):
    ``execute()`` re-solves them with the current strengths."""

    def __init__(self, lib, arr):
        self.lib, self.arr = lib, arr
        lib.fmm2dhip_plan_create.restype = C.c_void_p
        lib.fmm2dhip_plan_create.argtypes = [C.c_int, C.c_void_p, C.POINTER(C.c_int)]
        lib.fmm2dhip_plan_execute.restype = C.c_int
        lib.fmm2dhip_plan_execute.argtypes = [C.c_void_p]
        lib.fmm2dhip_plan_destroy.argtypes = [C.c_void_p]
        ier = C.c_int(0)
        self.h = lib.fmm2dhip_plan_create(len(arr), C.cast(arr, C.c_void_p), C.byref(ier))
        self.ier = ier.value
        if not self.h:
            raise RuntimeError(f"fmm2dhip_plan_create failed: ier {self.ier}")

    def execute(self):
        return self.lib.fmm2dhip_plan_execute(self.h)

    def close(self):
        if self.h:
            self.lib.fmm2dhip_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


def autotune(lib, problem, nrep=3):
    """Run the library's autotuner on one Problem; returns the number of
    tile choices changed (negative: -ier)."""
    lib.fmm2dhip_autotune.restype = C.c_int
    return lib.fmm2dhip_autotune(C.byref(problem.c), int(nrep))


def autotune_batch(lib, problems, nrep=3):
    """The autotuner on every fused group of a list of Problems (each group
    timed as one forest); returns the number of tile choices changed."""
    arr = (fmm2dhip_problem * len(problems))()
    for i, p in enumerate(problems):
        arr[i] = p.c
    lib.fmm2dhip_autotune_batch.restype = C.c_int
    return lib.fmm2dhip_autotune_batch(len(problems), arr, int(nrep))


TUNE_KEYS = ("m2l_laplace_fp64", "m2l_laplace_fp32", "m2l_helmholtz_fp64", "m2l_helmholtz_fp32")


def tune_table(lib):
    """The autotuner's table: {"m2l_laplace_fp64": {MT: tiles}, ...,
    "xlate_waves": n}."""
    v = (C.c_int * 37)()
    lib.fmm2dhip_tune_get(v, 37)
    out = {k: {mt: v[9 * j + mt] for mt in range(1, 9)} for j, k in enumerate(TUNE_KEYS)}
    out["xlate_waves"] = v[36]
    return out


STAT_KEYS = ("done", "failed", "points", "span_s", "mpts_s", "p50_s", "p99_s", "max_s",
             "peak_workspace_bytes", "budget_bytes", "tracked_tickets")


class Server:
    """The library's request loop (fmm2dhip_server_*)."""

    def __init__(self, lib, nworkers=8, hbm_budget_bytes=200e9):
        self.lib = lib
        lib.fmm2dhip_server_create.restype = C.c_void_p
        lib.fmm2dhip_server_create.argtypes = [C.c_int, C.c_double]
        lib.fmm2dhip_server_submit.restype = C.c_longlong
        lib.fmm2dhip_server_submit.argtypes = [C.c_void_p, C.POINTER(fmm2dhip_problem)]
        lib.fmm2dhip_server_wait.restype = C.c_int
        lib.fmm2dhip_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.c_double]
        lib.fmm2dhip_server_stats.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        lib.fmm2dhip_server_destroy.argtypes = [C.c_void_p]
        self.h = lib.fmm2dhip_server_create(int(nworkers), float(hbm_budget_bytes))
        if not self.h:
            raise RuntimeError("fmm2dhip_server_create failed")

    def submit(self, problem):
        """Queue a problem (kept alive by the caller until done); returns its ticket."""
        return self.lib.fmm2dhip_server_submit(self.h, C.byref(problem.c))

    def wait(self, ticket, timeout_s=-1.0):
        return self.lib.fmm2dhip_server_wait(self.h, ticket, float(timeout_s))

    def stats(self):
        v = np.zeros(len(STAT_KEYS))
        self.lib.fmm2dhip_server_stats(self.h, v.ctypes.data, len(v))
        return dict(zip(STAT_KEYS, v.tolist()))

    def close(self):
        if self.h:
            self.lib.fmm2dhip_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False
