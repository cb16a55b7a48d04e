"""Plans (fmm2dhip_plan_*): the front end once, the rest of each solve a
captured HIP graph replayed by every execute.

A plan over device arrays is executed three times, with new charges and
dipole strengths written into the same arrays before each execute and other
solves run between executes (the thread's own workspace is reused
meanwhile).  Every replay must equal, bit for bit, fmm2dhip_batch_fused on
the same inputs (the graph replays the kernels of that solve), and the
fused solve is itself checked against direct sums in test_gpu_fused.py.
The mix holds a Cauchy forest, a Helmholtz forest, a log-kernel problem
with dipoles alone in its group, and a complex-wavenumber problem (solved
alone), in FP64 and one group in mixed precision."""
import ctypes as C

import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import serve
from checks import Hip

pytestmark = pytest.mark.gpu


def _mix(seed):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(12):
        n = int(rng.integers(500, 20000))
        src = rng.random((2, n)) if i % 2 else 0.5 + 0.05 * rng.standard_normal((2, n))
        helm = i % 3 == 0
        specs.append(dict(kernel=serve.HELMHOLTZ if helm else serve.CAUCHY, eps=1e-6 if i < 8 else 1e-9,
                          zk=float(rng.uniform(1, 40)) if helm else 0.0, sources=src, pg=2,
                          precision=serve.MIXED if i >= 10 else serve.FP64, dip=False))
    n = 3000
    specs.append(dict(kernel=serve.LOG, eps=1e-6, zk=0.0, sources=rng.random((2, n)), pg=2,
                      precision=serve.FP64, dip=True))
    specs.append(dict(kernel=serve.HELMHOLTZ, eps=1e-6, zk=10 + 0.3j, sources=rng.random((2, n)),
                      pg=2, precision=serve.FP64, dip=False))
    return specs


def _strengths(rng, s):
    n = s["sources"].shape[1]
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dp = rng.standard_normal(n) + 1j * rng.standard_normal(n) if s["dip"] else None
    a = rng.random(n) * 2 * np.pi
    dv = np.asfortranarray(np.stack([np.cos(a), np.sin(a)])[None]) if s["dip"] else None
    return ch, dp, dv


class _DevProblems:
    """Device copies of the specs' points and strength/output buffers, and
    the fmm2dhip_problem array pointing at them."""

    def __init__(self, hip, specs):
        self.hip, self.specs = hip, specs
        self.arr = (serve.fmm2dhip_problem * len(specs))()
        self.bufs = []
        for i, s in enumerate(specs):
            n = s["sources"].shape[1]
            ng = 1 if s["kernel"] == serve.CAUCHY else 2
            b = dict(n=n, ng=ng, src=hip.put(np.asfortranarray(s["sources"])),
                     ch=hip.zeros(16 * n), dp=hip.zeros(16 * n), dv=hip.zeros(16 * n),
                     pot=hip.zeros(16 * n), grad=hip.zeros(16 * n * ng))
            p = self.arr[i]
            p.kernel, p.nd, p.eps = s["kernel"], 1, s["eps"]
            zk = complex(s["zk"])
            p.zk_re, p.zk_im = zk.real, zk.imag
            p.ns, p.sources = n, b["src"]
            p.ifcharge, p.charge = 1, b["ch"]
            p.ifdipole, p.dipstr, p.dipvec = int(s["dip"]), b["dp"], b["dv"]
            p.ifpgh, p.pot, p.grad, p.hess = 2, b["pot"], b["grad"], b["grad"]
            p.nt, p.targ, p.ifpghtarg = 0, b["src"], 0
            p.pottarg = p.gradtarg = p.hesstarg = b["pot"]
            p.ndiv, p.precision, p.device_ptrs = 0, s["precision"], 1
            self.bufs.append(b)

    def load(self, strengths):
        h = self.hip.h
        for b, (ch, dp, dv) in zip(self.bufs, strengths):
            assert h.hipMemcpy(C.c_void_p(b["ch"]), ch.ctypes.data, ch.nbytes, 1) == 0
            if dp is not None:
                assert h.hipMemcpy(C.c_void_p(b["dp"]), dp.ctypes.data, dp.nbytes, 1) == 0
                assert h.hipMemcpy(C.c_void_p(b["dv"]), dv.ctypes.data, dv.nbytes, 1) == 0

    def outputs(self):
        out = []
        for b in self.bufs:
            out.append((self.hip.get(b["pot"], np.zeros(b["n"], complex)),
                        self.hip.get(b["grad"], np.zeros(b["n"] * b["ng"], complex))))
        return out


def _eager(specs, strengths):
    """fmm2dhip_batch_fused on host copies of the same inputs."""
    probs = []
    for s, (ch, dp, dv) in zip(specs, strengths):
        kw = dict(charges=ch, pg=2, precision=s["precision"])
        if dp is not None:
            kw.update(dipstr=dp, dipvec=dv)
        probs.append(serve.Problem(s["kernel"], s["eps"], s["sources"], zk=s["zk"], **kw))
    assert serve.batch_fused(F.lib(), probs, 4) == 0
    out = []
    for p in probs:
        r = p.result()
        out.append((np.asarray(r["pot"]).ravel(), np.asarray(r["grad"]).ravel(order="F")))
    return out


def test_plan_replays_equal_fused_solves():
    lib = F.lib()
    specs = _mix(11)
    hip = Hip()
    try:
        dev = _DevProblems(hip, specs)
        rng = np.random.default_rng(12)
        dev.load([_strengths(rng, s) for s in specs])
        with serve.Plan(lib, dev.arr) as plan:
            print(f"plan: {len(specs)} problems", flush=True)
            for rep in range(3):
                st = [_strengths(rng, s) for s in specs]
                dev.load(st)
                assert plan.execute() == 0
                got = dev.outputs()
                want = _eager(specs, st)  # eager work between the replays
                for i, ((gp, gg), (wp, wg)) in enumerate(zip(got, want)):
                    assert np.array_equal(gp, wp), (rep, i, "pot")
                    assert np.array_equal(gg, wg), (rep, i, "grad")
                print(f"plan replay {rep}: {len(specs)} problems equal the fused solves",
                      flush=True)
    finally:
        hip.free()


def test_plan_rejects_host_arrays():
    lib = F.lib()
    rng = np.random.default_rng(1)
    p = serve.Problem(serve.CAUCHY, 1e-6, rng.random((2, 100)), charges=rng.standard_normal(100) + 0j)
    arr = (serve.fmm2dhip_problem * 1)()
    arr[0] = p.c
    with pytest.raises(RuntimeError, match="ier 102"):
        serve.Plan(lib, arr)
