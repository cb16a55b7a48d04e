#!/usr/bin/env python3
"""Throughput benchmark of the MI355X 2D FMM on one GPU.

Headline workload (one "step" = one full FMM solve, inputs already resident
in HBM): Helmholtz hfmm2d, ns = 1,000,000 sources and nt = 1,000,000
separate targets uniform in [0,1]^2 (seed 2), zk = 40, eps = 1e-9, complex
charges N(0,1) plus dipoles with strengths N(0,1) and unit directions at
angles U[0, 2 pi), potential and gradient at sources and targets, FP64.

Prints one JSON line.  ``value`` is points per second over the whole job
((ns + nt) / seconds per solve, in millions), with the headline solve's own
accuracy: relative l2 error against FP64 direct sums at 10^4 sources + 10^4
targets, and against the reference C FMM (compiled from its sources by
oracle/Makefile, where it travelled with the tree) on a bounded sample of
the same distribution.  ``roofline`` describes the near-field (list-1)
kernel, the dominant one, timed with the HIP events the library records
around each FMM step on the stream its kernels run on.  ``cpu_baseline``
times the OpenMP CPU restatement in oracle/ on the host cores, calibrated
against the compiled reference on one core.  Sub-lines cover the other
configurations: C1 (cfmm2d, 1e5 points), C3 (4e6 adaptive points, eps
ladder, FP64 vs mixed, CPU ratio at eps 1e-12), mixed vs FP64 at eps 1e-6 on
the headline points, C4 (512-problem batch, host and device arrays) and C5
(served Poisson stream, FP64 and mixed).
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Algorithmic flops per list-1 pair of the Helmholtz charge+dipole
# potential+gradient kernel (H0 and H1 counted as 120 flops together).
P2P_FLOPS_PER_PAIR = 180.0
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64, vector (the near field's unit) and matrix
HBM_PEAK_GBS = 8000.0
# HBM bytes per near-field launch at the headline workload (ns = nt = 1e6),
# from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE of k_hp2p, each scaled
# by what the same counter reads for a 1 GiB device copy (x2.0 and x1.0):
# profiles/hfmm2d_1M_zk40_eps1e-9_hbm_traffic.md
P2P_TRAFFIC_BYTES_1M = 789.1e6
TRAFFIC_SOURCE = "profiles/hfmm2d_1M_zk40_eps1e-9_hbm_traffic.md"
STEP_NAMES = ["form_mp", "form_lo", "m2m", "m2l", "l2l", "list3", "eval_lo", "p2p"]


def rel(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def e3(v):
    return float(f"{v:.3e}")


def host_cpus():
    """CPUs of this host as lscpu reports them (os.cpu_count() if absent)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("CPU(s):"):
                return int(line.split()[1])
    except Exception:
        pass
    return os.cpu_count() or 1


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def headline_host_arrays(lib, src, trg, ch, dp, dv, zk, eps, ndiv, reps=3):
    """The headline solve through the drop-in hfmm2d_ndiv_ with the caller's
    arrays in host memory (the reference's own interface): the transfers in
    and out are inside the timing.  Outputs are allocated once, outside it;
    the best of reps calls."""
    src = np.asfortranarray(src)
    trg = np.asfortranarray(trg)
    ch = np.ascontiguousarray(ch, dtype=np.complex128)
    dp = np.ascontiguousarray(dp, dtype=np.complex128)
    dv = np.asfortranarray(dv)
    n, nt = src.shape[1], trg.shape[1]
    outs = [np.zeros(n, np.complex128), np.zeros((2, n), np.complex128, order="F"),
            np.zeros(1, np.complex128), np.zeros(nt, np.complex128),
            np.zeros((2, nt), np.complex128, order="F"), np.zeros(1, np.complex128)]
    i32 = lambda v: C.byref(C.c_int32(v))
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    zkc = (C.c_double * 2)(zk, 0.0)
    ti = np.zeros(8)
    best = 1e9
    for _ in range(reps):
        ier, iper = C.c_int32(-1), C.c_int32(0)
        t0 = time.perf_counter()
        lib.hfmm2d_ndiv_(i32(1), C.byref(C.c_double(eps)), zkc, i32(n), p(src), i32(1), p(ch),
                         i32(1), p(dp), p(dv), C.byref(iper), i32(2), p(outs[0]), p(outs[1]),
                         p(outs[2]), i32(nt), p(trg), i32(2), p(outs[3]), p(outs[4]), p(outs[5]),
                         i32(ndiv), i32(0), i32(1), p(ti), C.byref(ier))
        best = min(best, time.perf_counter() - t0)
        assert ier.value == 0, ier.value
    return {"mpts_s": round((n + nt) / best / 1e6, 3), "ms_per_solve": round(best * 1e3, 3),
            "data": "host arrays through hfmm2d_ndiv_ (transfers in and out included); "
                    f"best of {reps}"}


def make_inputs(n, nt, seed=2):
    rng = np.random.default_rng(seed)
    src = rng.random((2, n))
    trg = rng.random((2, nt))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dp = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    th = rng.random(n) * 2 * np.pi
    dv = np.stack([np.cos(th), np.sin(th)])[None]  # (nd=1, 2, n)
    return src, trg, ch, dp, dv


def c1_inputs(n=100_000, seed=1):
    rng = np.random.default_rng(seed)
    return rng.random((2, n)), rng.standard_normal(n) + 1j * rng.standard_normal(n)


def c3_inputs(n=4_000_000, seed=3):
    """90% on a circle of radius 0.4 + N(0, 1e-3) about (0.5, 0.5), 10%
    uniform; real charges N(0,1)."""
    rng = np.random.default_rng(seed)
    th = rng.random(n) * 2 * np.pi
    r = 0.4 + 1e-3 * rng.standard_normal(n)
    src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
    k = n // 10
    src[:, :k] = rng.random((2, k))
    return src, rng.standard_normal(n) + 0j


def lndiv(eps):
    """Leaf size of the reference's rule (src/laplace/lndiv2d.f:17-35)."""
    for lim, v in ((0.5, 3), (0.5e-1, 5), (0.5e-2, 8), (0.5e-3, 10), (0.5e-6, 15),
                   (0.5e-9, 20), (0.5e-12, 25), (0.5e-15, 45)):
        if eps >= lim:
            return v
    raise ValueError(eps)


def set_threads(threads):
    """OpenMP threads of the oracle restatement's later calls (its libgomp
    read OMP_NUM_THREADS once, when it was loaded)."""
    import oracle
    os.environ["OMP_NUM_THREADS"] = str(threads)
    oracle.oracle_lib().omp_set_num_threads(C.c_int(int(threads)))


def time_oracle_hfmm(lib, sym, n_sample, zk, eps, threads):
    """One hfmm2d_ solve of the CPU library lib (the oracle restatement, or
    the compiled reference, which is single-threaded) on n_sample sources +
    n_sample targets of the headline distribution."""
    from fmm2d_c_translation_amd import abi
    set_threads(threads)
    src, trg, ch, dp, dv = make_inputs(n_sample, n_sample, seed=2)
    t0 = time.perf_counter()
    r = abi.hfmm2d_call(lib, sym, eps, zk, src, ch, dp, dv, 2, trg, 2)
    dt = time.perf_counter() - t0
    assert r.ier == 0, r.ier
    return 2 * n_sample / dt / 1e6, dt, r


class DeviceProblem:
    """Inputs resident in HBM in the reference layouts (points (2, n) as an
    (n, 2) row-major array, strengths (nd, n), dipole vectors (nd, 2, n)) and
    output buffers, for the device entry points of the library."""

    def __init__(self, lib, src, trg, ch, dp=None, dv=None):
        import torch
        dev = torch.device("cuda:0")
        self.lib, self.n, self.nt = lib, src.shape[1], (0 if trg is None else trg.shape[1])
        self.src = torch.from_numpy(np.ascontiguousarray(src.T)).to(dev)
        self.trg = (torch.from_numpy(np.ascontiguousarray(trg.T)).to(dev) if trg is not None
                    else torch.zeros((1, 2), dtype=torch.float64, device=dev))
        self.ch = torch.from_numpy(ch.astype(np.complex128)).to(dev)
        self.dp = None if dp is None else torch.from_numpy(dp.astype(np.complex128)).to(dev)
        self.dv = (None if dv is None else
                   torch.from_numpy(np.ascontiguousarray(np.transpose(dv, (2, 1, 0)))).to(dev))
        self.pot = torch.zeros(self.n, dtype=torch.complex128, device=dev)
        self.grad = torch.zeros(2 * self.n, dtype=torch.complex128, device=dev)
        self.pott = torch.zeros(max(self.nt, 1), dtype=torch.complex128, device=dev)
        self.gradt = torch.zeros(2 * max(self.nt, 1), dtype=torch.complex128, device=dev)
        self.stream = torch.cuda.current_stream(dev)

    def solve(self, kernel, eps, zk, ndiv, ti):
        """kernel "h": hfmm2d, "l": lfmm2d (charges + dipoles, pot + grad at
        sources and targets); "c": cfmm2d (charges, pot + grad at sources)."""
        vp, ier = C.c_void_p, C.c_int(0)
        d = lambda t: vp(t.data_ptr())
        pgt = 2 if self.nt else 0
        if kernel == "h":
            self.lib.fmm2dhip_hfmm2d_dev(
                1, C.c_double(eps), C.c_double(zk), C.c_double(0.0), self.n, d(self.src), 1,
                d(self.ch), 1, d(self.dp), d(self.dv), 2, d(self.pot), d(self.grad), None,
                self.nt, d(self.trg), pgt, d(self.pott), d(self.gradt), None, ndiv, 0, 1,
                ti.ctypes.data_as(vp), vp(self.stream.cuda_stream), C.byref(ier))
        elif kernel == "l":
            self.lib.fmm2dhip_lfmm2d_dev(
                1, C.c_double(eps), self.n, d(self.src), 1, d(self.ch), 1, d(self.dp),
                d(self.dv), 2, d(self.pot), d(self.grad), None, self.nt, d(self.trg), pgt,
                d(self.pott), d(self.gradt), None, ndiv, 0, 1, ti.ctypes.data_as(vp),
                vp(self.stream.cuda_stream), C.byref(ier))
        else:
            self.lib.fmm2dhip_cfmm2d_dev(
                1, C.c_double(eps), self.n, d(self.src), 1, d(self.ch), 0, d(self.ch), 2,
                d(self.pot), d(self.grad), None, self.nt, d(self.trg), pgt, d(self.pott),
                d(self.gradt), None, ndiv, 0, 1, ti.ctypes.data_as(vp),
                vp(self.stream.cuda_stream), C.byref(ier))
        if ier.value != 0:
            raise RuntimeError(f"{kernel}fmm2d returned ier={ier.value}")

    def outputs(self, kernel="h"):
        g = self.grad.cpu().numpy()
        g = g[:self.n] if kernel == "c" else g.reshape(self.n, 2).T
        gt = self.gradt.cpu().numpy()
        gt = gt[:self.nt] if kernel == "c" else gt.reshape(-1, 2).T[:, :self.nt]
        return self.pot.cpu().numpy(), g, self.pott.cpu().numpy()[:self.nt], gt


def timed(prob, kernel, eps, zk, ndiv, steps, warmup=3, warm_s=0.3):
    """Seconds per solve over `steps` solves (after at least `warmup` solves
    and `warm_s` seconds of them, so that the GPU clock has ramped up after
    the host-only phases between legs) and the last solve's step times."""
    import torch
    ti = np.zeros(8)
    t0 = time.perf_counter()
    k = 0
    while k < warmup or time.perf_counter() - t0 < warm_s:
        prob.solve(kernel, eps, zk, ndiv, ti)
        torch.cuda.synchronize()
        k += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prob.solve(kernel, eps, zk, ndiv, ti)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, ti


def direct_at(lib, kernel, zk, src, ch, dp, dv, pts):
    """FP64 direct sum (the library's GPU c2d/l2d/h2d_direct*g_) of every
    source at the points pts (2, m): pot (m,), grad ((m,) for the Cauchy
    kernel's d/dz, else (2, m))."""
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    ns, m = src.shape[1], pts.shape[1]
    s = np.asfortranarray(src)
    t = np.asfortranarray(pts)
    c = np.asfortranarray(np.asarray(ch, complex).reshape(1, ns))
    pot = np.zeros((1, m), complex, order="F")
    grad = np.zeros((1, 1 if kernel == "c" else 2, m), complex, order="F")
    nd, th = C.byref(C.c_int32(1)), C.byref(C.c_double(1e-14))
    i32 = lambda v: C.byref(C.c_int32(v))
    if kernel == "c":
        lib.c2d_directcg_(nd, P(s), i32(ns), P(c), P(t), i32(m), P(pot), P(grad), th)
        return pot[0], grad[0, 0]
    q = np.asfortranarray(dp.reshape(1, ns))
    v = np.asfortranarray(dv)
    if kernel == "h":
        lib.h2d_directcdg_(nd, (C.c_double * 2)(zk, 0.0), P(s), i32(ns), P(c), P(q), P(v), P(t),
                           i32(m), P(pot), P(grad), th)
    else:
        lib.l2d_directcdg_(nd, P(s), i32(ns), P(c), P(q), P(v), P(t), i32(m), P(pot), P(grad), th)
    return pot[0], grad[0]


def headline_accuracy(lib, prob, src, trg, ch, dp, dv, zk, nsub=10_000):
    """Relative l2 error of the headline solve (outputs as left by the timed
    loop) against FP64 direct sums at nsub sources and nsub targets."""
    rng = np.random.default_rng(99)
    i_s = np.sort(rng.choice(prob.n, nsub, replace=False))
    i_t = np.sort(rng.choice(prob.nt, nsub, replace=False))
    p, g, pt, gt = prob.outputs()
    ds, dgs = direct_at(lib, "h", zk, src, ch, dp, dv, src[:, i_s])
    dt, dgt = direct_at(lib, "h", zk, src, ch, dp, dv, trg[:, i_t])
    return {"reference": f"FP64 direct sum at {nsub} sources + {nsub} targets",
            "pot": e3(max(rel(p[i_s], ds), rel(pt[i_t], dt))),
            "grad": e3(max(rel(g[:, i_s], dgs), rel(gt[:, i_t], dgt)))}


def reference_leg(lib, zk, eps, n_ref, threads):
    """The compiled reference (single-threaded: the C translation has no
    OpenMP) and the oracle restatement (1 thread and all threads) timed on
    the same n_ref + n_ref sample of the headline distribution, and this
    engine's result on that sample compared with the reference's."""
    import oracle
    from fmm2d_c_translation_amd import abi
    out = {"sample": f"ns=nt={n_ref}, same distribution and seed as the headline",
           "host_cpus_lscpu": host_cpus()}
    src, trg, ch, dp, dv = make_inputs(n_ref, n_ref, seed=2)
    g = abi.hfmm2d_call(lib, "hfmm2d_", eps, zk, src, ch, dp, dv, 2, trg, 2)
    ref = oracle.ref_lib()
    if ref is not None:
        v, dt, r = time_oracle_hfmm(ref, "hfmm2d_", n_ref, zk, eps, 1)
        out["reference_1t"] = {"mpts_s": round(v, 5), "seconds": round(dt, 2)}
        out["gpu_vs_reference_rel_l2"] = {k: e3(rel(g[k], r[k])) for k in
                                          ("pot", "grad", "targ_pot", "targ_grad")}
    else:
        out["reference_1t"] = None
    v1, dt1, r1 = time_oracle_hfmm(oracle.oracle_lib(), "orc_hfmm2d_", n_ref, zk, eps, 1)
    out["port_1t"] = {"mpts_s": round(v1, 5), "seconds": round(dt1, 2)}
    vn, dtn, _ = time_oracle_hfmm(oracle.oracle_lib(), "orc_hfmm2d_", n_ref, zk, eps, threads)
    out["port_all"] = {"mpts_s": round(vn, 5), "seconds": round(dtn, 2), "threads": threads}
    out["gpu_vs_port_rel_l2"] = {k: e3(rel(g[k], r1[k])) for k in
                                 ("pot", "grad", "targ_pot", "targ_grad")}
    if out["reference_1t"]:
        out["port_1t_over_reference_1t_time"] = round(dt1 / out["reference_1t"]["seconds"], 3)
    return out


def c1_leg(lib, steps=50):
    """cfmm2d_ndiv on 1e5 uniform points, complex charges, pot + grad at the
    sources, eps 1e-6, ndiv 40: FP64 and mixed Mpts/s, errors against direct
    sums (gradient: the pot has a box-dependent gauge) and against the
    compiled reference (same tree: pot and grad), CPU times."""
    import fmm2d_c_translation_amd as F
    import oracle
    from fmm2d_c_translation_amd import abi
    src, ch = c1_inputs()
    n = src.shape[1]
    prob = DeviceProblem(lib, src, None, ch)
    eps, ndiv = 1e-6, 40
    rng = np.random.default_rng(11)
    sub = np.sort(rng.choice(n, 10_000, replace=False))
    _, dg = direct_at(lib, "c", 0, src, ch, None, None, src[:, sub])
    out = {"workload": "cfmm2d_ndiv C1: ns=1e5 uniform, complex charges, pot+grad, eps=1e-6, "
                       "ndiv=40", "eps": eps}
    for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
        with F.precision(mode):
            dt, ti = timed(prob, "c", eps, 0.0, ndiv, steps)
        p, g, _, _ = prob.outputs("c")
        out[label] = {"mpts_s": round(n / dt / 1e6, 3), "ms_per_solve": round(dt * 1e3, 3),
                      "grad_rel_l2_vs_direct": e3(rel(g[sub], dg))}
    ref = oracle.ref_lib()
    kw = dict(eps=eps, sources=src, charges=ch, pg=2, ndiv=ndiv)
    a = abi.cfmm2d_call(lib, "cfmm2d_ndiv_", **kw)
    if ref is not None:
        t0 = time.perf_counter()
        b = abi.cfmm2d_call(ref, "cfmm2d_ndiv_", **kw)
        tr = time.perf_counter() - t0
        out["reference_1t"] = {"mpts_s": round(n / tr / 1e6, 4), "seconds": round(tr, 3)}
        out["gpu_vs_reference_rel_l2"] = {k: e3(rel(a[k], b[k])) for k in ("pot", "grad")}
    threads = omp_threads()
    set_threads(threads)
    t0 = time.perf_counter()
    abi.cfmm2d_call(oracle.oracle_lib(), "orc_cfmm2d_ndiv_", **kw)
    tp = time.perf_counter() - t0
    out["port_all"] = {"mpts_s": round(n / tp / 1e6, 4), "seconds": round(tp, 3),
                       "threads": threads}
    return out


def c3_leg(lib, steps=3, with_cpu=True):
    """cfmm2d_ on 4e6 adaptive points (90% on a thin circle), real charges,
    pot + grad at the sources, eps in {1e-3, 1e-6, 1e-9, 1e-12} (leaf size by
    the reference's rule), FP64 and (eps >= 1e-6) mixed: Mpts/s and errors
    against FP64 direct sums at 1e4 sampled points (Re pot, which is
    gauge-free for real charges, and grad); at eps 1e-12 the CPU restatement
    on all host cores on the same 4e6 points."""
    import fmm2d_c_translation_amd as F
    import oracle
    from fmm2d_c_translation_amd import abi
    src, ch = c3_inputs()
    n = src.shape[1]
    prob = DeviceProblem(lib, src, None, ch)
    rng = np.random.default_rng(13)
    sub = np.sort(rng.choice(n, 10_000, replace=False))
    dp_, dg = direct_at(lib, "c", 0, src, ch, None, None, src[:, sub])
    out = {"workload": "cfmm2d C3: 4e6 points, 90% on an r=0.4 circle with 1e-3 jitter, "
                       "real charges, pot+grad", "ladder": {}}
    for eps in (1e-3, 1e-6, 1e-9, 1e-12):
        row = {"ndiv": lndiv(eps)}
        for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
            if mode == F.MIXED and eps < 1e-6:
                continue  # mixed precision is FP64 below eps 1e-6
            with F.precision(mode):
                dt, ti = timed(prob, "c", eps, 0.0, lndiv(eps), steps)
            p, g, _, _ = prob.outputs("c")
            row[label] = {"mpts_s": round(n / dt / 1e6, 3), "ms_per_solve": round(dt * 1e3, 3),
                          "re_pot_rel_l2": e3(rel(p[sub].real, dp_.real)),
                          "grad_rel_l2": e3(rel(g[sub], dg))}
        out["ladder"][f"{eps:.0e}"] = row
    if with_cpu:
        threads = omp_threads()
        set_threads(threads)
        t0 = time.perf_counter()
        r = abi.cfmm2d_call(oracle.oracle_lib(), "orc_cfmm2d_", 1e-12, src, charges=ch, pg=2)
        tc = time.perf_counter() - t0
        assert r.ier == 0
        gpu = out["ladder"]["1e-12"]["fp64"]["mpts_s"]
        cpu = n / tc / 1e6
        out["eps_1e-12_vs_cpu"] = {
            "gpu_mpts_s": gpu, "cpu_mpts_s": round(cpu, 4), "cores": threads,
            "ratio": round(gpu / cpu, 1), "target_ratio": 30,
            "cpu_sample": f"oracle/ OpenMP restatement on the same 4e6 points ({tc:.1f} s)"}
    return out


def mixed_vs_fp64(lib, prob, src, trg, ch, dp, dv, zk, eps=1e-6, steps=5, nsub=10_000):
    """The headline points at eps = 1e-6 with the near field and M2L in FP64
    and in mixed precision: Mpts/s of each and relative l2 errors against
    FP64 direct sums at nsub sources and nsub targets."""
    import fmm2d_c_translation_amd as F
    rng = np.random.default_rng(99)
    i_s = np.sort(rng.choice(prob.n, nsub, replace=False))
    i_t = np.sort(rng.choice(prob.nt, nsub, replace=False))
    out = {"eps": eps, "steps": steps,
           "error_reference": f"FP64 direct sum at {nsub} sources + {nsub} targets"}
    for kernel, name, ndiv in (("h", "hfmm2d zk=40", 15), ("l", "lfmm2d", 15)):
        ds, dgs = direct_at(lib, kernel, zk, src, ch, dp, dv, src[:, i_s])
        dt, dgt = direct_at(lib, kernel, zk, src, ch, dp, dv, trg[:, i_t])
        res = {}
        for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
            with F.precision(mode):
                dt_s, ti = timed(prob, kernel, eps, zk, ndiv, steps)
            p, g, pt, gt = prob.outputs()
            res[label] = {
                "mpts_s": round((prob.n + prob.nt) / dt_s / 1e6, 3),
                "ms_per_solve": round(dt_s * 1e3, 3),
                "step_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(STEP_NAMES, ti)},
                "rel_l2_err": {"pot": e3(max(rel(p[i_s], ds), rel(pt[i_t], dt))),
                               "grad": e3(max(rel(g[:, i_s], dgs), rel(gt[:, i_t], dgt)))},
            }
        res["speedup"] = round(res["mixed"]["mpts_s"] / res["fp64"]["mpts_s"], 3)
        out[name] = res
    return out


def _device_problems(specs, mode):
    """fmm2dhip_problem structs over device copies of the specs' arrays (the
    layouts of the _dev entry points), with the tensors that back them."""
    import torch
    from fmm2d_c_translation_amd import serve
    dev = torch.device("cuda:0")
    keep, arr = [], (serve.fmm2dhip_problem * len(specs))()
    for i, s in enumerate(specs):
        n = s["sources"].shape[1]
        src = torch.from_numpy(np.ascontiguousarray(s["sources"].T)).to(dev)
        ch = torch.from_numpy(np.asarray(s["charges"], np.complex128)).to(dev)
        pot = torch.zeros(n, dtype=torch.complex128, device=dev)
        grad = torch.zeros(n * (1 if s["kernel"] == serve.CAUCHY else 2),
                           dtype=torch.complex128, device=dev)
        keep.append((src, ch, pot, grad))
        p = arr[i]
        p.kernel, p.nd, p.eps = s["kernel"], 1, s["eps"]
        p.zk_re, p.zk_im = s["zk"], 0.0
        p.ns, p.sources = n, src.data_ptr()
        p.ifcharge, p.charge, p.ifdipole, p.dipstr, p.dipvec = 1, ch.data_ptr(), 0, None, None
        p.ifpgh, p.pot, p.grad, p.hess = 2, pot.data_ptr(), grad.data_ptr(), None
        p.nt, p.targ, p.ifpghtarg = 0, src.data_ptr(), 0
        p.ndiv, p.precision, p.device_ptrs, p.ier = 0, mode, 1, -1
    torch.cuda.synchronize()
    return arr, keep


def batched_c4(lib, nworkers=8, autotune=True):
    """The 512 problems of workloads.c4_specs, in FP64 and in mixed
    precision: on the worker pool (fmm2dhip_batch) from host arrays (PCIe
    transfers inside the timed region, as a drop-in caller would pay) and
    from device arrays; fused, one launch sequence per group of compatible
    problems (fmm2dhip_batch_fused), from device and host arrays; as a plan
    (fmm2dhip_plan_*: trees built once, each solve a replay of captured
    graphs); and one at a time through the drop-in calls for comparison."""
    import fmm2d_c_translation_amd as F
    from fmm2d_c_translation_amd import abi, serve, workloads
    specs = workloads.c4_specs()
    npts = sum(s["sources"].shape[1] for s in specs)
    out = {"problems": len(specs), "points": npts, "workers": nworkers, "eps": 1e-6}
    print(f"batched: {len(specs)} problems, {npts} points", file=sys.stderr, flush=True)
    results = {}
    for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
        if autotune:
            # tile choices tuned on the batch's own forests (one per group)
            t0 = time.perf_counter()
            probs = [serve.Problem(precision=mode, **s) for s in specs]
            changed = serve.autotune_batch(lib, probs, nrep=2)
            out[label + "_autotune"] = {"changed": changed,
                                        "seconds": round(time.perf_counter() - t0, 2)}
            print(f"autotune {label}: {changed} changed", file=sys.stderr, flush=True)
        probs = [serve.Problem(precision=mode, **s) for s in specs[:nworkers]]
        serve.batch(lib, probs, nworkers)  # workers' streams and workspaces
        probs = [serve.Problem(precision=mode, **s) for s in specs]
        t0 = time.perf_counter()
        nbad = serve.batch(lib, probs, nworkers)
        dt = time.perf_counter() - t0
        if nbad:
            raise RuntimeError(f"batched solve: {nbad} problems failed")
        results[label] = probs
        out[label] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
                      "data": "host arrays (PCIe transfers inside the timed region)"}
        arr, keep = _device_problems(specs, mode)
        lib.fmm2dhip_batch.restype = C.c_int
        t0 = time.perf_counter()
        nbad = lib.fmm2dhip_batch(len(specs), arr, nworkers)
        dt = time.perf_counter() - t0
        if nbad:
            raise RuntimeError(f"batched device solve: {nbad} problems failed")
        out[label + "_device"] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
                                  "data": "device arrays (resident in HBM)"}
        print(f"batched {label}: host arrays {out[label]['seconds']} s, device arrays "
              f"{dt:.3f} s", file=sys.stderr, flush=True)
        # fused: each group of compatible problems as one forest (one launch
        # sequence per group; here one Cauchy and one Helmholtz group)
        lib.fmm2dhip_batch_fused.restype = C.c_int
        if lib.fmm2dhip_batch_fused(len(specs), arr, nworkers):  # workspace sized once
            raise RuntimeError("fused device solve failed")
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            nbad = lib.fmm2dhip_batch_fused(len(specs), arr, nworkers)
            ts.append(time.perf_counter() - t0)
            if nbad:
                raise RuntimeError(f"fused device solve: {nbad} problems failed")
        dt = min(ts)
        out[label + "_fused_device"] = {
            "mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
            "seconds_all": [round(t, 4) for t in ts],
            "data": "device arrays (resident in HBM), fmm2dhip_batch_fused"}
        # a plan: the same problems with the trees built once and each solve a
        # replay of captured graphs (repeated solves on fixed points)
        t0 = time.perf_counter()
        plan = serve.Plan(lib, arr)
        t_create = time.perf_counter() - t0
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            if plan.execute():
                raise RuntimeError("plan execute failed")
            ts.append(time.perf_counter() - t0)
        plan.close()
        dt = min(ts)
        out[label + "_plan_replay"] = {
            "mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
            "seconds_all": [round(t, 4) for t in ts], "create_seconds": round(t_create, 3),
            "data": "device arrays, fmm2dhip_plan_execute (graph replay; trees built at create)"}
        print(f"plan {label}: create {t_create:.3f} s, replay {dt * 1e3:.2f} ms",
              file=sys.stderr, flush=True)
        del keep
        probs = [serve.Problem(precision=mode, **s) for s in specs]
        t0 = time.perf_counter()
        nbad = serve.batch_fused(lib, probs, nworkers)
        dt = time.perf_counter() - t0
        if nbad:
            raise RuntimeError(f"fused solve: {nbad} problems failed")
        results[label + "_fused"] = probs
        out[label + "_fused"] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
                                 "data": "host arrays (PCIe inside), fmm2dhip_batch_fused"}
        print(f"fused {label}: device arrays {out[label + '_fused_device']['seconds']} s, "
              f"host arrays {dt:.3f} s", file=sys.stderr, flush=True)
    worst = 0.0
    for a, b in zip(results["fp64"], results["mixed"]):
        ra, rb = a.result(), b.result()
        for k in ("pot", "grad"):
            worst = max(worst, rel(rb[k], ra[k]))
    out["mixed_max_rel_l2_diff_vs_fp64"] = e3(worst)
    # fused against the pool's solves (the same problems solved alone):
    # Laplace to rounding, Helmholtz within 2 eps (rescaled, enlarged root)
    worst = {}
    for s, a, b in zip(specs, results["fp64"], results["fp64_fused"]):
        ra, rb = a.result(), b.result()
        key = "helmholtz" if s["kernel"] == serve.HELMHOLTZ else "cauchy"
        for k in ("pot", "grad"):
            worst[key] = max(worst.get(key, 0.0), rel(rb[k], ra[k]))
    out["fused_max_rel_l2_diff_vs_alone_fp64"] = {k: e3(v) for k, v in worst.items()}
    out["tune_table"] = serve.tune_table(lib)
    t0 = time.perf_counter()
    for s in specs:
        if s["kernel"] == serve.HELMHOLTZ:
            abi.hfmm2d_call(lib, "hfmm2d_", 1e-6, s["zk"], s["sources"], charges=s["charges"], pg=2)
        else:
            abi.cfmm2d_call(lib, "cfmm2d_", 1e-6, s["sources"], charges=s["charges"], pg=2)
    dt = time.perf_counter() - t0
    print(f"sequential fp64: {dt:.3f} s", file=sys.stderr, flush=True)
    out["sequential_fp64"] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4),
                              "data": "host arrays, one drop-in call at a time"}
    return out


def served_c5(lib, seconds, precision, nworkers=8, budget=200e9):
    """workloads.C5Stream: Poisson arrivals at 200 requests/s for `seconds`,
    every request in the given precision; arrival to result written in the
    caller's host arrays."""
    import threading
    from fmm2d_c_translation_amd import serve, workloads
    stream = workloads.C5Stream(seconds, precision=precision)
    done, pending = [], []
    cv = threading.Condition()
    label = "mixed" if precision == serve.MIXED else "fp64"
    with serve.Server(lib, nworkers=nworkers, hbm_budget_bytes=budget) as srv:
        def collector():
            while True:
                with cv:
                    while not pending:
                        cv.wait()
                    item = pending.pop(0)
                if item is None:
                    return
                p, t = item
                srv.wait(t, -1.0)
                done.append((p.c.ier, p.c.latency))
        th = threading.Thread(target=collector)
        th.start()
        t0 = time.perf_counter()
        print(f"served {label}: {len(stream)} arrivals over {seconds} s", file=sys.stderr,
              flush=True)
        next_report = 5.0
        for ta in stream.arrivals:
            if ta >= next_report:
                print(f"served {label}: t={ta:.0f} s, {len(done)} done", file=sys.stderr,
                      flush=True)
                next_report += 5.0
            p = serve.Problem(**stream.spec())
            wait = ta - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            t = srv.submit(p)
            with cv:
                pending.append((p, t))
                cv.notify()
        with cv:
            pending.append(None)
            cv.notify()
        th.join()
        st = srv.stats()
    fails = sum(1 for ier, _ in done if ier != 0)
    return {"seconds": seconds, "rate_req_s": 200.0, "requests": len(done), "failed": fails,
            "workers": nworkers, "budget_bytes": budget,
            "mpts_s": round(st["mpts_s"], 3), "p50_ms": round(st["p50_s"] * 1e3, 2),
            "p99_ms": round(st["p99_s"] * 1e3, 2), "max_ms": round(st["max_s"] * 1e3, 2),
            "peak_workspace_gb": round(st["peak_workspace_bytes"] / 1e9, 3),
            "data": "host arrays, arrival to result written (PCIe transfers included)"}


def _stfmm_call(lib, eps, src, f, m, nv, pp):
    """stfmm2d_ on host arrays (its only interface); returns the outputs."""
    nd, ns = f.shape[0], src.shape[1]
    z = lambda *sh: np.zeros(sh, order="F")
    pot, pre, grad = z(nd, 2, ns), z(nd, ns), z(nd, 2, 2, ns)
    tg = z(2, 1)
    dum = z(nd, 2, 1)
    ier = C.c_int32(-1)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    I = lambda v: C.byref(C.c_int32(v))
    lib.stfmm2d_(I(nd), C.byref(C.c_double(eps)), I(ns), P(src), I(1), P(f), I(1), P(m), P(nv),
                 I(pp), P(pot), P(pre), P(grad), I(0), P(tg), I(0), P(dum), P(dum[0]), P(dum),
                 C.byref(ier))
    if ier.value != 0:
        raise RuntimeError(f"stfmm2d_ ier={ier.value}")
    return pot, pre, grad


def stokes_leg(lib, n=1_000_000, n_ref=200_000, steps=5, with_cpu=True):
    """stfmm2d_ (Stokes FMM on the biharmonic engine): n uniform sources with
    Stokeslets and stresslets, velocity + pressure + gradient at the sources,
    eps 1e-9, host arrays (PCIe-inclusive; the entry point takes host
    arrays); and the compiled reference (single thread) on an n_ref sample
    of the same distribution, with this engine's result there compared."""
    import oracle
    rng = np.random.default_rng(21)
    mk = lambda k: (np.asfortranarray(rng.random((2, k))),
                    *(np.asfortranarray(rng.standard_normal((1, 2, k))) for _ in range(3)))
    src, f, m, nv = mk(n)
    eps = 1e-9
    _stfmm_call(lib, eps, src, f, m, nv, 3)
    t0 = time.perf_counter()
    for _ in range(steps):
        _stfmm_call(lib, eps, src, f, m, nv, 3)
    dt = (time.perf_counter() - t0) / steps
    out = {"workload": f"stfmm2d: ns={n} uniform, Stokeslets+stresslets, ifppreg=3, eps={eps}, "
                       "host arrays", "mpts_s": round(n / dt / 1e6, 3),
           "ms_per_solve": round(dt * 1e3, 2)}
    ref = oracle.ref_lib()
    if with_cpu and ref is not None:
        rs, rf, rm, rn = mk(n_ref)
        a = _stfmm_call(lib, eps, rs, rf, rm, rn, 3)
        t0 = time.perf_counter()
        b = _stfmm_call(ref, eps, rs, rf, rm, rn, 3)
        tr = time.perf_counter() - t0
        out["reference_1t"] = {"mpts_s": round(n_ref / tr / 1e6, 4), "seconds": round(tr, 2),
                               "sample": f"ns={n_ref}, same distribution"}
        out["gpu_vs_reference_rel_l2"] = {k: e3(rel(x, y)) for k, x, y in
                                          zip(("pot", "pre", "grad"), a, b)}
        out["gpu_over_reference_1t"] = round(out["mpts_s"] / out["reference_1t"]["mpts_s"], 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000,
                    help="sources (= targets) of the OpenMP CPU baseline's sample")
    ap.add_argument("--ref-sample", type=int, default=50_000,
                    help="sources (= targets) of the compiled-reference calibration sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the eps = 1e-6 mixed-precision vs FP64 comparison")
    ap.add_argument("--no-configs", action="store_true", help="skip the C1 and C3 sub-lines")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-problem batch")
    ap.add_argument("--c4-fused-only", action="store_true",
                    help="only time the 512-problem batch fused from device arrays (FP64, "
                         "--steps solves after --warmup; for kernel profiles of the forests)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in tile choices (no fmm2dhip_autotune on the headline "
                         "workload before timing)")
    ap.add_argument("--served-seconds", type=float, default=60.0,
                    help="length of each served request stream, FP64 and mixed (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (rocprof counter "
                         "calibration: FETCH_SIZE/WRITE_SIZE of a kernel of known traffic)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))
        return
    if args.c4_fused_only:
        import torch
        import fmm2d_c_translation_amd as F
        from fmm2d_c_translation_amd import workloads
        torch.zeros(1, device="cuda:0")
        lib = F.lib()
        specs = workloads.c4_specs()
        npts = sum(s["sources"].shape[1] for s in specs)
        arr, keep = _device_problems(specs, F.FP64)
        lib.fmm2dhip_batch_fused.restype = C.c_int
        ts = []
        for k in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            if lib.fmm2dhip_batch_fused(len(specs), arr, 8):
                raise RuntimeError("fused device solve failed")
            if k >= args.warmup:
                ts.append(time.perf_counter() - t0)
        print(json.dumps({"workload": "C4 fused, device arrays, FP64", "problems": len(specs),
                          "points": npts, "solves": len(ts),
                          "ms_per_batch": round(1e3 * float(np.mean(ts)), 3),
                          "mpts_s": round(npts / float(np.mean(ts)) / 1e6, 3)}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; the multi-GPU path is a "
                                   "design only (see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)

    import torch
    import fmm2d_c_translation_amd as F
    from fmm2d_c_translation_amd import serve

    lib = F.lib()
    rc, arch = F.device_name()
    if rc != 0:
        raise SystemExit(f"no gfx950 device ({arch})")
    zk, eps = 40.0, 1e-9
    n = nt = args.n
    src, trg, ch, dp, dv = make_inputs(n, nt)
    tuned = None
    if not args.no_autotune:
        # the library's autotuner picks the translation kernels' tile shapes
        # on the headline problem itself (its choices hold for every solve)
        tp = serve.Problem(serve.HELMHOLTZ, eps, src, charges=ch, dipstr=dp, dipvec=dv, pg=2,
                           targets=trg, pgt=2, zk=zk)
        t0 = time.perf_counter()
        changed = serve.autotune(lib, tp, nrep=2)
        tuned = {"changed": changed, "seconds": round(time.perf_counter() - t0, 2),
                 "table": serve.tune_table(lib)}
        del tp
        print(f"autotune: {tuned}", file=sys.stderr, flush=True)
    prob = DeviceProblem(lib, src, trg, ch, dp, dv)
    ndiv = 20  # hndiv2d for eps = 1e-9 (c_translation/src/hndiv2d.c:31-47)
    ti = np.zeros(8)

    def solve():
        prob.solve("h", eps, zk, ndiv, ti)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    p2p_s = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
        p2p_s.append(ti[7])
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    print(f"headline: {args.steps} solves in {total:.3f} s", file=sys.stderr, flush=True)
    stats = np.zeros(26)
    lib.fmm2dhip_last_stats(stats.ctypes.data_as(C.c_void_p), 26)
    ms = total / args.steps * 1e3
    value = (n + nt) / (total / args.steps) / 1e6
    pairs = stats[0] + stats[1]
    p2p_time = float(np.mean(p2p_s))
    achieved = pairs * P2P_FLOPS_PER_PAIR / p2p_time / 1e12
    m2l_tf = stats[3] / stats[10] / 1e12 if stats[10] > 0 else None
    # near-field algorithmic bytes (each output point: 16 B coordinates and
    # 48 B pot + grad written; each list-1 source read: 16 B coordinates +
    # 48 B charge, dipole strength and direction), two launches per solve
    p2p_bytes = (n + nt) * 64.0 + (stats[24] + stats[25]) * 64.0
    out = {
        "metric": "Mpts/s at eps 1e-6 mixed vs FP64 GPU FMM; p50/p99 served solve latency",
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform [0,1]^2, seeded)",
        "config": {"workload": "hfmm2d C2: ns=nt=1e6 uniform, zk=40, eps=1e-9, charges+dipoles, "
                               "pot+grad at sources and targets",
                   "ns": n, "nt": nt, "zk": zk, "eps": eps, "ndiv": ndiv,
                   "nboxes": int(stats[4]), "nlevels": int(stats[5]), "pmax": int(stats[6])},
        "step_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(STEP_NAMES, ti)},
        "host_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(
            ["bbox", "keys_sort", "refine", "level_restrict", "boxes_lists_worklists",
             "leaf_order_gather", "operator_setup", "call_total"], stats[15:23])},
        "accuracy": headline_accuracy(lib, prob, src, trg, ch, dp, dv, zk),
        "host_arrays": headline_host_arrays(lib, src, trg, ch, dp, dv, zk, eps, ndiv),
        "roofline": {"kernel": "k_hp2p (list-1 near field, 2 launches per solve)",
                     "bound": "valu",
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "launch_us_avg": round(p2p_time / 2 * 1e6, 1),
                     "pairs_per_launch_avg": pairs / 2, "flops_per_pair": P2P_FLOPS_PER_PAIR,
                     "traffic": P2P_TRAFFIC_BYTES_1M if n == 1_000_000 else None,
                     "traffic_source": TRAFFIC_SOURCE if n == 1_000_000 else None,
                     "algorithmic_bytes_per_launch": round(p2p_bytes / 2),
                     "traffic_over_algorithmic": (round(P2P_TRAFFIC_BYTES_1M / (p2p_bytes / 2), 3)
                                                  if n == 1_000_000 else None),
                     "algorithmic_gbs": round(p2p_bytes / p2p_time / 1e9, 1),
                     "hbm_peak_gbs": HBM_PEAK_GBS,
                     "m2l_tflops": None if m2l_tf is None else round(m2l_tf, 3)},
        "cpu_baseline": None,
        "autotune": tuned,
    }
    if not args.no_mixed:
        out["mixed_vs_fp64"] = mixed_vs_fp64(lib, prob, src, trg, ch, dp, dv, zk)
        print("mixed vs fp64 done", file=sys.stderr, flush=True)
    del prob
    torch.cuda.empty_cache()
    if not args.no_configs:
        out["c1"] = c1_leg(lib)
        print("C1 done", file=sys.stderr, flush=True)
        out["c3"] = c3_leg(lib, with_cpu=not args.no_cpu)
        print("C3 done", file=sys.stderr, flush=True)
        out["stokes"] = stokes_leg(lib, with_cpu=not args.no_cpu)
        print("Stokes done", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_batch:
        out["batched_c4"] = batched_c4(lib, autotune=not args.no_autotune)
    if args.served_seconds > 0:
        out["served_c5"] = served_c5(lib, args.served_seconds, serve.FP64)
        out["served_c5_mixed"] = served_c5(lib, args.served_seconds, serve.MIXED)
    if not args.no_cpu:
        import oracle
        threads = omp_threads()
        v, dt, _ = time_oracle_hfmm(oracle.oracle_lib(), "orc_hfmm2d_", args.cpu_sample, zk, eps,
                                    threads)
        out["cpu_baseline"] = {"value": round(v, 4), "unit": "Mpts/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle/ OpenMP restatement, headline workload at "
                                         f"ns=nt={args.cpu_sample} ({dt:.1f} s)",
                               "host_cpus_lscpu": host_cpus(),
                               "calibration": reference_leg(lib, zk, eps, args.ref_sample,
                                                            threads)}
        out["cpu_baseline"]["gpu_over_cpu"] = round(value / v, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
