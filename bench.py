#!/usr/bin/env python3
"""Throughput benchmark of the MI355X 2D FMM on one GPU.

Workload (one "step" = one full FMM solve, inputs already resident in HBM):
Helmholtz hfmm2d, ns = 1,000,000 sources and nt = 1,000,000 separate targets
uniform in [0,1]^2 (seed 2), zk = 40, eps = 1e-9, complex charges N(0,1) plus
dipoles with strengths N(0,1) and unit directions at angles U[0, 2 pi),
potential and gradient at sources and targets, FP64.

Prints one JSON line.  ``value`` is points per second over the whole job
((ns + nt) / seconds per solve, in millions); ``roofline`` describes the
near-field (list-1) kernel, the dominant one, from HIP events recorded by
the library around each FMM step on the stream its kernels run on;
``cpu_baseline`` times the OpenMP CPU restatement in oracle/ on a bounded
sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# Algorithmic flops per list-1 pair of the Helmholtz charge+dipole
# potential+gradient kernel (H0 and H1 counted as 120 flops together).
P2P_FLOPS_PER_PAIR = 180.0
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0
# HBM bytes per near-field launch at the default workload (ns = nt = 1e6):
# rocprofv3 FETCH_SIZE and WRITE_SIZE of k_hp2p, each scaled by what the same
# counter reads for a 1 GiB device copy (x2.0 and x1.0); the commands and
# numbers are in profiles/hfmm2d_1M_zk40_eps1e-9_hbm_traffic.md
P2P_TRAFFIC_BYTES_1M = 747.2e6


def make_inputs(n, nt, seed=2):
    rng = np.random.default_rng(seed)
    src = rng.random((2, n))
    trg = rng.random((2, nt))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dp = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    th = rng.random(n) * 2 * np.pi
    dv = np.stack([np.cos(th), np.sin(th)])[None]  # (nd=1, 2, n)
    return src, trg, ch, dp, dv


def cpu_baseline(n_sample, zk, eps, threads):
    """Time the oracle (C restatement, OpenMP over boxes) on n_sample sources
    + n_sample targets of the same distribution; points/s scale linearly."""
    import oracle
    from fmm2d_c_translation_amd import abi
    os.environ["OMP_NUM_THREADS"] = str(threads)
    src, trg, ch, dp, dv = make_inputs(n_sample, n_sample, seed=2)
    t0 = time.perf_counter()
    r = abi.hfmm2d_call(oracle.oracle_lib(), "orc_hfmm2d_", eps, zk, src, ch, dp, dv, 2, trg, 2)
    dt = time.perf_counter() - t0
    assert r.ier == 0, r.ier
    return 2 * n_sample / dt / 1e6, dt


class DeviceProblem:
    """Inputs resident in HBM in the reference layouts (points (2, n) as an
    (n, 2) row-major array, strengths (nd, n), dipole vectors (nd, 2, n)) and
    output buffers, for the device entry points of the library."""

    def __init__(self, lib, src, trg, ch, dp, dv):
        import torch
        dev = torch.device("cuda:0")
        self.lib, self.n, self.nt = lib, src.shape[1], trg.shape[1]
        self.src = torch.from_numpy(np.ascontiguousarray(src.T)).to(dev)
        self.trg = torch.from_numpy(np.ascontiguousarray(trg.T)).to(dev)
        self.ch = torch.from_numpy(ch.astype(np.complex128)).to(dev)
        self.dp = torch.from_numpy(dp.astype(np.complex128)).to(dev)
        self.dv = torch.from_numpy(np.ascontiguousarray(np.transpose(dv, (2, 1, 0)))).to(dev)
        self.pot = torch.zeros(self.n, dtype=torch.complex128, device=dev)
        self.grad = torch.zeros(2 * self.n, dtype=torch.complex128, device=dev)
        self.pott = torch.zeros(self.nt, dtype=torch.complex128, device=dev)
        self.gradt = torch.zeros(2 * self.nt, dtype=torch.complex128, device=dev)
        self.stream = torch.cuda.current_stream(dev)

    def solve(self, kernel, eps, zk, ndiv, ti):
        """kernel "h": hfmm2d, "l": lfmm2d; charges + dipoles, pot + grad at
        sources and targets."""
        vp, ier = C.c_void_p, C.c_int(0)
        d = lambda t: vp(t.data_ptr())
        if kernel == "h":
            self.lib.fmm2dhip_hfmm2d_dev(
                1, C.c_double(eps), C.c_double(zk), C.c_double(0.0), self.n, d(self.src), 1,
                d(self.ch), 1, d(self.dp), d(self.dv), 2, d(self.pot), d(self.grad), None,
                self.nt, d(self.trg), 2, d(self.pott), d(self.gradt), None, ndiv, 0, 1,
                ti.ctypes.data_as(vp), vp(self.stream.cuda_stream), C.byref(ier))
        else:
            self.lib.fmm2dhip_lfmm2d_dev(
                1, C.c_double(eps), self.n, d(self.src), 1, d(self.ch), 1, d(self.dp),
                d(self.dv), 2, d(self.pot), d(self.grad), None, self.nt, d(self.trg), 2,
                d(self.pott), d(self.gradt), None, ndiv, 0, 1, ti.ctypes.data_as(vp),
                vp(self.stream.cuda_stream), C.byref(ier))
        if ier.value != 0:
            raise RuntimeError(f"{kernel}fmm2d returned ier={ier.value}")

    def outputs(self):
        return (self.pot.cpu().numpy(), self.grad.cpu().numpy().reshape(self.n, 2).T,
                self.pott.cpu().numpy(), self.gradt.cpu().numpy().reshape(self.nt, 2).T)


def direct_subsample(lib, kernel, zk, src, ch, dp, dv, pts):
    """FP64 direct sum (the library's GPU l2d/h2d_directcdg_) of every source
    at the points pts (2, m): pot (m,), grad (2, m)."""
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    ns, m = src.shape[1], pts.shape[1]
    s = np.asfortranarray(src)
    t = np.asfortranarray(pts)
    c = np.asfortranarray(ch.reshape(1, ns))
    q = np.asfortranarray(dp.reshape(1, ns))
    v = np.asfortranarray(dv)
    pot = np.zeros((1, m), complex, order="F")
    grad = np.zeros((1, 2, m), complex, order="F")
    nd = C.byref(C.c_int32(1))
    if kernel == "h":
        lib.h2d_directcdg_(nd, (C.c_double * 2)(zk, 0.0), P(s), C.byref(C.c_int32(ns)), P(c),
                           P(q), P(v), P(t), C.byref(C.c_int32(m)), P(pot), P(grad),
                           C.byref(C.c_double(1e-14)))
    else:
        lib.l2d_directcdg_(nd, P(s), C.byref(C.c_int32(ns)), P(c), P(q), P(v), P(t),
                           C.byref(C.c_int32(m)), P(pot), P(grad), C.byref(C.c_double(1e-14)))
    return pot[0], grad[0]


def mixed_vs_fp64(lib, prob, src, trg, ch, dp, dv, zk, eps=1e-6, steps=5, nsub=10_000):
    """The same points at eps = 1e-6 with the near field and M2L in FP64 and in
    mixed precision: Mpts/s of each and relative l2 errors against FP64 direct
    sums at nsub sources and nsub targets (fixed random subsample)."""
    import torch
    import fmm2d_c_translation_amd as F
    rng = np.random.default_rng(99)
    i_s = np.sort(rng.choice(prob.n, nsub, replace=False))
    i_t = np.sort(rng.choice(prob.nt, nsub, replace=False))
    out = {"eps": eps, "steps": steps,
           "error_reference": f"FP64 direct sum at {nsub} sources + {nsub} targets"}
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    for kernel, name, ndiv in (("h", "hfmm2d zk=40", 15), ("l", "lfmm2d", 15)):
        ds, dgs = direct_subsample(lib, kernel, zk, src, ch, dp, dv, src[:, i_s])
        dt, dgt = direct_subsample(lib, kernel, zk, src, ch, dp, dv, trg[:, i_t])
        res = {}
        for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
            ti = np.zeros(8)
            with F.precision(mode):
                prob.solve(kernel, eps, zk, ndiv, ti)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    prob.solve(kernel, eps, zk, ndiv, ti)
                torch.cuda.synchronize()
                dt_s = (time.perf_counter() - t0) / steps
            p, g, pt, gt = prob.outputs()
            res[label] = {
                "mpts_s": round((prob.n + prob.nt) / dt_s / 1e6, 3),
                "ms_per_solve": round(dt_s * 1e3, 3),
                "step_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(
                    ["form_mp", "form_lo", "m2m", "m2l", "l2l", "list3", "eval_lo", "p2p"], ti)},
                "rel_l2_err": {"pot": float(f"{max(rel(p[i_s], ds), rel(pt[i_t], dt)):.3e}"),
                               "grad": float(f"{max(rel(g[:, i_s], dgs), rel(gt[:, i_t], dgt)):.3e}")},
            }
        res["speedup"] = round(res["mixed"]["mpts_s"] / res["fp64"]["mpts_s"], 3)
        out[name] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the eps = 1e-6 mixed-precision vs FP64 comparison")
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
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; the multi-GPU path is a "
                                   "design only (see DESIGN.md), so --gpus must be 1"}))
        sys.exit(2)

    import torch
    import fmm2d_c_translation_amd as F

    lib = F.lib()
    rc, arch = F.device_name()
    if rc != 0:
        raise SystemExit(f"no gfx950 device ({arch})")
    zk, eps = 40.0, 1e-9
    n = nt = args.n
    src, trg, ch, dp, dv = make_inputs(n, nt)
    prob = DeviceProblem(lib, src, trg, ch, dp, dv)
    ndiv = 20  # hndiv2d for eps = 1e-9 (c_translation/src/hndiv2d.c:31-47)
    ti = np.zeros(8)

    def solve():
        prob.solve("h", eps, zk, ndiv, ti)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    steps = []
    p2p_s = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        solve()
        steps.append(time.perf_counter() - a)
        p2p_s.append(ti[7])
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    stats = np.zeros(23)
    vp = C.c_void_p
    lib.fmm2dhip_last_stats(stats.ctypes.data_as(vp), 23)
    ms = total / args.steps * 1e3
    value = (n + nt) / (total / args.steps) / 1e6
    pairs = stats[0] + stats[1]
    p2p_time = float(np.mean(p2p_s))
    achieved = pairs * P2P_FLOPS_PER_PAIR / p2p_time / 1e12
    m2l_tf = stats[3] / stats[10] / 1e12 if stats[10] > 0 else None
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
        "step_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(
            ["form_mp", "form_lo", "m2m", "m2l", "l2l", "list3", "eval_lo", "p2p"], ti)},
        "host_ms": {k: round(float(v) * 1e3, 3) for k, v in zip(
            ["bbox", "keys_sort", "refine", "level_restrict", "boxes_lists_worklists",
             "leaf_order_gather", "operator_setup", "call_total"], stats[15:23])},
        "roofline": {"kernel": "k_hp2p (list-1 near field)", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "traffic": P2P_TRAFFIC_BYTES_1M if n == 1_000_000 else None,
                     "pairs_per_launch_avg": pairs / 2, "flops_per_pair": P2P_FLOPS_PER_PAIR,
                     "m2l_tflops": None if m2l_tf is None else round(m2l_tf, 3)},
        "cpu_baseline": None,
    }
    if not args.no_mixed:
        out["mixed_vs_fp64"] = mixed_vs_fp64(lib, prob, src, trg, ch, dp, dv, zk)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v, dt = cpu_baseline(args.cpu_sample, zk, eps, threads)
        out["cpu_baseline"] = {"value": round(v, 4), "unit": "Mpts/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle/ OpenMP restatement, same workload at "
                                         f"ns=nt={args.cpu_sample} ({dt:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
