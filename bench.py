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
P2P_TRAFFIC_BYTES_1M = 780.8e6


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


def eps12_vs_cpu(prob, zk, threads, cpu_sample, steps=3):
    """hfmm2d on the same points at eps = 1e-12 (ndiv 25, the reference's
    leaf-size rule for that eps): GPU throughput against the oracle's OpenMP
    path on the host cores, on the same sample as cpu_baseline."""
    import torch
    ti = np.zeros(8)
    eps, ndiv = 1e-12, 25
    prob.solve("h", eps, zk, ndiv, ti)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        prob.solve("h", eps, zk, ndiv, ti)
    torch.cuda.synchronize()
    gpu = (prob.n + prob.nt) * steps / (time.perf_counter() - t0) / 1e6
    cpu, dt = cpu_baseline(cpu_sample, zk, eps, threads)
    return {"eps": eps, "gpu_mpts_s": round(gpu, 3), "cpu_mpts_s": round(cpu, 4),
            "cores": threads, "ratio": round(gpu / cpu, 1),
            "cpu_sample": f"oracle/ OpenMP restatement at ns=nt={cpu_sample} ({dt:.1f} s)"}


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


def batched_c4(lib, nworkers=8, nprob=512, seed=4):
    """512 independent problems (N lognormal, median 2e4, clipped to
    [2e3, 2e5]; uniform or Gaussian-blob points; half Helmholtz with
    zk ~ U[1, 50], half Cauchy; complex charges; eps = 1e-6; pot + grad at
    sources) solved by fmm2dhip_batch in FP64 and in mixed precision, and one
    at a time through the drop-in calls for comparison."""
    import fmm2d_c_translation_amd as F
    from fmm2d_c_translation_amd import serve
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nprob):
        n = int(np.clip(np.exp(rng.normal(np.log(2e4), 1.0)), 2e3, 2e5))
        src = rng.random((2, n)) if rng.random() < 0.5 else 0.5 + 0.05 * rng.standard_normal((2, n))
        helm = rng.random() < 0.5
        specs.append(dict(kernel=serve.HELMHOLTZ if helm else serve.CAUCHY, eps=1e-6,
                          zk=float(rng.uniform(1, 50)) if helm else 0.0, sources=src,
                          charges=rng.standard_normal(n) + 1j * rng.standard_normal(n), pg=2))
    npts = sum(s["sources"].shape[1] for s in specs)
    out = {"problems": nprob, "points": npts, "workers": nworkers, "eps": 1e-6}
    print(f"batched: {nprob} problems, {npts} points", file=sys.stderr, flush=True)
    results = {}
    for mode, label in ((F.FP64, "fp64"), (F.MIXED, "mixed")):
        probs = [serve.Problem(precision=mode, **s) for s in specs]
        serve.batch(lib, probs[:nworkers], nworkers)  # workers' streams and workspaces
        probs = [serve.Problem(precision=mode, **s) for s in specs]
        t0 = time.perf_counter()
        nbad = serve.batch(lib, probs, nworkers)
        dt = time.perf_counter() - t0
        if nbad:
            raise RuntimeError(f"batched solve: {nbad} problems failed")
        results[label] = probs
        out[label] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4)}
        print(f"batched {label}: {dt:.3f} s", file=sys.stderr, flush=True)
    # largest relative difference of the mixed results from the FP64 ones
    worst = 0.0
    for a, b in zip(results["fp64"], results["mixed"]):
        ra, rb = a.result(), b.result()
        for k in ("pot", "grad"):
            worst = max(worst, float(np.linalg.norm(rb[k] - ra[k]) / np.linalg.norm(ra[k])))
    out["mixed_max_rel_l2_diff_vs_fp64"] = float(f"{worst:.3e}")
    # the same problems one at a time on one thread (FP64, drop-in calls)
    from fmm2d_c_translation_amd import abi
    t0 = time.perf_counter()
    for s in specs:
        if s["kernel"] == serve.HELMHOLTZ:
            abi.hfmm2d_call(lib, "hfmm2d_", 1e-6, s["zk"], s["sources"], charges=s["charges"], pg=2)
        else:
            abi.cfmm2d_call(lib, "cfmm2d_", 1e-6, s["sources"], charges=s["charges"], pg=2)
    dt = time.perf_counter() - t0
    print(f"sequential fp64: {dt:.3f} s", file=sys.stderr, flush=True)
    out["sequential_fp64"] = {"mpts_s": round(npts / dt / 1e6, 3), "seconds": round(dt, 4)}
    out["data"] = "host arrays (PCIe transfers inside the timed region)"
    return out


def served_c5(lib, seconds, rate=200.0, nworkers=8, budget=200e9, seed=5):
    """Poisson arrivals at `rate` requests/s for `seconds`: 60% lfmm2d, 40%
    hfmm2d (zk ~ U[1, 100]); N log-uniform in [1e3, 2e6] uniform in [0,1]^2;
    charges + dipoles N(0,1); eps from {1e-4, 1e-6, 1e-10}; pot + grad at
    sources; HBM budget 200 GB.  Inputs are slices of pre-drawn pools."""
    import threading
    from fmm2d_c_translation_amd import serve
    rng = np.random.default_rng(seed)
    npool, M = 4, 2_000_000
    pools = []
    for _ in range(npool):
        a = rng.random(M) * 2 * np.pi
        pools.append((np.asfortranarray(rng.random((2, M))),
                      rng.standard_normal(M) + 1j * rng.standard_normal(M),
                      rng.standard_normal(M) + 1j * rng.standard_normal(M),
                      np.asfortranarray(np.stack([np.cos(a), np.sin(a)])[None])))
    t_arr = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 10))
    t_arr = t_arr[t_arr < seconds]
    done = []
    pending = []
    cv = threading.Condition()
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
        print(f"served: {len(t_arr)} arrivals over {seconds} s", file=sys.stderr, flush=True)
        next_report = 1.0
        for ta in t_arr:
            if ta >= next_report:
                print(f"served: t={ta:.1f} s, {len(done)} done", file=sys.stderr, flush=True)
                next_report += 1.0
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
            src, ch, dp, dv = pools[int(rng.integers(npool))]
            off = int(rng.integers(0, M - n + 1))
            helm = rng.random() < 0.4
            eps = float(rng.choice([1e-4, 1e-6, 1e-10]))
            p = serve.Problem(kernel=serve.HELMHOLTZ if helm else serve.LOG, eps=eps,
                              zk=float(rng.uniform(1, 100)) if helm else 0.0,
                              sources=src[:, off:off + n], charges=ch[off:off + n],
                              dipstr=dp[off:off + n], dipvec=dv[:, :, off:off + n], pg=2)
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
    return {"seconds": seconds, "rate_req_s": rate, "requests": len(done), "failed": fails,
            "workers": nworkers, "budget_bytes": budget,
            "mpts_s": round(st["mpts_s"], 3), "p50_ms": round(st["p50_s"] * 1e3, 2),
            "p99_ms": round(st["p99_s"] * 1e3, 2), "max_ms": round(st["max_s"] * 1e3, 2),
            "peak_workspace_gb": round(st["peak_workspace_bytes"] / 1e9, 3),
            "data": "host arrays, arrival to result written (PCIe transfers included)"}


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
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-problem batch")
    ap.add_argument("--served-seconds", type=float, default=10.0,
                    help="length of the served request stream (0: skip)")
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
    if not args.no_batch:
        out["batched_c4"] = batched_c4(lib)
    if args.served_seconds > 0:
        out["served_c5"] = served_c5(lib, args.served_seconds)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        v, dt = cpu_baseline(args.cpu_sample, zk, eps, threads)
        out["cpu_baseline"] = {"value": round(v, 4), "unit": "Mpts/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle/ OpenMP restatement, same workload at "
                                         f"ns=nt={args.cpu_sample} ({dt:.1f} s)"}
        out["eps_1e-12_vs_cpu"] = eps12_vs_cpu(prob, zk, threads, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
