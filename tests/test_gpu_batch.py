"""Batched (fmm2dhip_batch) and served (fmm2dhip_server_*) solves.

Every problem of a batch, and every request through the server, must give
the same result as the same problem solved alone through its drop-in entry
point: the same kernels run on the same data whatever the interleaving on
the device, so the comparison is exact.  The mixes cover the three kernels,
uneven sizes, uniform and clustered points, FP64 and mixed precision, and
real and complex Helmholtz wavenumbers.  The server tests also check
the admission bound (workspaces held within the HBM budget), rejection of a
request larger than the budget, and the latency statistics."""
import threading
import time

import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import abi, serve

pytestmark = pytest.mark.gpu


def _problem_specs(seed, count, nmin=500, nmax=30000):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        nt = int(rng.integers(0, n // 2 + 1)) if rng.random() < 0.5 else 0
        if rng.random() < 0.5:
            src = rng.random((2, n))
        else:  # Gaussian blob
            src = np.clip(0.5 + 0.05 * rng.standard_normal((2, n)), 0, 1)
        kernel = int(rng.integers(0, 3))
        eps = float(rng.choice([1e-4, 1e-6, 1e-10]))
        spec = dict(kernel=kernel, eps=eps, sources=src,
                    charges=rng.standard_normal(n) + 1j * rng.standard_normal(n),
                    dipstr=(rng.standard_normal(n) + 1j * rng.standard_normal(n))
                    if rng.random() < 0.5 else None,
                    pg=int(rng.integers(1, 3)),
                    targets=rng.random((2, nt)) if nt else None,
                    pgt=int(rng.integers(1, 3)) if nt else 0,
                    precision=int(rng.integers(0, 2)))
        if kernel != serve.CAUCHY and spec["dipstr"] is not None:
            a = rng.random(n) * 2 * np.pi
            spec["dipvec"] = np.stack([np.cos(a), np.sin(a)])[None]
        if kernel == serve.HELMHOLTZ:
            spec["zk"] = complex(rng.uniform(1, 50), float(rng.choice([0.0, 0.2])))
        specs.append(spec)
    return specs


def _alone(spec):
    """The same problem through its drop-in entry point."""
    lib = F.lib()
    kw = dict(sources=spec["sources"], charges=spec["charges"], dipstr=spec["dipstr"],
              pg=spec["pg"], targets=spec["targets"], pgt=spec["pgt"])
    with F.precision(spec["precision"]):
        if spec["kernel"] == serve.CAUCHY:
            return abi.cfmm2d_call(lib, "cfmm2d_", spec["eps"], **kw)
        if spec["kernel"] == serve.REAL:
            return abi.rfmm2d_call(lib, "rfmm2d_", spec["eps"], dipvec=spec.get("dipvec"), **kw)
        if spec["kernel"] == serve.LOG:
            return abi.lfmm2d_call(lib, "lfmm2d_", spec["eps"], dipvec=spec.get("dipvec"), **kw)
        return abi.hfmm2d_call(lib, "hfmm2d_", spec["eps"], spec["zk"],
                               dipvec=spec.get("dipvec"), **kw)


def _same(a, b, tag):
    assert a["ier"] == 0 and b["ier"] == 0, (tag, a["ier"], b["ier"])
    for k in ("pot", "grad", "hess", "targ_pot", "targ_grad", "targ_hess"):
        if k in b:
            assert k in a, (tag, k)
            assert np.array_equal(a[k], b[k]), (tag, k)


def test_batch_matches_solves_alone():
    specs = _problem_specs(31, 24)
    probs = [serve.Problem(**s) for s in specs]
    print(f"batch of {len(probs)}", flush=True)
    nbad = serve.batch(F.lib(), probs, nworkers=6)
    assert nbad == 0, [p.ier for p in probs]
    for i, (s, p) in enumerate(zip(specs, probs)):
        _same(p.result(), _alone(s), i)
        assert p.c.seconds > 0 and p.c.latency >= p.c.seconds * 0.999


def test_batch_real_kernel_matches_alone():
    """rfmm2d problems (real strengths) in a batch with the other kernels."""
    specs = _problem_specs(34, 8, nmax=8000)
    rng = np.random.default_rng(35)
    for s in specs[::2]:
        n = s["sources"].shape[1]
        s["kernel"] = serve.REAL
        s["charges"] = rng.standard_normal(n)
        if s["dipstr"] is not None:
            s["dipstr"] = rng.standard_normal(n)
            a = rng.random(n) * 2 * np.pi
            s["dipvec"] = np.stack([np.cos(a), np.sin(a)])[None]
        s.pop("zk", None)
    probs = [serve.Problem(**s) for s in specs]
    print(f"batch with real kernels: {len(probs)}", flush=True)
    assert serve.batch(F.lib(), probs, nworkers=4) == 0, [p.ier for p in probs]
    for i, (s, p) in enumerate(zip(specs, probs)):
        r = p.result()
        if s["kernel"] == serve.REAL:
            assert r["pot"].dtype == np.float64
        _same(r, _alone(s), i)


def test_batch_reuses_pool_and_grows_it():
    specs = _problem_specs(32, 6, nmax=5000)
    for nw in (2, 2, 5):
        probs = [serve.Problem(**s) for s in specs]
        assert serve.batch(F.lib(), probs, nworkers=nw) == 0
        for i, (s, p) in enumerate(zip(specs, probs)):
            _same(p.result(), _alone(s), (nw, i))


def test_server_continuous_batching_matches_alone():
    specs = _problem_specs(33, 30)
    probs = [serve.Problem(**s) for s in specs]
    rng = np.random.default_rng(5)
    tickets = [None] * len(probs)
    with serve.Server(F.lib(), nworkers=4, hbm_budget_bytes=60e9) as srv:
        def arrivals():  # requests keep arriving while earlier ones run
            for i, p in enumerate(probs):
                tickets[i] = srv.submit(p)
                time.sleep(rng.exponential(0.002))
        th = threading.Thread(target=arrivals)
        th.start()
        th.join()
        assert all(t is not None and t >= 0 for t in tickets)
        for t in tickets:
            assert srv.wait(t, 120.0) == 0
        st = srv.stats()
    print(st, flush=True)
    assert st["done"] == len(probs) and st["failed"] == 0
    assert 0 < st["p50_s"] <= st["p99_s"] <= st["max_s"]
    assert st["points"] == sum(p.ns + p.nt for p in probs)
    for i, (s, p) in enumerate(zip(specs, probs)):
        _same(p.result(), _alone(s), i)


def test_server_admission_within_budget():
    specs = _problem_specs(34, 12, nmin=20000, nmax=60000)
    probs = [serve.Problem(**s) for s in specs]
    budget = 4e8
    with serve.Server(F.lib(), nworkers=6, hbm_budget_bytes=budget) as srv:
        big = serve.Problem(kernel=serve.HELMHOLTZ, eps=1e-10, zk=30.0,
                            sources=np.random.default_rng(0).random((2, 1_000_000)),
                            charges=np.ones(1_000_000, complex), pg=2)
        assert srv.submit(big) == -1  # alone beyond the budget: rejected
        tickets = [srv.submit(p) for p in probs]
        assert all(t >= 0 for t in tickets)
        for t in tickets:
            assert srv.wait(t, 120.0) == 0
        st = srv.stats()
    print(st, flush=True)
    assert st["done"] == len(probs) and st["failed"] == 0
    assert st["peak_workspace_bytes"] <= budget
    for i, (s, p) in enumerate(zip(specs, probs)):
        _same(p.result(), _alone(s), i)
