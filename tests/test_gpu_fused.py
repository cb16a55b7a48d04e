"""Fused batches (fmm2dhip_batch_fused): groups of problems solved as one
forest, one level-synchronous launch sequence per group.

Checked three ways per problem:
  1. against the same problem solved alone through its drop-in entry point:
     Laplace kernels (cfmm2d, lfmm2d, rfmm2d) keep their own tree, centres
     and box sizes in the forest, so they agree to rounding (1e-12 relative
     l2, the Cauchy potential's box-dependent gauge included); a Helmholtz
     problem is rescaled to the group's wavenumber with its root box enlarged
     (a different tree): both results are within eps of the direct sums
     (check 2), so they agree within 2 eps;
  2. within eps of FP64 direct sums at sampled sources (checks.py);
  3. problems that cannot join a forest (no more sources than the leaf
     size, complex wavenumbers, a group of one) go to the worker pool and
     equal their solve alone bit for bit.
The batched workload (512 problems, workloads.c4_specs) runs fused too."""
import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import serve, workloads
import ctypes as C

from checks import Hip, alone, errors_vs_direct, rel_l2, same

pytestmark = pytest.mark.gpu
LAP_TOL = 1e-12


def _specs(seed, count, kernels, eps_list, nmin=200, nmax=30000, targets=False, dipoles=False):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        if rng.random() < 0.5:
            src = rng.random((2, n)) * rng.uniform(0.2, 3.0) + rng.uniform(-2, 2, (2, 1))
        else:
            src = rng.uniform(-1, 1, (2, 1)) + rng.uniform(0.01, 0.2) * rng.standard_normal((2, n))
        kern = int(rng.choice(kernels))
        s = dict(kernel=kern, eps=float(rng.choice(eps_list)), sources=src, pg=2,
                 charges=rng.standard_normal(n) + 1j * rng.standard_normal(n))
        if kern == serve.REAL:
            s["charges"] = rng.standard_normal(n)
        if dipoles:
            s["dipstr"] = (rng.standard_normal(n) if kern == serve.REAL
                           else rng.standard_normal(n) + 1j * rng.standard_normal(n))
            if kern != serve.CAUCHY:
                a = rng.random(n) * 2 * np.pi
                s["dipvec"] = np.stack([np.cos(a), np.sin(a)])[None]
        if targets:
            nt = int(rng.integers(1, n))
            s["targets"] = src[:, :1] + (src.max(1, keepdims=True) - src.min(1, keepdims=True)) * (
                rng.random((2, nt)) - 0.3)
            s["pgt"] = 2
        if kern == serve.HELMHOLTZ:
            s["zk"] = float(rng.uniform(1, 50))
        out.append(s)
    return out


def _run_fused(specs):
    probs = [serve.Problem(**s) for s in specs]
    nbad = serve.batch_fused(F.lib(), probs, nworkers=4)
    assert nbad == 0, [p.ier for p in probs]
    return [p.result() for p in probs]


def _check_vs_alone(specs, res, tag):
    lib = F.lib()
    worst = 0.0
    for i, s in enumerate(specs):
        a = alone(lib, s)
        assert a["ier"] == 0
        tol = LAP_TOL if s["kernel"] != serve.HELMHOLTZ else 2 * s["eps"]
        for k in ("pot", "grad", "targ_pot", "targ_grad"):
            if k in a:
                e = rel_l2(res[i][k], a[k])
                worst = max(worst, e / tol)
                assert e < tol, (tag, i, k, e, s["kernel"], s["sources"].shape[1])
    print(f"{tag}: {len(specs)} problems, worst error vs alone {worst:.2e} of the tolerance",
          flush=True)


def _check_vs_direct(specs, res, tag):
    lib = F.lib()
    worst = 0.0
    for i, s in enumerate(specs):
        if s["kernel"] == serve.REAL:
            continue
        e = errors_vs_direct(lib, s, res[i], nsub=300, seed=i)
        for k, v in e.items():
            worst = max(worst, v / s["eps"])
            assert v < s["eps"], (tag, i, k, v, s["kernel"], s["eps"])
    print(f"{tag}: worst error vs direct sums {worst:.2e} eps", flush=True)


@pytest.mark.parametrize("seed", [1, 2])
def test_fused_cauchy(seed):
    specs = _specs(seed, 24, [serve.CAUCHY], [1e-6, 1e-10], targets=seed == 2)
    res = _run_fused(specs)
    _check_vs_direct(specs, res, f"cauchy seed {seed}")
    _check_vs_alone(specs, res, f"cauchy seed {seed}")


def test_fused_log_and_real_with_dipoles():
    specs = _specs(3, 20, [serve.LOG, serve.REAL], [1e-6], dipoles=True, targets=True)
    res = _run_fused(specs)
    _check_vs_direct(specs, res, "log/real")
    _check_vs_alone(specs, res, "log/real")


@pytest.mark.parametrize("eps", [1e-6, 1e-9])
def test_fused_helmholtz(eps):
    specs = _specs(4, 20, [serve.HELMHOLTZ], [eps], dipoles=True, targets=eps == 1e-9)
    res = _run_fused(specs)
    _check_vs_direct(specs, res, f"helmholtz eps {eps}")
    _check_vs_alone(specs, res, f"helmholtz eps {eps}")


def test_fused_mixed_precision():
    specs = _specs(5, 16, [serve.CAUCHY, serve.HELMHOLTZ], [1e-6])
    for s in specs:
        s["precision"] = serve.MIXED
    res = _run_fused(specs)
    _check_vs_direct(specs, res, "mixed")


def test_unfusable_problems_go_to_the_pool():
    rng = np.random.default_rng(6)
    specs = _specs(6, 6, [serve.CAUCHY, serve.HELMHOLTZ], [1e-6])
    # too few sources for a forest, a complex wavenumber, a lone problem
    specs.append(dict(kernel=serve.CAUCHY, eps=1e-6, sources=rng.random((2, 10)), pg=2,
                      charges=rng.standard_normal(10) + 0j))
    specs.append(dict(kernel=serve.HELMHOLTZ, eps=1e-6, zk=10 + 0.5j, sources=rng.random((2, 3000)),
                      pg=2, charges=rng.standard_normal(3000) + 0j))
    specs.append(dict(kernel=serve.LOG, eps=1e-4, sources=rng.random((2, 4000)), pg=1,
                      charges=rng.standard_normal(4000) + 0j))
    res = _run_fused(specs)
    lib = F.lib()
    for i in (6, 7, 8):
        same(res[i], alone(lib, specs[i]), i)
    _check_vs_alone(specs[:6], res[:6], "fusable part")


def test_fused_deep_clusters_sort_again():
    """Problems whose trees run deeper than the forest's first point sort
    covers (a cluster of sigma 1e-4 inside a uniform background: some 20
    levels below the root, past the estimate from the point counts), so the
    forest sorts its points again on every bit; the results still equal the
    solves alone and are within eps of direct sums."""
    rng = np.random.default_rng(11)
    specs = []
    for i in range(6):
        bg = rng.random((2, 3000))
        cl = np.array([[0.3], [0.6]]) + 1e-4 * rng.standard_normal((2, 1000))
        src = np.concatenate([bg, cl], axis=1)
        n = src.shape[1]
        kern = serve.CAUCHY if i % 2 == 0 else serve.HELMHOLTZ
        s = dict(kernel=kern, eps=1e-9, sources=src, pg=2,
                 charges=rng.standard_normal(n) + 1j * rng.standard_normal(n))
        if kern == serve.HELMHOLTZ:
            s["zk"] = 5.0 + i
        specs.append(s)
    res = _run_fused(specs)
    _check_vs_direct(specs, res, "deep clusters")
    _check_vs_alone(specs, res, "deep clusters")


def test_fused_device_arrays_equal_host_arrays():
    """device_ptrs = 1 (arrays in HBM in the drop-in layouts): the same
    result as from host arrays, bit for bit."""
    specs = _specs(7, 12, [serve.CAUCHY, serve.HELMHOLTZ], [1e-6])
    host = _run_fused(specs)
    lib = F.lib()
    hip = Hip()
    arr = (serve.fmm2dhip_problem * len(specs))()
    outs = []
    try:
        for i, s in enumerate(specs):
            n = s["sources"].shape[1]
            ng = 1 if s["kernel"] == serve.CAUCHY else 2
            src = hip.put(np.asfortranarray(s["sources"]))
            ch = hip.put(np.asarray(s["charges"], complex))
            pot = hip.zeros(16 * n)
            grad = hip.zeros(16 * n * ng)
            p = arr[i]
            p.kernel, p.nd, p.eps = s["kernel"], 1, s["eps"]
            p.zk_re, p.zk_im = complex(s.get("zk", 0.0)).real, 0.0
            p.ns, p.sources = n, src
            p.ifcharge, p.charge = 1, ch
            p.ifdipole, p.dipstr, p.dipvec = 0, ch, src
            p.ifpgh, p.pot, p.grad, p.hess = 2, pot, grad, grad
            p.nt, p.targ, p.ifpghtarg = 0, src, 0
            p.pottarg = p.gradtarg = p.hesstarg = pot
            p.ndiv, p.precision, p.device_ptrs = 0, 0, 1
            outs.append((pot, grad, n, ng))
        lib.fmm2dhip_batch_fused.restype = C.c_int
        assert lib.fmm2dhip_batch_fused(len(specs), arr, 4) == 0
        for i, (pot, grad, n, ng) in enumerate(outs):
            pv = hip.get(pot, np.zeros(n, complex))
            gv = hip.get(grad, np.zeros((ng, n), complex, order="F"))
            assert np.array_equal(pv, host[i]["pot"].ravel()), i
            assert np.array_equal(gv.ravel(order="F"),
                                  np.asarray(host[i]["grad"]).reshape(ng, n).ravel(order="F")), i
    finally:
        hip.free()


def test_fused_c4_workload():
    specs = workloads.c4_specs()
    res = _run_fused(specs)
    print("C4 fused: solved", flush=True)
    lib = F.lib()
    worst = {}
    for i, s in enumerate(specs):
        e = errors_vs_direct(lib, s, res[i], nsub=300, seed=i)
        for k, v in e.items():
            assert v < s["eps"], (i, k, v, s["kernel"])
            worst[s["kernel"]] = max(worst.get(s["kernel"], 0.0), v)
        if i % 128 == 127:
            print(f"C4 fused direct: {i + 1} problems checked", flush=True)
    print(f"C4 fused: worst rel l2 vs direct by kernel {worst}", flush=True)
    for i in range(0, len(specs), 8):
        s = specs[i]
        a = alone(lib, s)
        tol = LAP_TOL if s["kernel"] != serve.HELMHOLTZ else 2 * s["eps"]
        for k in ("pot", "grad"):
            assert rel_l2(res[i][k], a[k]) < tol, (i, k, s["kernel"])
    print("C4 fused: every 8th problem agrees with its solve alone", flush=True)
