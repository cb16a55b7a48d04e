"""Seeded synthetic workloads of the batched and served configurations.

Both ``bench.py`` and the GPU tests draw their problems from here, so the
problems a test checks are the ones the benchmark times.

* ``c4_specs``: 512 independent problems.  N is lognormal with median 2e4,
  clipped to [2e3, 2e5]; points uniform in [0,1]^2 or a Gaussian blob
  (sigma 0.05) with equal odds; half Helmholtz with zk ~ U[1, 50], half
  Cauchy; complex charges N(0,1); eps 1e-6; pot + grad at the sources.
* ``C5Stream``: Poisson arrivals at 200 requests/s.  60% log-kernel Laplace
  (lfmm2d), 40% Helmholtz with zk ~ U[1, 100]; N log-uniform in [1e3, 2e6]
  uniform in [0,1]^2; charges and dipoles N(0,1) with unit dipole
  directions; eps from {1e-4, 1e-6, 1e-10}; pot + grad at the sources.
  Inputs are slices of four pre-drawn pools of 2e6 points (drawing 2e6
  fresh points per request would dominate the host time of the stream).

Each spec is a dict of ``serve.Problem`` keyword arguments.
"""
from __future__ import annotations

import numpy as np

from . import serve


def c4_specs(nprob=512, seed=4):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nprob):
        n = int(np.clip(np.exp(rng.normal(np.log(2e4), 1.0)), 2e3, 2e5))
        src = rng.random((2, n)) if rng.random() < 0.5 else 0.5 + 0.05 * rng.standard_normal((2, n))
        helm = rng.random() < 0.5
        specs.append(dict(kernel=serve.HELMHOLTZ if helm else serve.CAUCHY, eps=1e-6,
                          zk=float(rng.uniform(1, 50)) if helm else 0.0, sources=src,
                          charges=rng.standard_normal(n) + 1j * rng.standard_normal(n), pg=2))
    return specs


class C5Stream:
    """Arrival times and request specs of the served configuration.

    ``arrivals`` (seconds from the start) are Poisson at ``rate`` over
    ``seconds``; ``spec(i)`` draws request i's problem (deterministic in
    the order the specs are drawn).  ``precision`` is given to every request,
    or, when it is "alternate", FP64 and mixed precision alternate."""

    NPOOL, M = 4, 2_000_000

    def __init__(self, seconds, rate=200.0, seed=5, nmax=2_000_000, precision=serve.FP64):
        rng = np.random.default_rng(seed)
        self.pools = []
        for _ in range(self.NPOOL):
            a = rng.random(self.M) * 2 * np.pi
            self.pools.append((np.asfortranarray(rng.random((2, self.M))),
                               rng.standard_normal(self.M) + 1j * rng.standard_normal(self.M),
                               rng.standard_normal(self.M) + 1j * rng.standard_normal(self.M),
                               np.asfortranarray(np.stack([np.cos(a), np.sin(a)])[None])))
        t = np.cumsum(rng.exponential(1.0 / rate, int(rate * seconds * 1.5) + 10))
        self.arrivals = t[t < seconds]
        self.rng = rng
        self.nmax = nmax
        self.precision = precision
        self.drawn = 0

    def __len__(self):
        return len(self.arrivals)

    def spec(self):
        rng = self.rng
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
        n = min(n, self.nmax)
        src, ch, dp, dv = self.pools[int(rng.integers(self.NPOOL))]
        off = int(rng.integers(0, self.M - n + 1))
        helm = rng.random() < 0.4
        eps = float(rng.choice([1e-4, 1e-6, 1e-10]))
        zk = float(rng.uniform(1, 100)) if helm else 0.0
        prec = self.precision
        if prec == "alternate":
            prec = serve.FP64 if self.drawn % 2 == 0 else serve.MIXED
        self.drawn += 1
        return dict(kernel=serve.HELMHOLTZ if helm else serve.LOG, eps=eps, zk=zk,
                    sources=src[:, off:off + n], charges=ch[off:off + n],
                    dipstr=dp[off:off + n], dipvec=dv[:, :, off:off + n], pg=2,
                    precision=prec)
