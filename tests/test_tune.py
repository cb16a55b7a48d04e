"""The autotuner's table (fmm2dhip_tune_*): defaults, text round trip and
validation of what a file may set (no GPU needed), and on the GPU one
autotuning run whose choices leave results unchanged to rounding."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import serve


def _defaults():
    # the built-in rule: all row tiles in one wave up to 4, else two groups
    return {mt: (mt if mt <= 4 else (mt + 1) // 2) for mt in range(1, 9)}


def test_table_defaults_and_round_trip(tmp_path):
    # a fresh process, so that no other test has tuned the table
    path = tmp_path / "tune.txt"
    code = f"""
import sys; sys.path.insert(0, {os.getcwd()!r})
import fmm2d_c_translation_amd as F
from fmm2d_c_translation_amd import serve
lib = F.lib()
t = serve.tune_table(lib)
print(repr(t))
open({str(path)!r}, "w").write("m2l_mw laplace fp64 3 1\\nm2l_mw helmholtz fp32 6 4\\n"
                            "xlate_waves 2048\\nm2l_mw laplace fp64 5 9\\n"
                            "m2l_mw poisson fp64 2 1\\nbogus line\\n")
print(lib.fmm2dhip_tune_load({str(path)!r}.encode()))
print(repr(serve.tune_table(lib)))
print(lib.fmm2dhip_tune_save({str(path) + '.out'!r}.encode()))
"""
    env = dict(os.environ)
    env.pop("FMM2DHIP_TUNE_FILE", None)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    t0 = eval(lines[0])
    for k in serve.TUNE_KEYS:
        assert t0[k] == _defaults(), k
    assert t0["xlate_waves"] == 4096
    assert int(lines[1]) == 3  # the two valid m2l lines and the threshold
    t1 = eval(lines[2])
    assert t1["m2l_laplace_fp64"][3] == 1 and t1["m2l_helmholtz_fp32"][6] == 4
    assert t1["m2l_helmholtz_fp64"][3] == _defaults()[3]  # families apart
    assert t1["xlate_waves"] == 2048
    assert t1["m2l_laplace_fp64"][5] == _defaults()[5]  # 9 tiles per wave is refused
    assert int(lines[3]) == 0
    saved = open(str(path) + ".out").read()
    assert "m2l_mw laplace fp64 3 1" in saved and "m2l_mw helmholtz fp32 6 4" in saved
    assert "xlate_waves 2048" in saved


@pytest.mark.gpu
def test_autotune_keeps_results():
    lib = F.lib()
    rng = np.random.default_rng(3)
    n = 200_000
    src = rng.random((2, n))
    ch = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    probs = {"helm": serve.Problem(serve.HELMHOLTZ, 1e-9, src, charges=ch, pg=2, zk=40.0),
             "lap": serve.Problem(serve.CAUCHY, 1e-9, src, charges=ch, pg=2)}
    before = {}
    for k, p in probs.items():
        assert serve.batch(lib, [p], 1) == 0
        before[k] = p.result()
    for k, p in probs.items():
        changed = serve.autotune(lib, p, nrep=2)
        assert changed >= 0, changed
        print(f"autotune {k}: {changed} choices changed -> {serve.tune_table(lib)}", flush=True)
    t = serve.tune_table(lib)
    for half in serve.TUNE_KEYS:
        for mt, v in t[half].items():
            assert 1 <= v <= min(4, mt)
    assert t["xlate_waves"] in (1024, 2048, 4096, 8192)
    for k, p in probs.items():
        assert serve.batch(lib, [p], 1) == 0
        r = p.result()
        for f in ("pot", "grad"):
            a, b = np.asarray(r[f]).ravel(), np.asarray(before[k][f]).ravel()
            assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b), (k, f)
