"""ctypes calls of the biharmonic and Stokes entry points (bhfmm2d_,
stfmm2d_, bh2d_direct*_) shared by the tests; the same signatures in the
library, the compiled reference and (direct sums) the oracle."""
import ctypes as C

import numpy as np

P = lambda a: a.ctypes.data_as(C.c_void_p)
i32 = lambda v: C.byref(C.c_int32(int(v)))
f64 = lambda v: C.byref(C.c_double(float(v)))


def inputs(nd, ns, nt, seed, cluster=False):
    rng = np.random.default_rng(seed)
    if cluster:
        th = rng.random(ns) * 2 * np.pi
        r = 0.35 + 1e-3 * rng.standard_normal(ns)
        src = np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)])
        src[:, : ns // 10] = rng.random((2, ns // 10))
    else:
        src = rng.random((2, ns))
    src = np.asfortranarray(src)
    trg = np.asfortranarray(rng.random((2, nt)))
    cz = lambda *s: np.asfortranarray(rng.standard_normal(s) + 1j * rng.standard_normal(s))
    return src, trg, cz(nd, 2, ns), cz(nd, 3, ns)


def bh_direct(lib, nd, src, chg, dip, trg, pg, th=1e-14, oracle=False):
    ns, nt = src.shape[1], trg.shape[1]
    vel = np.zeros((nd, nt), complex, order="F")
    grad = np.zeros((nd, 3, nt), complex, order="F")
    if oracle:
        lib.oracle_bh2d_direct(C.c_int(nd), P(src), C.c_int(ns), C.c_int(int(chg is not None)),
                               P(chg) if chg is not None else None,
                               C.c_int(int(dip is not None)), P(dip) if dip is not None else None,
                               P(trg), C.c_int(nt), C.c_int(pg), P(vel), P(grad), C.c_double(th))
        return vel, grad
    var = ("c" if chg is not None else "") + ("d" if dip is not None else "")
    args = [i32(nd), P(src), i32(ns)]
    if chg is not None:
        args.append(P(chg))
    if dip is not None:
        args.append(P(dip))
    args += [P(trg), i32(nt), P(vel)] + ([P(grad)] if pg >= 2 else []) + [f64(th)]
    getattr(lib, f"bh2d_direct{var}{'pg'[pg - 1]}_")(*args)
    return vel, grad


def bhfmm(lib, eps, nd, src, chg, dip, pg, trg=None, pgt=0):
    ns = src.shape[1]
    nt = trg.shape[1] if trg is not None else 0
    tg = trg if trg is not None else np.zeros((2, 1), order="F")
    z = lambda *s: np.zeros(s, complex, order="F")
    out = dict(pot=z(nd, ns), grad=z(nd, 3, ns), hess=z(nd, 3, ns), pottarg=z(nd, max(nt, 1)),
               gradtarg=z(nd, 3, max(nt, 1)), hesstarg=z(nd, 3, max(nt, 1)))
    iper, ier = C.c_int32(7), C.c_int32(-1)
    cc = chg if chg is not None else z(nd, 2, 1)
    dd = dip if dip is not None else z(nd, 3, 1)
    lib.bhfmm2d_(i32(nd), f64(eps), i32(ns), P(src), i32(chg is not None), P(cc),
                 i32(dip is not None), P(dd), C.byref(iper), i32(pg), P(out["pot"]),
                 P(out["grad"]), P(out["hess"]), i32(nt), P(tg), i32(pgt), P(out["pottarg"]),
                 P(out["gradtarg"]), P(out["hesstarg"]), C.byref(ier))
    out["ier"], out["iper"] = ier.value, iper.value
    return out


def stfmm(lib, eps, nd, src, stok, strs, strv, pp, trg=None, ppt=0):
    ns = src.shape[1]
    nt = trg.shape[1] if trg is not None else 0
    tg = trg if trg is not None else np.zeros((2, 1), order="F")
    z = lambda *s: np.zeros(s, order="F")
    out = dict(pot=z(nd, 2, ns), pre=z(nd, ns), grad=z(nd, 2, 2, ns), pottarg=z(nd, 2, max(nt, 1)),
               pretarg=z(nd, max(nt, 1)), gradtarg=z(nd, 2, 2, max(nt, 1)))
    ier = C.c_int32(-1)
    f = stok if stok is not None else z(nd, 2, ns)
    m = strs if strs is not None else z(nd, 2, ns)
    n = strv if strv is not None else z(nd, 2, ns)
    lib.stfmm2d_(i32(nd), f64(eps), i32(ns), P(src), i32(stok is not None), P(f),
                 i32(strs is not None), P(m), P(n), i32(pp), P(out["pot"]), P(out["pre"]),
                 P(out["grad"]), i32(nt), P(tg), i32(ppt), P(out["pottarg"]), P(out["pretarg"]),
                 P(out["gradtarg"]), C.byref(ier))
    out["ier"] = ier.value
    return out
