"""Argument marshalling for the gfortran-ABI entry points.

Every entry point of this library (and of the reference C translation it
replaces) takes all arguments by pointer, column-major arrays with ``nd``
fastest, ``int32`` flags and ``complex128`` strengths
(c_translation/include/fmm2d_c.h:35-61).  The helpers here build those
arguments from numpy arrays and call one exported symbol of a
``ctypes.CDLL``.  They are the Python mirror of the C interface: the same
names, the same argument meaning, outputs overwritten (not accumulated),
``iper`` written 0 and ``ier`` returned.

Array conventions on the Python side follow the reference's own Python
wrapper (python/fmm2dpy/fmm2d.py): points are ``(2, n)``, strengths
``(nd, n)`` (or ``(n,)`` when ``nd == 1``), dipole vectors ``(nd, 2, n)``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_VP = C.c_void_p


def _i(v: int):
    return C.byref(C.c_int32(int(v)))


def _d(v: float):
    return C.byref(C.c_double(float(v)))


def _farr(a, dtype, shape):
    """Return a Fortran-ordered copy/view of ``a`` with the given shape."""
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a.reshape(shape, order="F"))


def _p(a):
    return a.ctypes.data_as(_VP)


class FmmResult(dict):
    """Outputs of one FMM call (``pot``, ``grad``, ... as requested, plus ``ier``)."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e


def _squeeze(a, nd):
    return a[0] if nd == 1 else a


def _strengths(nd, n, arr, flag):
    if flag and arr is not None:
        return _farr(arr, np.complex128, (nd, n))
    return np.zeros((nd, 1), dtype=np.complex128, order="F")


def _outs(nd, n, flag, ncomp_grad, ncomp_hess):
    """Allocate pot/grad/hess; unrequested ones get a (nd,1) placeholder."""
    n1 = max(n, 1)
    pot = np.zeros((nd, n1 if flag >= 1 else 1), np.complex128, order="F")
    gshape = (nd, n1) if ncomp_grad == 1 else (nd, ncomp_grad, n1)
    hshape = (nd, n1) if ncomp_hess == 1 else (nd, ncomp_hess, n1)
    grad = np.zeros(gshape if flag >= 2 else (nd,) + (1,) * (len(gshape) - 1), np.complex128, order="F")
    hess = np.zeros(hshape if flag >= 3 else (nd,) + (1,) * (len(hshape) - 1), np.complex128, order="F")
    return pot, grad, hess


def _collect(res, prefix, nd, n, flag, pot, grad, hess):
    if flag >= 1:
        res[prefix + "pot"] = _squeeze(pot[:, :n], nd)
    if flag >= 2:
        res[prefix + "grad"] = _squeeze(grad[..., :n], nd)
    if flag >= 3:
        res[prefix + "hess"] = _squeeze(hess[..., :n], nd)


def cfmm2d_call(lib, symbol, eps, sources, charges=None, dipstr=None, pg=1, targets=None,
                pgt=0, nd=1, ndiv=None, idivflag=0, ifnear=1, timeinfo=None):
    """Call a ``cfmm2d_``-family symbol (Cauchy kernel; grad = d/dz, hess = d2/dz2).

    With ``ndiv`` given the ``*_ndiv_`` signature is used (cfmm2d_ndiv.c:94),
    otherwise the plain one (cfmm2d.c:1053).
    """
    src = _farr(sources, np.float64, (2, -1))
    ns = src.shape[1]
    tg = _farr(targets, np.float64, (2, -1)) if targets is not None else np.zeros((2, 1), order="F")
    nt = tg.shape[1] if targets is not None else 0
    ifc = int(charges is not None)
    ifd = int(dipstr is not None)
    ch = _strengths(nd, ns, charges, ifc)
    dp = _strengths(nd, ns, dipstr, ifd)
    pot, grad, hess = _outs(nd, ns, pg, 1, 1)
    pt, gt, ht = _outs(nd, nt, pgt, 1, 1)
    iper = C.c_int32(7)
    ier = C.c_int32(-1)
    args = [_i(nd), _d(eps), _i(ns), _p(src), _i(ifc), _p(ch), _i(ifd), _p(dp), C.byref(iper),
            _i(pg), _p(pot), _p(grad), _p(hess), _i(nt), _p(tg), _i(pgt), _p(pt), _p(gt), _p(ht)]
    ti = np.zeros(8) if timeinfo is None else timeinfo
    if ndiv is not None:
        args += [_i(ndiv), _i(idivflag), _i(ifnear), _p(ti)]
    args += [C.byref(ier)]
    getattr(lib, symbol)(*args)
    res = FmmResult(ier=ier.value, iper=iper.value, timeinfo=ti)
    _collect(res, "", nd, ns, pg, pot, grad, hess)
    _collect(res, "targ_", nd, nt, pgt, pt, gt, ht)
    return res


def lfmm2d_call(lib, symbol, eps, sources, charges=None, dipstr=None, dipvec=None, pg=1,
                targets=None, pgt=0, nd=1, ndiv=None, idivflag=0, ifnear=1, timeinfo=None):
    """Call an ``lfmm2d_``-family symbol (log kernel, complex densities;
    grad (nd,2,n), hess (nd,3,n)); signature of lfmm2d_ndiv.c:57 with
    ``ndiv``, of src/laplace/lfmm2d.f:29 without."""
    src = _farr(sources, np.float64, (2, -1))
    ns = src.shape[1]
    tg = _farr(targets, np.float64, (2, -1)) if targets is not None else np.zeros((2, 1), order="F")
    nt = tg.shape[1] if targets is not None else 0
    ifc = int(charges is not None)
    ifd = int(dipstr is not None)
    ch = _strengths(nd, ns, charges, ifc)
    dp = _strengths(nd, ns, dipstr, ifd)
    dv = _farr(dipvec, np.float64, (nd, 2, ns)) if ifd else np.zeros((nd, 2, 1), order="F")
    pot, grad, hess = _outs(nd, ns, pg, 2, 3)
    pt, gt, ht = _outs(nd, nt, pgt, 2, 3)
    iper = C.c_int32(7)
    ier = C.c_int32(-1)
    args = [_i(nd), _d(eps), _i(ns), _p(src), _i(ifc), _p(ch), _i(ifd), _p(dp), _p(dv),
            C.byref(iper), _i(pg), _p(pot), _p(grad), _p(hess), _i(nt), _p(tg), _i(pgt),
            _p(pt), _p(gt), _p(ht)]
    ti = np.zeros(8) if timeinfo is None else timeinfo
    if ndiv is not None:
        args += [_i(ndiv), _i(idivflag), _i(ifnear), _p(ti)]
    args += [C.byref(ier)]
    getattr(lib, symbol)(*args)
    res = FmmResult(ier=ier.value, iper=iper.value, timeinfo=ti)
    _collect(res, "", nd, ns, pg, pot, grad, hess)
    _collect(res, "targ_", nd, nt, pgt, pt, gt, ht)
    return res


def rfmm2d_call(lib, symbol, eps, sources, charges=None, dipstr=None, dipvec=None, pg=1,
                targets=None, pgt=0, nd=1, ndiv=None, idivflag=0, ifnear=1, timeinfo=None):
    """Call an ``rfmm2d_``-family symbol (log kernel, real densities and
    outputs; grad (nd,2,n), hess (nd,3,n), all float64); signature of
    rfmm2d_ndiv.c:51 with ``ndiv``, of src/laplace/rfmm2d.f:27 without."""
    src = _farr(sources, np.float64, (2, -1))
    ns = src.shape[1]
    tg = _farr(targets, np.float64, (2, -1)) if targets is not None else np.zeros((2, 1), order="F")
    nt = tg.shape[1] if targets is not None else 0
    ifc = int(charges is not None)
    ifd = int(dipstr is not None)
    ch = _farr(charges, np.float64, (nd, ns)) if ifc else np.zeros((nd, 1), order="F")
    dp = _farr(dipstr, np.float64, (nd, ns)) if ifd else np.zeros((nd, 1), order="F")
    dv = _farr(dipvec, np.float64, (nd, 2, ns)) if ifd else np.zeros((nd, 2, 1), order="F")
    pot, grad, hess = (a.real.copy(order="F") for a in _outs(nd, ns, pg, 2, 3))
    pt, gt, ht = (a.real.copy(order="F") for a in _outs(nd, nt, pgt, 2, 3))
    iper = C.c_int32(7)
    ier = C.c_int32(-1)
    args = [_i(nd), _d(eps), _i(ns), _p(src), _i(ifc), _p(ch), _i(ifd), _p(dp), _p(dv),
            C.byref(iper), _i(pg), _p(pot), _p(grad), _p(hess), _i(nt), _p(tg), _i(pgt),
            _p(pt), _p(gt), _p(ht)]
    ti = np.zeros(8) if timeinfo is None else timeinfo
    if ndiv is not None:
        args += [_i(ndiv), _i(idivflag), _i(ifnear), _p(ti)]
    args += [C.byref(ier)]
    getattr(lib, symbol)(*args)
    res = FmmResult(ier=ier.value, iper=iper.value, timeinfo=ti)
    _collect(res, "", nd, ns, pg, pot, grad, hess)
    _collect(res, "targ_", nd, nt, pgt, pt, gt, ht)
    return res


def hfmm2d_call(lib, symbol, eps, zk, sources, charges=None, dipstr=None, dipvec=None, pg=1,
                targets=None, pgt=0, nd=1, ndiv=None, idivflag=0, ifnear=1, timeinfo=None):
    """Call an ``hfmm2d_``-family symbol (Helmholtz, H0(k r) i/4 kernel;
    grad (nd,2,n), hess (nd,3,n)); signature of hfmm2d.c:1529 or, with
    ``ndiv``, hfmm2d_ndiv.c:68."""
    src = _farr(sources, np.float64, (2, -1))
    ns = src.shape[1]
    tg = _farr(targets, np.float64, (2, -1)) if targets is not None else np.zeros((2, 1), order="F")
    nt = tg.shape[1] if targets is not None else 0
    ifc = int(charges is not None)
    ifd = int(dipstr is not None)
    ch = _strengths(nd, ns, charges, ifc)
    dp = _strengths(nd, ns, dipstr, ifd)
    dv = _farr(dipvec, np.float64, (nd, 2, ns)) if ifd else np.zeros((nd, 2, 1), order="F")
    pot, grad, hess = _outs(nd, ns, pg, 2, 3)
    pt, gt, ht = _outs(nd, nt, pgt, 2, 3)
    zkc = (C.c_double * 2)(complex(zk).real, complex(zk).imag)
    iper = C.c_int32(7)
    ier = C.c_int32(-1)
    args = [_i(nd), _d(eps), zkc, _i(ns), _p(src), _i(ifc), _p(ch), _i(ifd), _p(dp), _p(dv),
            C.byref(iper), _i(pg), _p(pot), _p(grad), _p(hess), _i(nt), _p(tg), _i(pgt),
            _p(pt), _p(gt), _p(ht)]
    ti = np.zeros(8) if timeinfo is None else timeinfo
    if ndiv is not None:
        args += [_i(ndiv), _i(idivflag), _i(ifnear), _p(ti)]
    args += [C.byref(ier)]
    getattr(lib, symbol)(*args)
    res = FmmResult(ier=ier.value, iper=iper.value, timeinfo=ti)
    _collect(res, "", nd, ns, pg, pot, grad, hess)
    _collect(res, "targ_", nd, nt, pgt, pt, gt, ht)
    return res


def ndiv_rule(lib, symbol, eps, ns, nt):
    """``hndiv2d_``/``lndiv2d_``: leaf size and idivflag for a tolerance."""
    nd = C.c_int32(0)
    idiv = C.c_int32(0)
    getattr(lib, symbol)(_d(eps), _i(ns), _i(nt), _i(1), _i(0), _i(1), _i(0), C.byref(nd), C.byref(idiv))
    return nd.value, idiv.value
