"""MI355X-native 2D FMM engine with the reference C translation's entry points.

``lib()`` loads ``lib/libfmm2d_hip.so`` (hand-written HIP for gfx950 behind a
C ABI, ``include/fmm2d_hip.h``); there is no other compute path: when the
library is missing, every call raises.

The functions below mirror the reference's C entry points one to one
(``cfmm2d_``, ``cfmm2d_ndiv_``, ``lfmm2d_``, ``lfmm2d_ndiv_``): the same
argument meaning, outputs overwritten, ``iper`` written 0 and ``ier``
returned in the result.  Array conventions follow the reference's Python
wrapper (python/fmm2dpy/fmm2d.py): points ``(2, n)``, strengths ``(nd, n)``
(``(n,)`` for ``nd == 1``), dipole vectors ``(nd, 2, n)``.
"""
from __future__ import annotations

import ctypes as _C
import os as _os

from . import abi

HERE = _os.path.dirname(_os.path.abspath(__file__))
LIB_PATH = _os.path.join(HERE, "lib", "libfmm2d_hip.so")
_lib = None


class ExtensionMissing(RuntimeError):
    pass


def build(force=True):
    """Compile the HIP library in-tree (``make`` in this package)."""
    import subprocess
    cmd = ["make", "-C", HERE, "-j8"] + (["-B"] if force else [])
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("HIP build failed:\n" + out.stdout[-6000:] + out.stderr[-6000:])


def lib():
    """The loaded engine library; raises ExtensionMissing when it is not built."""
    global _lib
    if _lib is None:
        if not _os.path.exists(LIB_PATH):
            raise ExtensionMissing(f"{LIB_PATH} is not built; run fmm2d_c_translation_amd.build()")
        _lib = _C.CDLL(LIB_PATH)
    return _lib


FP64, MIXED = 0, 1


class precision:
    """Context manager: arithmetic of the near-field and M2L steps for the
    solves this thread makes inside the block.  ``FP64`` (the default) or
    ``MIXED`` (FP32 near field and M2L when eps >= 1e-6, results within eps
    of FP64 direct sums; tighter eps stays FP64)."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = lib().fmm2dhip_set_precision(self.mode)
        return self

    def __exit__(self, *exc):
        lib().fmm2dhip_set_precision(self.prev)
        return False


def device_name():
    buf = _C.create_string_buffer(64)
    rc = lib().fmm2dhip_device_info(buf, 64)
    return rc, buf.value.decode()


def cfmm2d(eps, sources, charges=None, dipstr=None, pg=1, targets=None, pgt=0, nd=1,
           ndiv=None, idivflag=0, ifnear=1):
    """Cauchy-kernel FMM (``cfmm2d_`` or, with ``ndiv``, ``cfmm2d_ndiv_``)."""
    sym = "cfmm2d_ndiv_" if ndiv is not None else "cfmm2d_"
    return abi.cfmm2d_call(lib(), sym, eps, sources, charges, dipstr, pg, targets, pgt, nd, ndiv,
                           idivflag, ifnear)


def lfmm2d(eps, sources, charges=None, dipstr=None, dipvec=None, pg=1, targets=None, pgt=0,
           nd=1, ndiv=None, idivflag=0, ifnear=1):
    """Log-kernel Laplace FMM, complex densities (``lfmm2d_`` / ``lfmm2d_ndiv_``)."""
    sym = "lfmm2d_ndiv_" if ndiv is not None else "lfmm2d_"
    return abi.lfmm2d_call(lib(), sym, eps, sources, charges, dipstr, dipvec, pg, targets, pgt, nd,
                           ndiv, idivflag, ifnear)


def rfmm2d(eps, sources, charges=None, dipstr=None, dipvec=None, pg=1, targets=None, pgt=0,
           nd=1, ndiv=None, idivflag=0, ifnear=1):
    """Log-kernel Laplace FMM, real densities (``rfmm2d_`` / ``rfmm2d_ndiv_``)."""
    sym = "rfmm2d_ndiv_" if ndiv is not None else "rfmm2d_"
    return abi.rfmm2d_call(lib(), sym, eps, sources, charges, dipstr, dipvec, pg, targets, pgt, nd,
                           ndiv, idivflag, ifnear)


def hfmm2d(eps, zk, sources, charges=None, dipstr=None, dipvec=None, pg=1, targets=None, pgt=0,
           nd=1, ndiv=None, idivflag=0, ifnear=1):
    """Helmholtz FMM (``hfmm2d_`` / ``hfmm2d_ndiv_``)."""
    sym = "hfmm2d_ndiv_" if ndiv is not None else "hfmm2d_"
    return abi.hfmm2d_call(lib(), sym, eps, zk, sources, charges, dipstr, dipvec, pg, targets, pgt,
                           nd, ndiv, idivflag, ifnear)
