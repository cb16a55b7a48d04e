"""CPU checkers for the 2D FMM (test infrastructure only).

THIS PACKAGE IS THE ORACLE, NOT THE PRODUCT.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
it.  Two libraries are reachable from here:

* ``liboracle.so`` -- our own C99 restatement of the reference algorithm
  (``oracle/*.c``; every routine cites the reference file:line it follows).
  Its entry points carry the reference ABI under an ``orc_`` prefix.
* ``_ref/libfmm2d_ref.so`` -- the reference C translation compiled from its
  own sources by ``oracle/Makefile`` (only where the reference checkout is
  present when ``build()`` runs; the built file then travels with the tree).

Parity status of each oracle function is recorded in the header of the C
file that holds it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libfmm2d_ref.so")

_VP = C.c_void_p
_cache = {}


def build(quiet=True):
    """Compile liboracle.so (and the reference library when its sources exist)."""
    import subprocess
    cmd = ["make", "-C", HERE, "-j8", "all"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout[-4000:] + out.stderr[-4000:])
    if not quiet:
        print(out.stdout)


def oracle_lib():
    if "orc" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise FileNotFoundError(f"{ORACLE_SO} missing: run oracle.build()")
        _cache["orc"] = C.CDLL(ORACLE_SO)
    return _cache["orc"]


def ref_lib():
    """The compiled reference, or None where it was never built."""
    if "ref" not in _cache:
        _cache["ref"] = C.CDLL(REF_SO) if os.path.exists(REF_SO) else None
    return _cache["ref"]


def _p(a):
    return a.ctypes.data_as(_VP)


def _i(v):
    return C.byref(C.c_int32(int(v)))


class Tree:
    """A tree in the reference itree layout (pts_tree2d.c:484-491), 1-based ids."""

    def __init__(self, nlevels, nboxes, itree, centers, boxsize, isrc, isrcse, itarg, itargse):
        self.nlevels, self.nboxes = nlevels, nboxes
        self.itree, self.centers, self.boxsize = itree, centers, boxsize
        self.isrc, self.isrcse, self.itarg, self.itargse = isrc, isrcse, itarg, itargse
        n, L = nboxes, nlevels
        o = 2 * (L + 1)
        self.laddr = itree[:o].reshape(L + 1, 2)
        self.ilevel = itree[o:o + n]
        self.iparent = itree[o + n:o + 2 * n]
        self.nchild = itree[o + 2 * n:o + 3 * n]
        self.ichild = itree[o + 3 * n:o + 7 * n].reshape(n, 4)
        self.nnbors = itree[o + 7 * n:o + 8 * n]
        self.nbors = itree[o + 8 * n:o + 17 * n].reshape(n, 9)


def oracle_tree(src, targ, ndiv, idivflag=0):
    """Tree + point sort from our restatement."""
    lib = oracle_lib()
    src = np.asfortranarray(np.asarray(src, np.float64).reshape(2, -1, order="F"))
    ns = src.shape[1]
    tg = np.zeros((2, 1), order="F") if targ is None else np.asfortranarray(
        np.asarray(targ, np.float64).reshape(2, -1, order="F"))
    nt = 0 if targ is None else tg.shape[1]
    nl, nb = C.c_int32(0), C.c_int32(0)
    lib.oracle_tree_new.restype = _VP
    h = lib.oracle_tree_new(_p(src), C.c_int32(ns), _p(tg), C.c_int32(nt), C.c_int32(idivflag),
                            C.c_int32(ndiv), C.byref(nl), C.byref(nb))
    L, n = nl.value, nb.value
    itree = np.zeros(17 * n + 2 * (L + 1), np.int32)
    centers = np.zeros((n, 2))
    boxsize = np.zeros(L + 1)
    isrc = np.zeros(max(ns, 1), np.int32)
    itarg = np.zeros(max(nt, 1), np.int32)
    isrcse = np.zeros((n, 2), np.int32)
    itargse = np.zeros((n, 2), np.int32)
    lib.oracle_tree_export(_VP(h), _p(itree), _p(centers), _p(boxsize), _p(isrc), _p(isrcse),
                           _p(itarg), _p(itargse))
    nlist = np.zeros((n, 4), np.int32)
    lists = np.zeros((n, 65), np.int32)
    lib.oracle_lists_export(_VP(h), _p(nlist), _p(lists))
    lib.oracle_tree_free(_VP(h))
    t = Tree(L, n, itree, centers, boxsize, isrc[:ns], isrcse, itarg[:nt], itargse)
    t.nlist, t.lists = nlist, lists
    return t


def ref_tree(src, targ, ndiv, idivflag=0):
    """Tree + point sort + lists from the compiled reference
    (pts_tree_mem_/pts_tree_build_/pts_tree_sort_/computelists_)."""
    lib = ref_lib()
    if lib is None:
        return None
    src = np.asfortranarray(np.asarray(src, np.float64).reshape(2, -1, order="F"))
    ns = src.shape[1]
    tg = np.zeros((2, 1), order="F") if targ is None else np.asfortranarray(
        np.asarray(targ, np.float64).reshape(2, -1, order="F"))
    nt = 0 if targ is None else tg.shape[1]
    nl, nb, lt = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    iper = C.c_int32(0)
    lib.pts_tree_mem_(_p(src), _i(ns), _p(tg), _i(nt), _i(idivflag), _i(ndiv), _i(0), _i(51),
                      _i(0), C.byref(iper), C.byref(nl), C.byref(nb), C.byref(lt))
    L, n = nl.value, nb.value
    itree = np.zeros(lt.value, np.int32)
    iptr = np.zeros(8, np.int32)
    centers = np.zeros((n, 2))
    boxsize = np.zeros(L + 1)
    lib.pts_tree_build_(_p(src), _i(ns), _p(tg), _i(nt), _i(idivflag), _i(ndiv), _i(0), _i(51),
                        _i(0), C.byref(iper), C.byref(nl), C.byref(nb), C.byref(lt), _p(itree),
                        _p(iptr), _p(centers), _p(boxsize))
    isrc = np.zeros(max(ns, 1), np.int32)
    isrcse = np.zeros((n, 2), np.int32)
    itarg = np.zeros(max(nt, 1), np.int32)
    itargse = np.zeros((n, 2), np.int32)
    lib.pts_tree_sort_(_i(ns), _p(src), _p(itree), C.byref(lt), C.byref(nb), C.byref(nl), _p(iptr),
                       _p(centers), _p(isrc), _p(isrcse))
    lib.pts_tree_sort_(_i(nt), _p(tg), _p(itree), C.byref(lt), C.byref(nb), C.byref(nl), _p(iptr),
                       _p(centers), _p(itarg), _p(itargse))
    mn = [C.c_int32(13), C.c_int32(27), C.c_int32(20), C.c_int32(5)]
    ns_ = [np.zeros(n, np.int32) for _ in range(4)]
    ls_ = [np.zeros((n, m.value), np.int32) for m in mn]
    lib.computelists_(C.byref(nl), C.byref(nb), _p(itree), C.byref(lt), _p(iptr), _p(centers),
                      _p(boxsize), C.byref(iper),
                      C.byref(mn[0]), _p(ns_[0]), _p(ls_[0]), C.byref(mn[1]), _p(ns_[1]), _p(ls_[1]),
                      C.byref(mn[2]), _p(ns_[2]), _p(ls_[2]), C.byref(mn[3]), _p(ns_[3]), _p(ls_[3]))
    t = Tree(L, n, itree[:17 * n + 2 * (L + 1)], centers, boxsize, isrc[:ns], isrcse, itarg[:nt], itargse)
    t.nlist = np.stack(ns_, axis=1)
    lists = np.zeros((n, 65), np.int32)
    off = 0
    for k, m in enumerate([13, 27, 20, 5]):
        for b in range(n):
            c = t.nlist[b, k]
            lists[b, off:off + c] = ls_[k][b, :c]
        off += m
    t.lists = lists
    return t
