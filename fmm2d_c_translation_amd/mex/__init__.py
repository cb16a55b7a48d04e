"""The reference's MATLAB/numbl front end, driven from Python against
libfmm2d_hip.so.

The gateway ``lib/fmm2d.so`` is the reference's own mwrap MEX gateway
(matlab/fmm2d.c) and numbl mex shim (matlab/numbl/mex_shim.cpp) compiled by
this directory's Makefile and linked against the HIP engine in place of
libfmm2d.a (matlab/numbl/build_native.sh:86-134).  ``Gateway.call`` is what
fmm2d.numbl.js's callFmm2d does (matlab/numbl/fmm2d.numbl.js:361-377): build
one mxArray per argument, call ``mex_dispatch`` (mex_shim.cpp:396-410) with
the mwrap identifier first, read the outputs back.  The functions below
restate the MATLAB wrappers that a numbl or MATLAB user calls, argument for
argument: cfmm2d/lfmm2d/rfmm2d/hfmm2d (matlab/{c,l,r,h}fmm2d.m: hndiv2d for
ndiv/idivflag, then the *_ndiv routine), the direct sums c2ddir/l2ddir/
r2ddir/h2ddir/st2ddir, and stfmm2d (matlab/stfmm2d.m).

Arrays follow MATLAB's layout: column-major, sources (2, ns), charges
(nd, ns), dipvec (nd, 2, ns) or (2, ns), and the results are squeezed the
way MATLAB's squeeze is (singleton dimensions dropped from arrays of three or
more dimensions, two dimensions kept).  A wrapper returns a ``Result`` whose
fields are the MATLAB struct U's (pot, grad, hess, pottarg, ...), with ier
and timeinfo (the wrappers' varargout) as attributes.

A failed call raises ``MexError`` with the gateway's message, as
mexErrMsgTxt does in MATLAB.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GATEWAY = os.path.join(HERE, "lib", "fmm2d.so")
MATLABDIR = "/root/reference/matlab"


class MexError(RuntimeError):
    pass


def build(matlabdir=MATLABDIR):
    """Compiles the gateway (needs the reference's matlab/ sources and the
    engine library already built)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", HERE, f"MATLABDIR={matlabdir}"], check=True)


def available():
    return os.path.exists(GATEWAY)


def mwrap_id(name, spec):
    """The mwrap dispatch identifier of a stub (the string mexFunction
    compares, matlab/fmm2d.c:11975-12000): ``spec`` lists the arguments as
    int/double/dcomplex for one-index arrays, int2/double2/dcomplex2 for
    two-index arrays, '+' in front for inout."""
    parts = []
    for tok in spec.split():
        io = "io" if tok.startswith("+") else "i"
        tok = tok.lstrip("+")
        rank = "[xx]" if tok.endswith("2") else "[x]"
        parts.append(f"{io} {tok.rstrip('2')}{rank}")
    return f"{name}({', '.join(parts)})"


class Gateway:
    """ctypes view of the numbl entry points of lib/fmm2d.so."""

    def __init__(self, path=GATEWAY):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is not built: run make -C {HERE} (needs the reference's matlab/ sources)")
        L = C.CDLL(path)
        vp = C.c_void_p
        for name, res, args in [
            ("mex_make_real_matrix", vp, [C.c_int, C.c_int, vp]),
            ("mex_make_complex_matrix", vp, [C.c_int, C.c_int, vp, vp]),
            ("mex_make_string", vp, [C.c_char_p]),
            ("mex_alloc_args", vp, [C.c_int]),
            ("mex_free_args", None, [vp]),
            ("mex_set_arg", None, [vp, C.c_int, vp]),
            ("mex_get_arg", vp, [vp, C.c_int]),
            ("mex_dispatch", C.c_int, [C.c_int, vp, C.c_int, vp]),
            ("mex_get_error", C.c_char_p, []),
            ("mex_get_m", C.c_int, [vp]),
            ("mex_get_n", C.c_int, [vp]),
            ("mex_get_is_complex", C.c_int, [vp]),
            ("mex_read_real", None, [vp, vp]),
            ("mex_read_complex", None, [vp, vp, vp]),
            ("mex_free_array", None, [vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    def _matrix(self, x, cplx):
        a = np.asarray(x)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        elif a.ndim > 2:
            a = a.reshape(a.shape[0], -1, order="F")
        m, n = a.shape
        if cplx:
            a = np.asarray(a, dtype=complex)
            re = np.asfortranarray(a.real, dtype=float)
            im = np.asfortranarray(a.imag, dtype=float)
            return self.L.mex_make_complex_matrix(m, n, re.ctypes.data, im.ctypes.data), (m, n)
        r = np.asfortranarray(np.real(a), dtype=float)
        return self.L.mex_make_real_matrix(m, n, r.ctypes.data), (m, n)

    def _read(self, p):
        m, n = self.L.mex_get_m(p), self.L.mex_get_n(p)
        if self.L.mex_get_is_complex(p):
            re = np.zeros((m, n), order="F")
            im = np.zeros((m, n), order="F")
            self.L.mex_read_complex(p, re.ctypes.data, im.ctypes.data)
            return re + 1j * im
        out = np.zeros((m, n), order="F")
        self.L.mex_read_real(p, out.ctypes.data)
        return out

    def call(self, name, spec, *args, dims=None):
        """fmm2d(mex_id_, args..., dims...) of the MATLAB wrappers: the
        dimension arguments are taken from the arrays themselves (numel for
        a one-index argument, rows and columns for a two-index one) unless
        ``dims`` gives them.  Returns the inout arguments, in order, as
        arrays."""
        toks = spec.split()
        if len(toks) != len(args):
            raise ValueError(f"{name}: {len(toks)} arguments expected, {len(args)} given")
        made, got = [], []
        nout = sum(t.startswith("+") for t in toks)
        try:
            for t, a in zip(toks, args):
                p, (m, n) = self._matrix(a, "dcomplex" in t)
                made.append(p)
                got += [m, n] if t.endswith("2") else [m * n]
            if dims is not None:
                got = dims
            made.insert(0, self.L.mex_make_string(mwrap_id(name, spec).encode()))
            made += [self.L.mex_make_real_matrix(1, 1, np.array([float(d)]).ctypes.data)
                     for d in got]
            prhs = self.L.mex_alloc_args(len(made))
            plhs = self.L.mex_alloc_args(max(nout, 1))
            for i, p in enumerate(made):
                self.L.mex_set_arg(prhs, i, p)
            rc = self.L.mex_dispatch(nout, plhs, len(made), prhs)
            outs = [self.L.mex_get_arg(plhs, i) for i in range(nout)]
            try:
                if rc != 0:
                    raise MexError(self.L.mex_get_error().decode())
                return [self._read(p) for p in outs]
            finally:
                for p in outs:
                    if p:
                        self.L.mex_free_array(p)
                self.L.mex_free_args(plhs)
                self.L.mex_free_args(prhs)
        finally:
            for p in made:
                self.L.mex_free_array(p)


_gw = None


def gateway():
    global _gw
    if _gw is None:
        _gw = Gateway()
    return _gw


class Result(dict):
    """MATLAB's U struct: fields by key or attribute; ier and timeinfo are
    the wrappers' extra outputs."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError:
            raise AttributeError(k) from None


def squeeze(a):
    """MATLAB squeeze: drops singleton dimensions of arrays with three or
    more dimensions, never below two (a lone vector becomes a column)."""
    a = np.asarray(a)
    if a.ndim <= 2:
        return a
    shape = [s for s in a.shape if s != 1]
    while len(shape) < 2:
        shape.append(1)
    return a.reshape(shape, order="F")


def _nd(srcinfo):
    return int(srcinfo.get("nd", 1))


def _field(srcinfo, k, shape, dtype):
    """srcinfo field k as a column-major array of shape, or None if absent."""
    if k not in srcinfo or srcinfo[k] is None:
        return None
    a = np.asarray(srcinfo[k], dtype=dtype)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"{k} must hold {shape} values, got shape {a.shape}")
    return np.asfortranarray(a).reshape(shape, order="F")


# -- FMMs: matlab/{c,l,r,h}fmm2d.m -------------------------------------------

SPECS = {
    "hndiv2d": "double int int int int int int +int +int",
    "cfmm2d_ndiv": "int double int double2 int dcomplex2 int dcomplex2 int int +dcomplex2 "
                   "+dcomplex2 +dcomplex2 int double2 int +dcomplex2 +dcomplex2 +dcomplex2 "
                   "int int int +double +int",
    "lfmm2d_ndiv": "int double int double2 int dcomplex2 int dcomplex2 double2 int int "
                   "+dcomplex2 +dcomplex2 +dcomplex2 int double2 int +dcomplex2 +dcomplex2 "
                   "+dcomplex2 int int int +double +int",
    "rfmm2d_ndiv": "int double int double2 int double2 int double2 double2 int int +double2 "
                   "+double2 +double2 int double2 int +double2 +double2 +double2 int int int "
                   "+double +int",
    "hfmm2d_ndiv": "int double dcomplex int double2 int dcomplex2 int dcomplex2 double2 int int "
                   "+dcomplex2 +dcomplex2 +dcomplex2 int double2 int +dcomplex2 +dcomplex2 "
                   "+dcomplex2 int int int +double +int",
    "st2ddirectstokg": "int double2 double2 int double2 int +double2 +double2 +double2 double",
    "st2ddirectstokstrsg": "int double2 int double2 int double2 double2 int double2 int "
                           "+double2 +double2 +double2 double",
    "stfmm2d": "int double int double2 int double2 int double2 double2 int +double2 +double2 "
               "+double2 int double2 int +double2 +double2 +double2 +int",
}


def _direct_spec(kind, var, pgt):
    """Identifier and argument list of {c,l,r,h}2d_direct{c,d,cd}{p,g,h}."""
    tp = "double" if kind == "r" else "dcomplex"
    spec = ["int"] + (["dcomplex"] if kind == "h" else []) + ["double2", "int"]
    spec += [tp + "2"] * len(var)
    if "d" in var and kind != "c":
        spec.append("double2")
    spec += ["double2", "int"] + ["+" + tp + "2"] * pgt + ["double"]
    return f"{kind}2d_direct{var}{'pgh'[pgt - 1]}", " ".join(spec)


for _k in "clrh":
    for _v in ("c", "d", "cd"):
        for _p in (1, 2, 3):
            _n, _s = _direct_spec(_k, _v, _p)
            SPECS[_n] = _s


def _fmm(kind, eps, zk, srcinfo, pg, targ, pgt, opts):
    """Shared body of cfmm2d.m / lfmm2d.m / rfmm2d.m / hfmm2d.m."""
    opts = opts or {}
    src = np.asfortranarray(np.asarray(srcinfo["sources"], dtype=float))
    if src.ndim != 2 or src.shape[0] != 2:
        raise ValueError("The first dimension of sources must be 2")
    ns, nd = src.shape[1], _nd(srcinfo)
    dt = float if kind == "r" else complex
    ng, nh = (1, 1) if kind == "c" else (2, 3)
    pot = np.zeros((nd, ns), dt, order="F")
    grad = np.zeros((nd * ng, ns), dt, order="F")
    hess = np.zeros((nd * nh, ns), dt, order="F")
    if targ is None:
        nt, pgt, targ = 0, 0, np.zeros((2, 1), order="F")
    else:
        targ = np.asfortranarray(np.asarray(targ, dtype=float))
        if targ.ndim != 2 or targ.shape[0] != 2:
            raise ValueError("First dimension of targets must be 2")
        nt = targ.shape[1]
    ntuse = max(nt, 1)
    pottarg = np.zeros((nd, ntuse), dt, order="F")
    gradtarg = np.zeros((nd * ng, ntuse), dt, order="F")
    hesstarg = np.zeros((nd * nh, ntuse), dt, order="F")
    if (pg == 0 and pgt == 0) or ns == 0:
        print("Nothing to compute, set eigher pg or pgt to 1 or 2")
        return Result()
    charges = _field(srcinfo, "charges", (nd, ns), dt)
    ifcharge = int(charges is not None)
    if charges is None:
        charges = np.zeros((nd, ns), dt, order="F")
    dipstr = _field(srcinfo, "dipstr", (nd, ns), dt)
    ifdipole = int(dipstr is not None or (kind != "c" and "dipvec" in srcinfo))
    if kind == "c":
        if dipstr is None:
            dipstr = np.zeros((nd, ns), dt, order="F")
        dipvec = None
    elif ifdipole:
        dipvec = _field(srcinfo, "dipvec", (2 * nd, ns), float)
        if dipstr is None or dipvec is None:
            raise ValueError("dipoles need both dipstr and dipvec")
    else:
        dipvec = np.zeros((2 * nd, ns), order="F")
        dipstr = np.zeros((nd, ns), dt, order="F")
    gw = gateway()
    ndiv, idivflag = gw.call("hndiv2d", SPECS["hndiv2d"], eps, ns, nt, ifcharge, ifdipole, pg, pgt, 20, 0)
    ndiv, idivflag = int(ndiv.real.ravel()[0]), int(idivflag.real.ravel()[0])
    ndiv = int(opts.get("ndiv", ndiv))
    idivflag = int(opts.get("idivflag", idivflag))
    ifnear = int(opts.get("ifnear", 1))
    iper, timeinfo, ier = 1, np.zeros(8), 0
    head = [nd, eps] + ([zk] if kind == "h" else []) + [ns, src, ifcharge, charges, ifdipole,
                                                          dipstr]
    head += [] if kind == "c" else [dipvec]
    out = gw.call(f"{kind}fmm2d_ndiv", SPECS[f"{kind}fmm2d_ndiv"], *head, iper, pg, pot, grad,
                  hess, nt, targ, pgt, pottarg, gradtarg, hesstarg, ndiv, idivflag, ifnear,
                  timeinfo, ier)
    pot, grad, hess, pottarg, gradtarg, hesstarg, timeinfo, ier = out
    F = lambda a, *s: squeeze(a.reshape(s, order="F"))
    U = Result(pot=None, grad=None, hess=None, pottarg=None, gradtarg=None, hesstarg=None)
    if pg >= 1:
        U["pot"] = F(pot, nd, ns)
    if pg >= 2:
        U["grad"] = F(grad, nd, ns) if kind == "c" else F(grad, nd, 2, ns)
    if pg >= 3:
        U["hess"] = F(hess, nd, ns) if kind == "c" else F(hess, nd, 3, ns)
    if pgt >= 1:
        U["pottarg"] = F(pottarg, nd, nt)
    if pgt >= 2:
        U["gradtarg"] = F(gradtarg, nd, nt) if kind == "c" else F(gradtarg, nd, 2, nt)
    if pgt >= 3:
        U["hesstarg"] = F(hesstarg, nd, nt) if kind == "c" else F(hesstarg, nd, 3, nt)
    U.ier = int(np.real(ier).ravel()[0])
    U.timeinfo = np.real(timeinfo).ravel()
    return U


def cfmm2d(eps, srcinfo, pg, targ=None, pgt=0, opts=None):
    """matlab/cfmm2d.m: Cauchy sums sum_j c_j log(z - z_j) + v_j / (z - z_j)."""
    return _fmm("c", eps, None, srcinfo, pg, targ, pgt, opts)


def lfmm2d(eps, srcinfo, pg, targ=None, pgt=0, opts=None):
    """matlab/lfmm2d.m: Laplace, complex charges and dipole strengths."""
    return _fmm("l", eps, None, srcinfo, pg, targ, pgt, opts)


def rfmm2d(eps, srcinfo, pg, targ=None, pgt=0, opts=None):
    """matlab/rfmm2d.m: Laplace, real charges and dipole strengths."""
    return _fmm("r", eps, None, srcinfo, pg, targ, pgt, opts)


def hfmm2d(eps, zk, srcinfo, pg, targ=None, pgt=0, opts=None):
    """matlab/hfmm2d.m: Helmholtz at wavenumber zk."""
    return _fmm("h", eps, complex(zk), srcinfo, pg, targ, pgt, opts)


# -- direct sums: matlab/{c,l,r,h}2ddir.m, st2ddir.m ------------------------

def _direct(kind, zk, srcinfo, targ, pgt):
    """Shared body of c2ddir.m / l2ddir.m / r2ddir.m / h2ddir.m."""
    thresh = 1e-15
    src = np.asfortranarray(np.asarray(srcinfo["sources"], dtype=float))
    ns, nd = src.shape[1], _nd(srcinfo)
    targ = np.asfortranarray(np.asarray(targ, dtype=float))
    if targ.ndim != 2 or targ.shape[0] != 2:
        raise ValueError("First dimension of targets must be 2")
    nt = targ.shape[1]
    if pgt == 0:
        print("Nothing to compute, set pgt to 1,2 or 3")
        return Result()
    dt = float if kind == "r" else complex
    ng, nh = (1, 1) if kind == "c" else (2, 3)
    charges = _field(srcinfo, "charges", (nd, ns), dt)
    dipstr = _field(srcinfo, "dipstr", (nd, ns), dt)
    dipvec = None if kind == "c" else _field(srcinfo, "dipvec", (2 * nd, ns), float)
    if kind != "c" and (dipstr is None) != (dipvec is None):
        raise ValueError("dipoles need both dipstr and dipvec")
    var = ("c" if charges is not None else "") + ("d" if dipstr is not None else "")
    if not var:
        raise ValueError("srcinfo has neither charges nor dipoles")
    outs = [np.zeros((nd, nt), dt, order="F"), np.zeros((nd * ng, nt), dt, order="F"),
            np.zeros((nd * nh, nt), dt, order="F")][:pgt]
    args = [nd] + ([complex(zk)] if kind == "h" else []) + [src, ns]
    if charges is not None:
        args.append(charges)
    if dipstr is not None:
        args.append(dipstr)
        if kind != "c":
            args.append(dipvec)
    args += [targ, nt] + outs + [thresh]
    res = gateway().call(*_direct_spec(kind, var, pgt), *args)
    U = Result(pottarg=res[0])
    if pgt >= 2:
        U["gradtarg"] = squeeze(res[1].reshape((nd, nt) if kind == "c" else (nd, 2, nt),
                                               order="F"))
    if pgt >= 3:
        U["hesstarg"] = squeeze(res[2].reshape((nd, nt) if kind == "c" else (nd, 3, nt),
                                               order="F"))
    return U


def c2ddir(srcinfo, targ, pgt):
    return _direct("c", None, srcinfo, targ, pgt)


def l2ddir(srcinfo, targ, pgt):
    return _direct("l", None, srcinfo, targ, pgt)


def r2ddir(srcinfo, targ, pgt):
    return _direct("r", None, srcinfo, targ, pgt)


def h2ddir(zk, srcinfo, targ, pgt):
    return _direct("h", zk, srcinfo, targ, pgt)


def st2ddir(srcinfo, targ, ifppregtarg=3):
    """matlab/st2ddir.m: Stokeslets (stoklet) and stresslets (strslet with
    strsvec) to velocity, pressure and velocity gradient at targ."""
    thresh = 1e-15
    src = np.asfortranarray(np.asarray(srcinfo["sources"], dtype=float))
    ns, nd = src.shape[1], _nd(srcinfo)
    targ = np.asfortranarray(np.asarray(targ, dtype=float))
    nt = targ.shape[1]
    if ifppregtarg == 0:
        print("Nothing to compute, set eigher ifppregtarg to 1 or 2 or 3")
        return Result()
    pottarg = np.zeros((2 * nd, nt), order="F")
    pretarg = np.zeros((nd, nt), order="F")
    gradtarg = np.zeros((4 * nd, nt), order="F")
    stoklet = _field(srcinfo, "stoklet", (2 * nd, ns), float)
    ifstoklet = int(stoklet is not None)
    if stoklet is None:
        stoklet = np.zeros((2 * nd, 1), order="F")
    strslet = _field(srcinfo, "strslet", (2 * nd, ns), float)
    strsvec = _field(srcinfo, "strsvec", (2 * nd, ns), float)
    ifstrslet = int(strslet is not None and strsvec is not None)
    if not ifstrslet:
        strslet = strsvec = np.zeros((2 * nd, 1), order="F")
    gw = gateway()
    if ifstoklet == 1 and ifstrslet == 0:
        out = gw.call("st2ddirectstokg", SPECS["st2ddirectstokg"], nd, src, stoklet, ns, targ,
                      nt, pottarg, pretarg, gradtarg, thresh)
    else:
        out = gw.call("st2ddirectstokstrsg", SPECS["st2ddirectstokstrsg"], nd, src, ifstoklet,
                      stoklet, ifstrslet, strslet, strsvec, ns, targ, nt, pottarg, pretarg,
                      gradtarg, thresh)
    pottarg, pretarg, gradtarg = out
    U = Result(pottarg=None, pretarg=None, gradtarg=None)
    if ifppregtarg >= 1:
        U["pottarg"] = squeeze(pottarg.reshape((nd, 2, nt), order="F"))
    if ifppregtarg >= 2:
        U["pretarg"] = pretarg
    if ifppregtarg >= 3:
        U["gradtarg"] = squeeze(gradtarg.reshape((nd, 2, 2, nt), order="F"))
    return U


def stfmm2d(eps, srcinfo, ifppreg, targ=None, ifppregtarg=0):
    """matlab/stfmm2d.m: Stokeslets (stoklet) and stresslets (strslet with
    strsvec) to velocity, pressure and velocity gradient at the sources
    (ifppreg 1/2/3) and at targ (ifppregtarg), through the gateway's
    stfmm2d_ (the biharmonic FMM of the library).  Output arrays are sized
    by the flags as the wrapper sizes them; no targets means nt = 0."""
    src = np.asfortranarray(np.asarray(srcinfo["sources"], dtype=float))
    ns, nd = src.shape[1], _nd(srcinfo)
    if targ is None:
        nt, ifppregtarg = 0, 0
        targ = np.zeros((2, 0), order="F")
    else:
        targ = np.asfortranarray(np.asarray(targ, dtype=float))
        nt = targ.shape[1]
    z = lambda m, n: np.zeros((m, n), order="F")
    ns_pot, ns_pre, ns_grad = (ns if ifppreg >= k else 1 for k in (1, 2, 3))
    nt_pot, nt_pre, nt_grad = (nt if ifppregtarg >= k else 1 for k in (1, 2, 3))
    if ifppreg == 0 and ifppregtarg == 0:
        print("Nothing to compute, set eigher ifppreg or ifppregtarg to 1 or 2 or 3")
        return Result()
    stoklet = _field(srcinfo, "stoklet", (2 * nd, ns), float)
    strslet = _field(srcinfo, "strslet", (2 * nd, ns), float)
    strsvec = _field(srcinfo, "strsvec", (2 * nd, ns), float)
    ifst, ifsr = int(stoklet is not None), int(strslet is not None and strsvec is not None)
    stoklet = stoklet if ifst else z(2 * nd, 1)
    if not ifsr:
        strslet = strsvec = z(2 * nd, 1)
    out = gateway().call(
        "stfmm2d", SPECS["stfmm2d"], nd, eps, ns, src,
        ifst, stoklet, ifsr, strslet, strsvec, ifppreg, z(2 * nd, ns_pot), z(nd, ns_pre),
        z(4 * nd, ns_grad), nt, targ if nt else z(2, 1), ifppregtarg, z(2 * nd, nt_pot),
        z(nd, nt_pre), z(4 * nd, nt_grad), 0)
    pot, pre, grad, pottarg, pretarg, gradtarg, ier = out
    U = Result(pot=None, pre=None, grad=None, pottarg=None, pretarg=None, gradtarg=None)
    U.ier = int(np.asarray(ier).ravel()[0])
    if ifppreg >= 1:
        U["pot"] = squeeze(pot.reshape((nd, 2, ns), order="F"))
    if ifppreg >= 2:
        U["pre"] = pre
    if ifppreg >= 3:
        U["grad"] = squeeze(grad.reshape((nd, 2, 2, ns), order="F"))
    if ifppregtarg >= 1:
        U["pottarg"] = squeeze(pottarg.reshape((nd, 2, nt), order="F"))
    if ifppregtarg >= 2:
        U["pretarg"] = pretarg
    if ifppregtarg >= 3:
        U["gradtarg"] = squeeze(gradtarg.reshape((nd, 2, 2, nt), order="F"))
    return U
