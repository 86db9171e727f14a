"""The SIRT driver and its method wrappers, with the call surface of AIR Tools II.

    X, info, ext_info = sirt(sirt_method, A, b, K, x0=None, options=None)
    X, info, ext_info = cimmino(A, b, K, x0=None, options=None)   # also cav, drop, landweber, sart

mirror sirt/sirt.m:1-281 and the one-line wrappers sirt/cimmino.m:121,
sirt/cav.m:122, sirt/drop.m:122, sirt/landweber.m:119, sirt/sart.m:129.
``options`` is a dict with the MATLAB struct's fields (check_inputs.m:101-260):
``lbound``, ``ubound``, ``relaxpar`` (scalar, 'line', 'psi1', 'psi1mod',
'psi2', 'psi2mod', or a per-iteration vector), ``rho``, ``stoprule``
(dict with ``type``, ``taudelta``, ``res_dims``, ``ncp_smooth``), ``verbose``.
Engine-only keys: ``graph`` (iterations per captured HIP graph, 0 = direct
launches), and ``mlp``/``noise_norm``/``x_true`` for the learned relaxation.

The whole K-loop runs on the GPU (libairhip.so); this module only validates
arguments and converts results, raising ``ValueError`` with the reference's
messages where check_inputs.m calls ``error``.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from . import _lib
from ._lib import METHODS, RELAX, STOP, SirtInfo, SirtOptions, check, f64, load, ptr
from .operators import AirOperator, as_operator


class ExtInfo:
    """ext_info of sirt.m:171-191: diagonals of M and D (computed on access)."""

    def __init__(self, A, method, Mc, Dc):
        self._A, self._method, self._Mc, self._Dc = A, method, Mc, Dc
        self._M = self._D = None

    def _fill(self):
        if self._M is not None:
            return
        m, n = self._A.shape
        if self._method == METHODS["custom"]:
            self._M = np.ones(m) if self._Mc is None else self._Mc.copy()
            self._D = np.ones(n) if self._Dc is None else self._Dc.copy()
        else:
            self._M, self._D = np.empty(m), np.empty(n)
            check(load().air_weights(self._A.engine.h, self._method, ptr(self._M), ptr(self._D)))

    @property
    def M(self):
        self._fill()
        return self._M

    @property
    def D(self):
        self._fill()
        return self._D

    def __getitem__(self, k):
        return getattr(self, k)


def _method_code(sirt_method, m, n):
    """get_mfun_dfun.m:47-310 dispatch -> (code, M vector or None, D vector or None)."""
    if sirt_method is None or (isinstance(sirt_method, str) and sirt_method == ""):
        sirt_method = "sart"                                   # sirt.m:138-140
    if isinstance(sirt_method, str):
        if sirt_method not in ("landweber", "cimmino", "cav", "drop", "sart"):
            raise ValueError("SIRT method not defined.")
        return METHODS[sirt_method], None, None, sirt_method
    if isinstance(sirt_method, dict):
        def vec(key, size):
            v = sirt_method.get(key)
            if v is None or np.size(v) == 0:
                return None
            a = np.asarray(v, dtype=np.float64)
            if a.ndim == 2 and min(a.shape) > 1:
                if np.count_nonzero(a - np.diag(np.diag(a))) == 0:
                    a = np.diag(a)
                else:
                    raise NotImplementedError(f"non-diagonal {key} matrices are not supported by the engine")
            a = np.ravel(a)
            if a.size != size:
                raise ValueError(f"{key} must hold {size} diagonal entries")
            return np.ascontiguousarray(a)
        return METHODS["custom"], vec("M", m), vec("D", n), "custom"
    raise ValueError("First input must be a string with the SIRT method name or a struct with fields M and D "
                     "characterizing a custom method.")


def _relax(relaxpar):
    """calc_relaxpar.m:105-171 strategy dispatch (strncmpi prefixes)."""
    if relaxpar is None or (isinstance(relaxpar, (list, tuple, np.ndarray)) and np.size(relaxpar) == 0):
        return RELAX["default"], math.nan, None
    if isinstance(relaxpar, str):
        s = relaxpar.lower()
        if s[:4] == "line":
            return RELAX["line"], math.nan, None
        if s[:4] == "psi1":
            return (RELAX["psi1mod"] if s[:7] == "psi1mod" else RELAX["psi1"]), math.nan, None
        if s[:4] == "psi2":
            return (RELAX["psi2mod"] if s[:7] == "psi2mod" else RELAX["psi2"]), math.nan, None
        if s == "learned":          # the network (options['mlp'] parameters); no reference counterpart
            return RELAX["learned"], math.nan, None
        if s == "greedy":           # training trajectories (options['x_true']); no reference counterpart
            return RELAX["greedy"], math.nan, None
        raise ValueError("The chosen relaxation strategy is not valid for this method.")
    arr = np.ravel(np.asarray(relaxpar, dtype=np.float64))
    if arr.size == 1:
        return RELAX["scalar"], float(arr[0]), None
    return RELAX["vector"], math.nan, np.ascontiguousarray(arr)


def _lib_mlp_params():
    return 5378


def _build_options(A, sirt_method, options, kmax):
    m, n = A.shape
    o = SirtOptions()
    keep = []
    code, Mc, Dc, name = _method_code(sirt_method, m, n)
    o.method = code
    if Mc is not None:
        keep.append(Mc); o.M = Mc.ctypes.data
    if Dc is not None:
        keep.append(Dc); o.D = Dc.ctypes.data
    opts = dict(options or {})
    o.relax_kind, o.relaxpar, rv = _relax(opts.get("relaxpar"))
    o.noise_norm = float(opts.get("noise_norm", 0.0) or 0.0)
    if o.relax_kind == RELAX["learned"] and opts.get("mlp") is None:
        raise ValueError("relaxpar 'learned' needs options['mlp'] (the network parameters)")
    if o.relax_kind in (RELAX["learned"], RELAX["greedy"]) and opts.get("mlp") is not None:
        pm = np.ascontiguousarray(np.asarray(opts["mlp"], np.float32).ravel())
        if pm.size != _lib_mlp_params():
            raise ValueError("options['mlp'] must hold 5378 parameters")
        keep.append(pm); o.mlp_params = pm.ctypes.data
    if o.relax_kind == RELAX["greedy"] and opts.get("x_true") is None:
        raise ValueError("relaxpar 'greedy' needs options['x_true']")
    if o.relax_kind in (RELAX["greedy"], RELAX["learned"]) and opts.get("x_true") is not None:
        xt = f64(opts["x_true"])
        if xt.size != n:
            raise ValueError("options['x_true'] must have n values")
        keep.append(xt); o.x_true = xt.ctypes.data
    if rv is not None:
        if rv.size < kmax:
            raise ValueError("a relaxpar vector must hold at least K(end) values")
        keep.append(rv); o.relax_vec = rv.ctypes.data
    o.rho = float(opts.get("rho", math.nan)) if opts.get("rho") is not None else math.nan
    for key, kind, val, vec in (("lbound", "lb_kind", "lb", "lb_vec"), ("ubound", "ub_kind", "ub", "ub_vec")):
        if key in opts and opts[key] is not None and np.size(opts[key]) > 0:
            bv = f64(opts[key])
            if bv.size == 1:
                setattr(o, kind, _lib.BOUND_SCALAR); setattr(o, val, float(bv[0]))
            elif bv.size == n:
                keep.append(bv); setattr(o, kind, _lib.BOUND_VECTOR); setattr(o, vec, bv.ctypes.data)
            else:
                which = "lbound (lower bound)" if key == "lbound" else "ubound (upper bound)"
                raise ValueError(f"{which} must either be scalar or same length as vector of unknowns x.")
    st = opts.get("stoprule") or {}
    stype = str(st.get("type", "none")).upper()
    if stype not in STOP:
        raise ValueError("The chosen stopping rule is not valid.")
    o.stoprule = STOP[stype]
    o.taudelta = math.nan
    if stype in ("DP", "ME"):
        td = st.get("taudelta")
        if td is None or np.size(td) != 1:
            raise ValueError("For stopping rules DP and ME a scalar taudelta must be given in options.")
        o.taudelta = float(np.ravel(td)[0])
    ncp_options(st, stype, m, o)
    o.verbose = int(opts.get("verbose", 0) or 0)
    return o, keep, code, Mc, Dc, name


def ncp_options(st, stype, m, o):
    """res_dims and ncp_smooth of the NCP rule into an options struct
    (check_inputs.m:171-180, check_stoprules.m:99-120)."""
    if stype == "NCP":
        if "res_dims" not in st:
            raise ValueError("options.stoprule.res_dims must be specified for NCP stopping rule.")
        rd = np.atleast_1d(np.asarray(st["res_dims"])).astype(np.int64).ravel()
        if rd.size == 1:            # 1-D residual: one column of length m (check_stoprules.m:99-105)
            o.res_dims[0], o.res_dims[1] = int(m), 0
        elif rd.size == 2:
            o.res_dims[0], o.res_dims[1] = int(rd[0]), int(rd[1])
        else:
            raise ValueError("NCP stopping rule only implemented for 1D and 2D residuals, the size of which "
                             "should be given using options.stoprule.res_dims.")
    o.ncp_smooth = int(st.get("ncp_smooth", 2))


def sirt(sirt_method, A, b, K, x0=None, options=None):
    """[X,info,ext_info] = sirt(sirt_method,A,b,K,x0,options), sirt/sirt.m."""
    if A is None or b is None or K is None:
        raise ValueError("Too few input arguments.")
    if not hasattr(A, "shape"):
        A = as_operator(A)                            # a function handle: FunctionOperator
    m, n = A.shape
    b = f64(b)
    if b.shape[0] != m:
        raise ValueError("The size of A and b do not match.")
    K = np.atleast_1d(np.asarray(K)).astype(np.int64).ravel()
    if K.size == 0:
        raise ValueError("Max no. of iterations must ALWAYS be specified.")
    if not np.all(np.diff(K) > 0):
        raise ValueError("Elements of K must be strictly increasing.")
    kmax = int(K[-1])
    if x0 is not None and np.size(x0) > 0:
        x0 = f64(x0)
        if x0.shape[0] != n:
            raise ValueError("The size of x0 does not match the problem.")
    else:
        x0 = None
    opts = dict(options or {})
    o, keep, code, Mc, Dc, name = _build_options(A, sirt_method, opts, kmax)
    A = as_operator(A)
    check(load().air_set_graph(A.engine.h, int(opts.get("graph", 0) or 0)))
    K32 = np.ascontiguousarray(K, np.int32)
    X = np.zeros((len(K), n))
    its = np.zeros(len(K) + 1, np.int32)
    nits = np.zeros(1, np.int32)
    rn = np.zeros(kmax + 1)
    info = SirtInfo()
    check(load().air_sirt(A.engine.h, o, ptr(b), ptr(x0), ptr(K32), len(K), ptr(X), ptr(its),
                          nits.ctypes.data_as(_lib.C.POINTER(_lib.C.c_int32)), ptr(rn), info))
    X = X[: info.nsaved].T
    rp_in = opts.get("relaxpar")
    if info.finaliter == 0:
        relaxpar = rp_in                              # check_stoprules at k = 0 sees relaxparinput
    else:
        relaxpar = info.relaxpar
    if o.relax_kind == RELAX["scalar"] and not (0 < o.relaxpar < 2 / info.rho):
        warnings.warn(f"The relaxpar value is outside the interval (0,{2 / info.rho:f})", RuntimeWarning)
    out = {"stoprule": info.stoprule, "finaliter": info.finaliter, "relaxpar": relaxpar, "rho": info.rho,
           "itersaved": its[: nits[0]].astype(np.int64), "timetaken": info.timetaken,
           "resnorm": rn[: info.finaliter + 1]}
    if o.verbose:
        _print_verbose(o.verbose, out, kmax, info)
    return X, out, ExtInfo(A, code, Mc, Dc)


def _print_verbose(verbose, info, kmax, cinfo):
    """sirt.m:252-260 lines, printed after the device loop has finished."""
    fin = info["finaliter"]
    per = cinfo.iter_time / max(fin, 1)
    width = max(1, math.ceil(math.log10(max(kmax, 1))))
    for k in range(1, fin + 1):
        stop = k == fin
        if verbose == 1 or (verbose > 1 and (k == 1 or stop or k % verbose == 1)):
            print(f"{k:{width}d}/{kmax}:  resnorm={info['resnorm'][k]:1.2e}  relaxpar={float(info['relaxpar'] or 0):1.2e}"
                  f"  time(iter/total)={per:1.2e}/{cinfo.setup_time + per * k:1.2e} secs.")


def landweber(A, b, K, x0=None, options=None):
    """sirt/landweber.m:119."""
    return sirt("landweber", A, b, K, x0, options)


def cimmino(A, b, K, x0=None, options=None):
    """sirt/cimmino.m:121."""
    return sirt("cimmino", A, b, K, x0, options)


def cav(A, b, K, x0=None, options=None):
    """sirt/cav.m:122."""
    return sirt("cav", A, b, K, x0, options)


def drop(A, b, K, x0=None, options=None):
    """sirt/drop.m:122."""
    return sirt("drop", A, b, K, x0, options)


def sart(A, b, K, x0=None, options=None):
    """sirt/sart.m:129."""
    return sirt("sart", A, b, K, x0, options)


def weights(method, A):
    """Diagonals (M, D) of get_mfun_dfun.m for a named method, FP64."""
    A = as_operator(A)
    m, n = A.shape
    M, D = np.empty(m), np.empty(n)
    check(load().air_weights(A.engine.h, METHODS[method], ptr(M), ptr(D)))
    return M, D


def spectral_radius(method, A, M=None, D=None):
    """rho of calc_relaxpar.m:91-103 (largest eigenvalue of D A' M A), by Lanczos on the GPU."""
    A = as_operator(A)
    code = METHODS["custom"] if method in (None, "custom") else METHODS[method]
    Mv = None if M is None else f64(M)
    Dv = None if D is None else f64(D)
    r = _lib.C.c_double()
    check(load().air_spectral_radius(A.engine.h, code, ptr(Mv), ptr(Dv), _lib.C.byref(r)))
    return r.value


def calczeta(kmax):
    """z_k for k = 2..kmax-1 (calc_relaxpar.m:179-244)."""
    z = np.empty(max(kmax - 2, 1))
    check(load().air_calczeta(int(kmax), ptr(z)))
    return z[: kmax - 2]


class SirtRun:
    """Device-resident SIRT for callers that time or drive the iterations
    themselves: prepare (weights, rho, r0) once, then ``iterate(n)``."""

    def __init__(self, sirt_method, A, b, kmax, x0=None, options=None):
        self.A = as_operator(A)
        self.kmax = int(kmax)
        self._o, self._keep, *_ = _build_options(self.A, sirt_method, options, self.kmax)
        self._b = f64(b)
        self._x0 = None if x0 is None else f64(x0)
        check(load().air_set_graph(self.A.engine.h, int(dict(options or {}).get("graph", 0) or 0)))
        check(load().air_sirt_prepare(self.A.engine.h, self._o, ptr(self._b), ptr(self._x0), self.kmax))

    def iterate(self, niter):
        ms = _lib.C.c_double()
        check(load().air_sirt_iterate(self.A.engine.h, int(niter), _lib.C.byref(ms)))
        return ms.value

    def fetch(self):
        x = np.empty(self.A.shape[1])
        rn = np.zeros(self.kmax + 1)
        info = SirtInfo()
        check(load().air_sirt_fetch(self.A.engine.h, ptr(x), ptr(rn), info))
        return x, {"stoprule": info.stoprule, "finaliter": info.finaliter, "relaxpar": info.relaxpar,
                   "rho": info.rho, "resnorm": rn[: info.finaliter + 1], "setup_time": info.setup_time,
                   "iter_time": info.iter_time}

    def records(self):
        """Per-iteration records of a learned / greedy run: features of the state
        before each update (k-1 = 0..), greedy targets lambda* rho, network
        outputs and ||x_{k-1} - x_true||."""
        K = self.kmax
        feat = np.full((K, 16), np.nan, np.float32)
        target = np.full(K, np.nan, np.float32)
        out = np.full((K, 2), np.nan, np.float32)
        err = np.full(K, np.nan)
        check(load().air_sirt_records(self.A.engine.h, ptr(feat), ptr(target), ptr(out), ptr(err)))
        return {"features": feat, "target": target, "out": out, "err": err}

    def kernel_timing(self, on=True):
        """Per-kernel event timing: False/0 off, True/1 every iteration, an
        integer s > 1 every s-th iteration."""
        stride = (1 if on else 0) if isinstance(on, bool) else int(on)
        check(load().air_set_kernel_timing(self.A.engine.h, stride))

    def kernel_times(self):
        c, r, d, n = _lib.C.c_double(), _lib.C.c_double(), _lib.C.c_double(), _lib.C.c_int64()
        check(load().air_get_kernel_timing(self.A.engine.h, _lib.C.byref(c), _lib.C.byref(r), _lib.C.byref(d),
                                           _lib.C.byref(n)))
        return {"cols_ms": c.value, "rows_ms": r.value, "reduce_ms": d.value, "launches": n.value}


def sirt_multi(sirt_method, A, B, K, X0=None, options=None):
    """SIRT for the columns of B (m x nrhs, nrhs <= 64) sharing A, every pass
    reading A once for all of them (SpMM on the device).  Options as for sirt
    with the stop rule 'none', a default or scalar relaxpar and scalar bounds.
    Returns (X n x nrhs, info with 'resnorm' per column and 'gpu_ms')."""
    A = as_operator(A)
    m, n = A.shape
    B = np.asarray(B, np.float64)
    if B.ndim == 1:
        B = B[:, None]
    if B.shape[0] != m or not 1 <= B.shape[1] <= 64:
        raise ValueError("B must be m x nrhs with 1 <= nrhs <= 64")
    nrhs = B.shape[1]
    K = int(np.ravel(K)[-1])
    o, keep, *_ = _build_options(A, sirt_method, options, K)
    Bf = np.asfortranarray(B)
    X0f = None if X0 is None else np.asfortranarray(np.asarray(X0, np.float64).reshape(n, nrhs))
    X = np.zeros((n, nrhs), order="F")
    rn = np.zeros(nrhs)
    ms = _lib.C.c_double()
    check(load().air_sirt_multi(A.engine.h, o, nrhs, _lib.C.c_void_p(Bf.ctypes.data),
                                None if X0f is None else _lib.C.c_void_p(X0f.ctypes.data), K,
                                _lib.C.c_void_p(X.ctypes.data), ptr(rn), _lib.C.byref(ms)))
    return X, {"resnorm": rn, "gpu_ms": ms.value, "finaliter": K}
