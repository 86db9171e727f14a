"""ART (art/art.m) and CART (cart/cart.m) on the engine.

    [X, info] = kaczmarz(A, b, K, x0, options)        art/kaczmarz.m
    [X, info] = symkaczmarz(A, b, K, x0, options)     art/symkaczmarz.m
    [X, info] = randkaczmarz(A, b, K, x0, options)    art/randkaczmarz.m
    [X, info] = art(art_method, A, b, K, x0, options)  art_method: a name or a row order
    [X, info] = columnaction(A, b, K, x0, options)    cart/columnaction.m
    [X, info] = cart(cart_method, A, b, K, x0, options)

Each sweep runs on the GPU in the reference's step order (the methods are
sequential: every step reads what the one before wrote).  Options follow
check_inputs.m: relaxpar (scalar; ART also a function of the step index
(k-1)*m + i), lbound, ubound, stoprule ('none', 'NCP' with res_dims and
ncp_smooth, 'DP' and 'ME' with taudelta),
damp, and for CART THR, Kbegin, Nunflag.  ``seed`` (no counterpart in the
reference) fixes the random row orders of randkaczmarz and CART's unflagging
draws, which come from a SplitMix64 stream instead of MATLAB's rand.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings

import numpy as np

from . import _lib
from ._lib import RowOptions, SirtInfo, check, f64, load, ptr
from .operators import as_operator
from .sirt import ncp_options

ART_METHODS = {"kaczmarz": 0, "symkaczmarz": 1, "randkaczmarz": 2}
ART_CUSTOM = 3
STOPS = {"NONE": 0, "NCP": 1, "DP": 2, "ME": 3}


def _prepare(A, b, K, x0, options):
    if A is None or b is None or K is None:
        raise ValueError("Too few input arguments.")
    A = as_operator(A)
    m, n = A.shape
    b = f64(b)
    if b.shape[0] != m:
        raise ValueError("The size of A and b do not match.")
    K = np.atleast_1d(np.asarray(K)).astype(np.int64).ravel()
    if K.size == 0:
        raise ValueError("Max no. of iterations must ALWAYS be specified.")
    if not np.all(np.diff(K) > 0):
        raise ValueError("Elements of K must be strictly increasing.")
    if x0 is not None and np.size(x0) > 0:
        x0 = f64(x0)
        if x0.shape[0] != n:
            raise ValueError("The size of x0 does not match the problem.")
    else:
        x0 = None
    return A, b, K, x0, dict(options or {})


def _options(A, opts, cart, sweeps):
    m, n = A.shape
    o = RowOptions()
    keep = []
    rp = opts.get("relaxpar")
    if rp is None or (np.size(rp) == 0 and not callable(rp)):
        o.relax_kind = _lib.RELAX["default"]
    elif callable(rp):
        if cart:
            raise ValueError("For CART methods the relaxpar must be a scalar.")
        idx = np.arange(1, sweeps * m + 1, dtype=np.float64)
        try:
            vals = np.asarray(rp(idx), np.float64).ravel()
        except Exception:
            vals = None
        if vals is None or vals.size != idx.size:
            vals = np.array([float(rp(int(i))) for i in idx])
        keep.append(np.ascontiguousarray(vals))
        o.relax_kind = _lib.RELAX["vector"]
        o.relax_vec = keep[-1].ctypes.data
    elif np.size(rp) == 1:
        o.relax_kind = _lib.RELAX["scalar"]
        o.relaxpar = float(np.ravel(rp)[0])
        if not 0 < o.relaxpar < 2:                     # calc_relaxpar.m:59-63, 78-82
            warnings.warn("The relaxpar value is outside the interval (0,2)", RuntimeWarning)
    else:
        raise ValueError("For ART methods the relaxpar must be a scalar or function handle." if not cart else
                         "For CART methods the relaxpar must be a scalar.")
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
    if stype not in STOPS:
        raise ValueError("The chosen stopping rule is not valid.")
    o.stoprule = STOPS[stype]
    o.taudelta = math.nan
    if stype in ("DP", "ME"):
        td = st.get("taudelta")
        if td is None or np.size(td) != 1:
            raise ValueError("For stopping rules DP and ME a scalar taudelta must be given in options.")
        o.taudelta = float(np.ravel(td)[0])
    ncp_options(st, stype, m, o)
    o.damp = float(opts.get("damp", 0.0))
    if o.damp < 0:
        raise ValueError("Damping must be positive")
    o.THR = float(opts.get("THR", 1e-4))
    o.Kbegin = int(opts.get("Kbegin", 10))
    o.Nunflag = int(opts.get("Nunflag", 0) or 0)
    o.seed = int(opts.get("seed", 0)) & (2 ** 64 - 1)
    return o, keep


def _run(fn, A, b, K, x0, o, sweeps, sym=False):
    n = A.shape[1]
    K32 = np.ascontiguousarray(K, np.int32)
    X = np.zeros((len(K), n))
    its = np.zeros(len(K) + 1, np.int32)
    nits = np.zeros(1, np.int32)
    rn = np.zeros(sweeps + 1)
    info = SirtInfo()
    check(fn(K32, X, its, nits, rn, info))
    done = info.finaliter // 2 if sym else info.finaliter          # sweeps run
    out = {"stoprule": info.stoprule, "finaliter": info.finaliter, "relaxpar": info.relaxpar,
           "itersaved": its[: nits[0]].astype(np.int64), "timetaken": info.timetaken,
           "resnorm": rn[: done + 1]}
    return X[: info.nsaved].T, out


def art(art_method, A, b, K, x0=None, options=None):
    """[X,info] = art(art_method,A,b,K,x0,options), art/art.m:1-330."""
    A, b, K, x0, opts = _prepare(A, b, K, x0, options)
    order = None
    if art_method is None or (isinstance(art_method, str) and art_method == ""):
        art_method = "kaczmarz"
    if isinstance(art_method, str):
        key = art_method.lower()
        if key not in ART_METHODS:
            raise ValueError("Unknown ART method specified")
        code = ART_METHODS[key]
    else:
        code = ART_CUSTOM
        order = np.ascontiguousarray(np.ravel(np.asarray(art_method)), np.int32)
    sweeps = int(K[-1])
    if code == ART_METHODS["symkaczmarz"]:
        if np.any(K % 2):
            raise ValueError("For symkaczmarz only even iteration numbers can be requested.")
        sweeps //= 2
    o, keep = _options(A, opts, False, sweeps)
    h = A.engine.h

    def fn(K32, X, its, nits, rn, info):
        return load().air_art(h, code, ptr(order), 0 if order is None else order.size, o, ptr(b), ptr(x0),
                              ptr(K32), len(K32), ptr(X), ptr(its), nits.ctypes.data_as(C.POINTER(C.c_int32)),
                              ptr(rn), info)
    return _run(fn, A, b, K, x0, o, sweeps, code == ART_METHODS["symkaczmarz"])


def kaczmarz(A, b, K, x0=None, options=None):
    """art/kaczmarz.m"""
    return art("kaczmarz", A, b, K, x0, options)


def symkaczmarz(A, b, K, x0=None, options=None):
    """art/symkaczmarz.m"""
    return art("symkaczmarz", A, b, K, x0, options)


def randkaczmarz(A, b, K, x0=None, options=None):
    """art/randkaczmarz.m"""
    return art("randkaczmarz", A, b, K, x0, options)


def cart(cart_method, A, b, K, x0=None, options=None):
    """[X,info] = cart(cart_method,A,b,K,x0,options), cart/cart.m:1-323."""
    A, b, K, x0, opts = _prepare(A, b, K, x0, options)
    order = None
    if cart_method is None or (isinstance(cart_method, str) and cart_method == ""):
        cart_method = "columnaction"
    if isinstance(cart_method, str):
        if cart_method.lower() != "columnaction":
            raise ValueError("Unknown CART method specified")
    else:
        order = np.ascontiguousarray(np.ravel(np.asarray(cart_method)), np.int32)
    sweeps = int(K[-1])
    o, keep = _options(A, opts, True, sweeps)
    h = A.engine.h

    def fn(K32, X, its, nits, rn, info):
        return load().air_cart(h, ptr(order), 0 if order is None else order.size, o, ptr(b), ptr(x0), ptr(K32),
                               len(K32), ptr(X), ptr(its), nits.ctypes.data_as(C.POINTER(C.c_int32)), ptr(rn),
                               info)
    return _run(fn, A, b, K, x0, o, sweeps)


def columnaction(A, b, K, x0=None, options=None):
    """cart/columnaction.m"""
    return cart("columnaction", A, b, K, x0, options)
