"""train_relaxpar (train/train_relaxpar.m:43-271) and train_dpme
(train/train_dpme.m:1-137), SIRT branch, on the engine.

train_relaxpar: step 1 runs the method in chunks of 10 iterations with warm
starts until the error to ``x_ex`` stops decreasing; step 2 runs the augmented
golden-section search over (0, 2/rho) with the error bound (1+0.02)*minE.
train_dpme: for s noisy right-hand sides it finds the iteration of smallest
error in the same chunked way and returns the mean of the tau that makes the
DP or ME rule stop there.  Every run is a device-resident SIRT run; iterates
come back to the host only to measure their error.
"""
from __future__ import annotations

import numpy as np

from .operators import as_operator
from .art import art, cart
from .sirt import sirt

_SIRT = ("landweber", "cimmino", "cav", "drop", "sart")


ROW_ACTION = ("kaczmarz", "symkaczmarz", "randkaczmarz", "columnaction")


def _method_name(method):
    name = method if isinstance(method, str) else getattr(method, "__name__", "")
    if name not in _SIRT and name not in ROW_ACTION:
        raise ValueError(f"Unknown method {name}")
    return name


def _run(name, A, b, K, x0, opts):
    """[X, info] of the named SIRT, ART or CART method."""
    if name == "columnaction":
        return cart(name, A, b, K, x0, opts)
    if name in ROW_ACTION:
        return art(name, A, b, K, x0, opts)
    X, info, _ = sirt(name, A, b, K, x0, opts)
    return X, info


def train_relaxpar(A, b, x_ex, method, kmax, options=None):
    """relaxpar = train_relaxpar(A,b,x_ex,method,kmax,options), train_relaxpar.m:43-271.

    SIRT methods start from the default relaxpar with relaxparmax = 2/rho; the
    ART methods and columnaction (:90-96) start from relaxpar 0.25 with
    relaxparmax = 2, and their golden section search tests the upper probe
    first (:213-268)."""
    name = _method_name(method)
    is_sirt = name not in ROW_ACTION

    def run(K_, x0_, opts_):
        return _run(name, A, b, K_, x0_, opts_)
    A = as_operator(A)
    x_ex = np.ravel(np.asarray(x_ex, np.float64))
    if options is not None:
        opts = dict(options)
        opts["stoprule"] = {"type": "none"}                  # :44-49
        opts.pop("relaxpar", None)
    else:
        opts = {}
    fudge = 0.02                                             # :54
    stepk = 10
    K = np.arange(1, stepk + 1)
    k = 0
    Er = []
    x0 = None
    relaxparmax = None
    while True:                                              # :65-122
        if k + stepk > kmax:
            K = np.arange(1, kmax - k + 1)
            k = kmax
        else:
            k = k + stepk
        if relaxparmax is None and not is_sirt:
            opts["relaxpar"] = 0.25
        X, info = run(K, x0, opts)
        if relaxparmax is None:
            if is_sirt:
                opts["rho"] = info["rho"]
                relaxparmax = 2.0 / info["rho"]
            else:
                relaxparmax = 2.0
        Er.extend(np.sqrt(np.sum((X - x_ex[:, None]) ** 2, axis=0)).tolist())
        kopt = int(np.argmin(Er)) + 1
        minE = Er[kopt - 1]
        if kopt == k and k < kmax:
            x0 = X[:, -1]
        else:
            break
    meub = (1 + fudge) * minE                                # :125

    def probe(lam):
        o2 = dict(opts)
        o2["relaxpar"] = lam
        X, _ = run(np.arange(1, kopt + 1), None, o2)
        ee = np.sqrt(np.sum((X - x_ex[:, None]) ** 2, axis=0))
        kk = int(np.sum((ee[:-1] > meub) & (np.diff(ee) < 0))) + 1
        return kk, float(ee[kk - 1])

    r = (3 - np.sqrt(5)) / 2                                 # :131-136
    alpha, beta = 0.0, relaxparmax
    alphap = alpha + r * (beta - alpha)
    betap = alpha + (1 - r) * (beta - alpha)
    k_a, e_a = probe(alphap)
    k_b, e_b = probe(betap)
    while abs(alphap - betap) > relaxparmax * 0.01:
        if is_sirt:
            right = e_a > meub or (not e_b > meub and k_a >= k_b)     # :160-211
        else:
            right = not e_b > meub and (e_a > meub or k_a >= k_b)     # :213-268
        if right:
            z = alphap + (1 - r) * abs(beta - alphap)
            alpha, alphap, betap = alphap, betap, z
            k_a, e_a = k_b, e_b
            k_b, e_b = probe(betap)
        else:
            z = alpha + r * abs(betap - alpha)
            beta, betap, alphap = betap, alphap, z
            k_b, e_b = k_a, e_a
            k_a, e_a = probe(alphap)
    return (alphap + betap) / 2                              # :270


def train_dpme(A, b_exact, x_exact, method, type, delta, s, kmax, options=None, *, noise=None, seed=0):
    """tau = train_dpme(A,b_exact,x_exact,method,type,delta,s,kmax,options).

    The reference draws the noise with MATLAB's randn (train_dpme.m:63-65),
    which is not reproduced: ``noise`` (m x s) gives the unscaled noise
    columns, otherwise they come from NumPy's PCG64 with ``seed``.  Each
    column is scaled to norm ``delta``.  As in the reference, X keeps the
    starting vector as its first column (:82,99), so the residual at the
    optimal iteration kopt is that of iterate kopt-1 (:120) and the ME rule
    also uses iterate kopt-2 (:126)."""
    name = _method_name(method)
    kind = str(type).upper()[:2]
    if kind not in ("DP", "ME"):
        raise ValueError("The chosen type is invalid")
    A = as_operator(A)
    m = A.shape[0]
    b_exact = np.ravel(np.asarray(b_exact, np.float64))
    x_exact = np.ravel(np.asarray(x_exact, np.float64))
    opts = dict(options) if options is not None else {}
    opts["stoprule"] = {"type": "none"}                      # :45-50
    E = np.random.default_rng(seed).standard_normal((m, s)) if noise is None else np.asarray(noise, np.float64)
    if E.shape != (m, s):
        raise ValueError("noise must be m x s")
    E = delta * E / np.sqrt(np.sum(E * E, axis=0))[None, :]  # :64
    kstep = kmax if kmax < 10 else 10                        # :71-75
    tauh = np.zeros(s)
    for i in range(s):                                       # :78-134
        bi = b_exact + E[:, i]
        cols = [np.zeros(x_exact.shape[0])]
        x0 = np.zeros(x_exact.shape[0])
        k = kstep
        K = np.arange(1, k + 1)
        Ek = []
        while True:
            Xny, _ = _run(name, A, bi, K, x0, opts)
            Ek.extend(np.sqrt(np.sum((Xny - x_exact[:, None]) ** 2, axis=0)).tolist())
            cols.extend(Xny.T)
            I = int(np.argmin(Ek)) + 1
            if I != k or k >= kmax:
                kopt = I
                break
            if k + kstep > kmax:
                K = np.arange(1, kmax - k + 1)
                k = kmax
            else:
                k = k + kstep
            x0 = Xny[:, -1]
        rk = bi - A.matvec(cols[kopt - 1])                   # X(:,kopt)
        if kind == "DP":
            tauh[i] = np.linalg.norm(rk) / delta
        else:
            if kopt < 2:
                raise IndexError("Index in position 2 is invalid: ME needs X(:,kopt-1) with kopt >= 2")
            rkm1 = bi - A.matvec(cols[kopt - 2])
            tauh[i] = 0.5 * (rkm1 @ (rkm1 + rk)) / (delta * np.linalg.norm(rkm1))
    return float(np.mean(tauh))
