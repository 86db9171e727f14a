"""train_relaxpar (train/train_relaxpar.m:43-271), SIRT branch, on the engine.

Step 1 runs the method in chunks of 10 iterations with warm starts until the
error to ``x_ex`` stops decreasing; step 2 runs the augmented golden-section
search over (0, 2/rho) with the error bound (1+0.02)*minE.  Every probe is a
full device-resident SIRT run; only the error curves come back to the host.
"""
from __future__ import annotations

import numpy as np

from .operators import as_operator
from .sirt import sirt

_SIRT = ("landweber", "cimmino", "cav", "drop", "sart")


def _method_name(method):
    name = method if isinstance(method, str) else getattr(method, "__name__", "")
    if name not in _SIRT:
        raise ValueError(f"Unknown method {name}")
    return name


def train_relaxpar(A, b, x_ex, method, kmax, options=None):
    """relaxpar = train_relaxpar(A,b,x_ex,method,kmax,options)."""
    name = _method_name(method)
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
        X, info, _ = sirt(name, A, b, K, x0, opts)
        if relaxparmax is None:
            opts["rho"] = info["rho"]
            relaxparmax = 2.0 / info["rho"]
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
        X, _, _ = sirt(name, A, b, np.arange(1, kopt + 1), None, o2)
        ee = np.sqrt(np.sum((X - x_ex[:, None]) ** 2, axis=0))
        kk = int(np.sum((ee[:-1] > meub) & (np.diff(ee) < 0))) + 1
        return kk, float(ee[kk - 1])

    r = (3 - np.sqrt(5)) / 2                                 # :131-136
    alpha, beta = 0.0, relaxparmax
    alphap = alpha + r * (beta - alpha)
    betap = alpha + (1 - r) * (beta - alpha)
    k_a, e_a = probe(alphap)
    k_b, e_b = probe(betap)
    while abs(alphap - betap) > relaxparmax * 0.01:          # :160-229 (SIRT order)
        if e_a > meub or (not e_b > meub and k_a >= k_b):
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
