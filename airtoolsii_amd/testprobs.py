"""Tomography test problems of testprobs/ traced on the GPU.

``paralleltomo`` (testprobs/paralleltomo.m) and ``fanlineartomo``
(testprobs/fanlineartomo.m) keep the reference argument order and defaults.
With ``isMatrix=1`` the system matrix is traced by the HIP ray tracer into
CSR (and CSC) in HBM; with ``isMatrix=0`` only the ray geometry is kept and
every product re-traces the rays (the reference's matrix-free mode).
``dtype`` chooses the value storage ('f32', 'bf16', 'f16', 'f64').
"""
from __future__ import annotations

import numpy as np

from ._lib import DTYPES, Engine, check, f64, load, ptr
from .operators import AirOperator
from .phantoms import _matlab_round, shepplogan


def _finish(A, return_all, extra):
    if not return_all:
        return A
    x = shepplogan(A.geometry["N"]).ravel(order="F")
    b = A.matvec(x)
    return (A, b, x) + extra


def paralleltomo(N, theta=None, p=None, d=None, isDisp=0, isMatrix=1, *, dtype="f32", device=0,
                 return_all=False):
    """[A,b,x,theta,p,d] = paralleltomo(N,theta,p,d,isDisp,isMatrix), paralleltomo.m:62-110."""
    N = int(N)
    theta = np.arange(0.0, 180.0) if theta is None or np.size(theta) == 0 else f64(theta)
    p = int(_matlab_round(np.sqrt(2) * N)) if p is None else int(p)
    d = float(p - 1) if d is None else float(d)
    e = Engine(device)
    check(load().air_build_paralleltomo(e.h, N, ptr(theta), len(theta), p, d, 0 if isMatrix else 1,
                                        DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix, {"kind": "parallel", "N": N, "theta": theta, "p": p, "d": d})
    return _finish(A, return_all, (theta, p, d))


def fanlineartomo(N, theta=None, p=None, R=None, dw=None, sd=None, isDisp=0, isMatrix=1, *, dtype="f32",
                  device=0, return_all=False):
    """[A,b,x,theta,p,R,dw,sd] = fanlineartomo(N,theta,p,R,dw,sd,isDisp,isMatrix),
    fanlineartomo.m:70-136 (R, dw, sd in units of N)."""
    N = int(N)
    sd = 3.0 if sd is None else float(sd)
    dw = 2.5 if dw is None else float(dw)
    R = 2.0 if R is None else float(R)
    p = int(_matlab_round(np.sqrt(2) * N)) if p is None else int(p)
    theta = np.arange(0.0, 359.0, 2.0) if theta is None or np.size(theta) == 0 else f64(theta)
    if R * N < np.sqrt(2) / 2 * N:
        raise ValueError("R must be greater than half squareroot 2")
    e = Engine(device)
    check(load().air_build_fanlineartomo(e.h, N, ptr(theta), len(theta), p, R, dw, sd, 0 if isMatrix else 1,
                                         DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix,
                    {"kind": "fanlinear", "N": N, "theta": theta, "p": p, "R": R, "dw": dw, "sd": sd})
    # fanlineartomo.m:75-90,133-135: dw and sd come back multiplied by N, R in units of N
    return _finish(A, return_all, (theta, p, R, dw * N, sd * N))
