"""Tomography test problems of testprobs/ traced on the GPU.

``paralleltomo`` (testprobs/paralleltomo.m), ``fanlineartomo``
(testprobs/fanlineartomo.m), ``fancurvedtomo`` (testprobs/fancurvedtomo.m) and
``seismictomo`` (testprobs/seismictomo.m) keep the reference argument order
and defaults; ``sphericaltomo`` (testprobs/sphericaltomo.m) and
``seismicwavetomo`` (testprobs/seismicwavetomo.m) build their circle-quadrature
and Fresnel-zone matrices on the GPU as well (matrix form only).
With ``isMatrix=1`` the system matrix is traced by the HIP ray tracer into
CSR (and CSC) in HBM; with ``isMatrix=0`` only the ray geometry is kept and
every product re-traces the rays (the reference's matrix-free mode).
``dtype`` chooses the value storage ('f32', 'bf16', 'f16', 'f64').
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import DTYPES, Engine, check, f64, load, ptr
from .operators import AirOperator
from .phantoms import _matlab_round, shepplogan, tectonic


def _finish(A, return_all, extra, phantom=shepplogan):
    if not return_all:
        return A
    x = phantom(A.geometry["N"]).ravel(order="F")
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


def fancurvedtomo(N, theta=None, p=None, R=None, d=None, isDisp=0, isMatrix=1, *, dtype="f32", device=0,
                  return_all=False):
    """[A,b,x,theta,p,R,d] = fancurvedtomo(N,theta,p,R,d,isDisp,isMatrix),
    fancurvedtomo.m:69-137 (R in units of N, d in degrees)."""
    N = int(N)
    R = 2.0 if R is None else float(R)
    if R < math.sqrt(2) / 2:
        raise ValueError("R must be greater than half squareroot 2")
    # fancurvedtomo.m:79-83: first and last rays touch the corners (-N/2,N/2), (N/2,N/2)
    d = 2.0 * ((180.0 / math.pi) * math.atan(1.0 / (2.0 * R - 1.0))) if d is None else float(d)
    p = int(_matlab_round(np.sqrt(2) * N)) if p is None else int(p)
    theta = np.arange(0.0, 359.0, 2.0) if theta is None or np.size(theta) == 0 else f64(theta)
    if d < 0 or d > 180:
        raise ValueError("The angle of the source must be in the interval [0 180]")
    e = Engine(device)
    check(load().air_build_fancurvedtomo(e.h, N, ptr(theta), len(theta), p, R, d, 0 if isMatrix else 1,
                                         DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix, {"kind": "fancurved", "N": N, "theta": theta, "p": p, "R": R, "d": d})
    return _finish(A, return_all, (theta, p, R, d))


def seismictomo(N, s=None, p=None, isDisp=0, isMatrix=1, *, dtype="f32", device=0, return_all=False):
    """[A,b,x,s,p] = seismictomo(N,s,p,isDisp,isMatrix), seismictomo.m:57-93
    (x is the tectonic phantom)."""
    N = int(N)
    if s is not None and np.size(s) > 1:
        raise ValueError("s must be a scaler")
    s = N if s is None else int(s)
    p = 2 * N if p is None else int(p)
    e = Engine(device)
    check(load().air_build_seismictomo(e.h, N, s, p, 0 if isMatrix else 1, DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix, {"kind": "seismic", "N": N, "s": s, "p": p})
    return _finish(A, return_all, (s, p), tectonic)


def sphericaltomo(N, angles=None, numCircles=None, isDisp=0, isMatrix=1, *, dtype="f32", device=0,
                  return_all=False):
    """[A,b,x,angles,numCircles] = sphericaltomo(N,angles,numCircles,isDisp,isMatrix),
    sphericaltomo.m:63-105.  isMatrix=0 gives the matrix-free operator, which
    re-forms every row's circle samples on each product (sphericaltomo.m:90-93)."""
    N = int(N)
    angles = np.arange(0.0, 359.0, 2.0) if angles is None or np.size(angles) == 0 else f64(angles)
    nC = int(_matlab_round(np.sqrt(2) * N)) if numCircles is None else int(numCircles)
    e = Engine(device)
    check(load().air_build_sphericaltomo(e.h, N, ptr(angles), len(angles), nC, 0 if isMatrix else 1,
                                         DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix, {"kind": "spherical", "N": N, "angles": angles, "numCircles": nC})
    return _finish(A, return_all, (angles, nC))


def seismicwavetomo(N, s=None, p=None, omega=None, isDisp=0, isMatrix=1, *, dtype="f32", device=0,
                    return_all=False):
    """[A,b,x,s,p,omega] = seismicwavetomo(N,s,p,omega,isDisp,isMatrix),
    seismicwavetomo.m:70-110 (x is the tectonic phantom).  isMatrix=0 gives the
    matrix-free operator, which re-evaluates the Fresnel kernels on every
    product (seismicwavetomo.m:94-99)."""
    N = int(N)
    s = N if s is None else int(s)
    p = 2 * N if p is None else int(p)
    omega = 10.0 if omega is None else float(omega)
    e = Engine(device)
    check(load().air_build_seismicwavetomo(e.h, N, s, p, omega, 0 if isMatrix else 1, DTYPES[dtype]))
    A = AirOperator(e, dtype, not isMatrix, {"kind": "seismicwave", "N": N, "s": s, "p": p, "omega": omega})
    return _finish(A, return_all, (s, p, omega), tectonic)
