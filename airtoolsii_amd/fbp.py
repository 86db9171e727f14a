"""fbp (auxil/fbp.m:1-80): filtered back projection through the engine.

The sinogram filter (zero padding, FFT, ramp with an optional window, inverse
FFT) runs as a batched FP64 hipFFT transform over the angles and the back
projection is the engine's A' product; only b and x cross the host boundary.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, f64, load, ptr
from .operators import as_operator

FILTERS = {"none": 0, "ram-lak": 1, "shepp-logan": 2, "cosine": 3, "hamming": 4, "hann": 5}


def fbp(A, b, theta, filter="ram-lak"):
    """x = fbp(A,b,theta,filter) with filter 'Ram-Lak' (default), 'Shepp-Logan',
    'cosine', 'Hamming', 'Hann' or 'none' (plain back projection A'*b)."""
    if A is None or b is None or theta is None:
        raise ValueError("Input A, b, theta must me specified")
    key = str(filter).lower()                              # fbp.m:34
    if key not in FILTERS:
        raise ValueError("Iinvalid filter")
    A = as_operator(A)
    b = f64(b)
    if b.shape[0] != A.shape[0]:
        raise ValueError("Incorrect length of input vector")
    na = int(np.size(theta))
    if key != "none" and (na == 0 or A.shape[0] % na):
        raise ValueError("fbp: size(A,1) is not a multiple of length(theta)")
    x = np.empty(A.shape[1])
    check(load().air_fbp(A.engine.h, ptr(b), max(na, 1), FILTERS[key], ptr(x)))
    return x
