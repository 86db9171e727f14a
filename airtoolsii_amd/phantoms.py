"""Test images of testprobs/phantomgallery.m (host-side input generator).

Images are N x N arrays whose column-major flattening ``im.ravel('F')`` is the
unknown vector x, as ``x = im(:)`` in MATLAB.  ``rng(seed)`` + ``rand`` is
MT19937 (init_genrand(seed), seed 0 meaning 5489) with 53-bit doubles, which
is what NumPy's legacy ``RandomState(seed).random_sample`` produces.
"""
from __future__ import annotations

import numpy as np


def _rng(seed):
    return np.random.RandomState(5489 if (seed is None or int(seed) == 0) else int(seed))


def _matlab_round(x):
    return np.where(np.asarray(x) < 0, -np.floor(-np.asarray(x) + 0.5), np.floor(np.asarray(x) + 0.5))


def shepplogan(N):
    """myphantom, phantomgallery.m:155-214 (modified Shepp-Logan)."""
    e = np.array([[1, .69, .92, 0, 0, 0], [-.8, .6624, .8740, 0, -.0184, 0], [-.2, .1100, .3100, .22, 0, -18],
                  [-.2, .1600, .4100, -.22, 0, 18], [.1, .2100, .2500, 0, .35, 0], [.1, .0460, .0460, 0, .1, 0],
                  [.1, .0460, .0460, 0, -.1, 0], [.1, .0460, .0230, -.08, -.605, 0],
                  [.1, .0230, .0230, 0, -.606, 0], [.1, .0230, .0460, .06, -.605, 0]])
    xn = (np.arange(N) - (N - 1) / 2.0) / ((N - 1) / 2.0)
    Xn = np.tile(xn, (N, 1))
    Yn = np.rot90(Xn)
    X = np.zeros((N, N))
    for A, a, b, x0, y0, ph in e:
        phi = ph * np.pi / 180.0
        x = Xn - x0
        y = Yn - y0
        u = x * np.cos(phi) + y * np.sin(phi)
        v = y * np.cos(phi) - x * np.sin(phi)
        X[(u * u) / (a * a) + (v * v) / (b * b) <= 1] += A
    X[X < 0] = 0
    return X


def smooth(N, p=4):
    """phantomgallery.m:216-232."""
    I, J = np.meshgrid(np.arange(1, N + 1), np.arange(1, N + 1))
    sigma = 0.25 * N
    c = np.array([[0.6 * N, 0.6 * N], [0.5 * N, 0.3 * N], [0.2 * N, 0.7 * N], [0.8 * N, 0.2 * N]])
    a = [1, 0.5, 0.7, 0.9]
    im = np.zeros((N, N))
    for i in range(p):
        im = im + a[i] * np.exp(-((I - c[i, 0]) ** 2) / (1.2 * sigma) ** 2 - ((J - c[i, 1]) ** 2) / sigma ** 2)
    return im / im.max()


def threephases(N, p=100, seed=None):
    """phantomgallery.m:257-292 (rand(p,2) fills column-major)."""
    rs = _rng(seed)
    # meshgrid(1:N): I(r, c) = c + 1 varies along columns, J(r, c) = r + 1 along rows
    ax = np.arange(1, N + 1, dtype=np.float64)
    I, J = ax[None, :], ax[:, None]
    cube = lambda t: t * t * t
    s1 = 0.025 * N
    c1 = rs.random_sample(2 * p).reshape((p, 2), order="F") * N
    d1, d2 = cube(2.5 * s1), cube(s1)
    # one blob at a time, exp(-f - g) on the whole grid, accumulated in blob order
    im1 = np.zeros((N, N))
    for i in range(p):
        im1 = im1 + np.exp(-cube(np.abs(I - c1[i, 0])) / d1 - cube(np.abs(J - c1[i, 1])) / d2)
    im1 = np.where(im1 < 0.35, 0.0, 2.0)
    s2 = 0.025 * N
    c2 = rs.random_sample(2 * p).reshape((p, 2), order="F") * N
    e1, e2 = (2 * s2) * (2 * s2), s2 * s2
    im2 = np.zeros((N, N))
    for i in range(p):
        u, v = I - c2[i, 0], J - c2[i, 1]
        im2 = im2 + np.exp(-(u * u) / e1 - (v * v) / e2)
    im2 = np.where(im2 < 0.55, 0.0, 1.0)
    im = im1 + im2
    im[im == 3] = 1
    return im / im.max()


def grains(N, numGrains=None, seed=None):
    """phantomgallery.m:352-389 (Voronoi cells, first centre wins ties)."""
    if numGrains is None:
        numGrains = int(_matlab_round(3 * np.sqrt(N)))
    rs = _rng(seed)
    dN = int(_matlab_round(N / 10))
    Nbig = N + 2 * dN
    xG = np.ceil(Nbig * rs.random_sample(numGrains))
    yG = np.ceil(Nbig * rs.random_sample(numGrains))
    # only the cropped window is needed: big-grid rows/cols dN..dN+N-1 (0-based)
    X, Y = np.meshgrid(np.arange(dN + 1, dN + N + 1, dtype=np.float64), np.arange(dN + 1, dN + N + 1, dtype=np.float64))
    best = np.zeros((N, N), np.int64)
    bd = np.full((N, N), np.inf)
    for k in range(numGrains):
        d = (X - xG[k]) ** 2 + (Y - yG[k]) ** 2
        upd = d < bd
        best[upd] = k
        bd[upd] = d[upd]
    im = best.astype(np.float64)
    return im / im.max()


def tectonic(N):
    """phantomgallery.m:422-454: two plates, the right one (0.75) tilted up at its
    left end, the left one (1) bending down past column 5*round(N/13)."""
    N = int(N)
    r = lambda v: int(_matlab_round(v))
    N5, N13, N7, N20 = r(N / 5), r(N / 13), r(N / 7), r(N / 20)
    if N5 < 1 or N13 < 1:
        raise ValueError("tectonic needs N >= 7")
    x = np.zeros((N, N))
    c0 = 5 * N13 - 1                                   # 0-based first column of the right plate
    x[N5 - 1:N5 + N7, c0:] = 0.75
    steps = (N20 + 1) // 2                             # odd j in 1..N20 lift the plate by one row
    if N5 - 1 - steps < 0 or 2 * N5 + (7 * N13 + 2) // 2 > N:
        raise ValueError("tectonic: plate outside the image for this N")
    for q in range(steps):
        j = 2 * q + 1
        x[N5 - 2 - q, c0 + j:] = 0.75
    x[N5 - 1:2 * N5, :5 * N13] = 1.0
    shift = 0
    for j in range(5 * N13, min(12 * N13, N) + 1):
        shift += j % 2
        x[N5 - 1 + shift:2 * N5 + shift, j - 1] = 1.0
    return x


def phantomgallery(name, N, P1=None, P2=None, P3=None):
    """phantomgallery.m:79-151 dispatch (shepplogan, smooth, threephases, grains, tectonic)."""
    if name == "shepplogan":
        return shepplogan(N)
    if name == "smooth":
        return smooth(N, 4 if P1 is None else int(P1))
    if name == "threephases":
        return threephases(N, 100 if P1 is None else int(P1), P2)
    if name == "grains":
        return grains(N, P1, P2)
    if name == "tectonic":
        return tectonic(N)
    if name in ("binary", "threephasessmooth", "fourphases", "ppower"):
        raise NotImplementedError(f"phantom '{name}' is not provided by this engine")
    raise ValueError("Illegal phantom name")


def add_noise(b, eta, seed=0):
    """Seeded Gaussian noise scaled to eta*||b|| (the reference demos use
    MATLAB randn, e.g. demos/demo_stoprules.m:43-46; randn streams are not
    reproduced, so the noise is drawn from NumPy's PCG64 with ``seed``)."""
    b = np.asarray(b, np.float64)
    e = np.random.default_rng(seed).standard_normal(b.shape[0])
    e = eta * np.linalg.norm(b) * e / np.linalg.norm(e)
    return b + e, e
