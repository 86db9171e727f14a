"""Test images of testprobs/phantomgallery.m (host-side input generator).

Images are N x N arrays whose column-major flattening ``im.ravel('F')`` is the
unknown vector x, as ``x = im(:)`` in MATLAB.  ``rng(seed)`` + ``rand`` is
MT19937 (init_genrand(seed), seed 0 meaning 5489) with 53-bit doubles, which
is what NumPy's legacy ``RandomState(seed).random_sample`` produces.
"""
from __future__ import annotations

import os

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


_TI_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "binary_training_image.txt")


def binary_training_image():
    """The 81 x 84 training image of phantomgallery.m:677-2641 (getTI)."""
    with open(_TI_PATH) as f:
        return np.array([[int(c) for c in line.strip()] for line in f if line.strip()], np.uint8)


def clique_histogram(Z):
    """CompHistTrain(Z, ones(3), 1), phantomgallery.m:458-509: counts of the
    3 x 3 patterns of Z, pattern id = 1 + sum_i n_i 2^(i-1) over the window
    read column-major; H[id - 1] is returned (512 values)."""
    Z = np.asarray(Z, np.int64)
    r, c = Z.shape
    ids = np.zeros((r - 2, c - 2), np.int64)
    for dc in range(3):
        for dr in range(3):
            ids += Z[dr:dr + r - 2, dc:dc + c - 2] << (dr + 3 * dc)
    return np.bincount(ids.ravel(), minlength=512).astype(np.float64)


def binary(N, seed=None):
    """binary, phantomgallery.m:236-253 with sequential_simulation_full_clique
    (:511-611): a 3 x 3 clique histogram of the training image and its mirror
    (+eps) drives a sequential simulation.  The first 3 x 3 pattern is drawn
    from the whole histogram, then the first three rows are extended column by
    column (three new pixels per draw), then each later row starts with three
    new pixels at its left end and continues one pixel at a time.  Every draw
    takes one rand: the patterns compatible with the known pixels (marg2id,
    :630-659, ids ascending) get pdf = H/sum(H), cdf = cumsum(pdf)/sum(pdf), and
    the first pattern with cdf > rand is taken.  Sums run left to right.

    A one-pixel draw has two candidates, ids v and v + 256 for the 8 known
    pixels v, so its outcome is 1 exactly when rand >= cdf(v), the first cdf
    entry; that value is tabulated once per v with the same operations."""
    N = int(N)
    if N < 3:
        raise ValueError("binary needs N >= 3")
    rnd = iter(_rng(seed).random_sample(1 + (N - 3) * (N - 1)).tolist())
    TI = binary_training_image()
    Hl = (clique_histogram(TI) + clique_histogram(TI[:, ::-1]) + np.finfo(float).eps).tolist()

    def seqsum(v):
        s = 0.0
        for t in v:
            s += t
        return s

    def pick(ids, hs, total):
        """First id whose cdf = cumsum(hs)/total exceeds the next rand."""
        r = next(rnd)
        c = 0.0
        for q, t in zip(ids, hs):
            c += t
            if c / total > r:
                return q
        raise ValueError("Zero probabiltiy pattern event! Please add a homogeneous distribution")

    def draw(ids):
        h = [Hl[q] for q in ids]
        sh = seqsum(h)
        pdf = [t / sh for t in h]
        return pick(ids, pdf, seqsum(pdf))

    def first_cdf(v):
        h0, h1 = Hl[v], Hl[v + 256]
        sh = h0 + h1
        p0, p1 = h0 / sh, h1 / sh
        return p0 / (p0 + p1)

    cut = [first_cdf(v) for v in range(256)]
    sim = [[-1] * N for _ in range(N)]

    def window(i0, j0):
        """0-based ids compatible with the known pixels of the window at (i0, j0)."""
        known = value = 0
        for dc in range(3):
            for dr in range(3):
                v = sim[i0 + dr][j0 + dc]
                if v >= 0:
                    bit = 1 << (dr + 3 * dc)
                    known |= bit
                    if v:
                        value |= bit
        return [q for q in range(512) if (q & known) == value]

    # the first pattern: cdf = cumsum(H)/sum(H) over all 512 ids
    q = pick(range(512), Hl, seqsum(Hl))
    for dc in range(3):
        for dr in range(3):
            sim[dr][dc] = (q >> (dr + 3 * dc)) & 1
    for j in range(3, N):                          # rows 1..3, columns 4..N: column j of the window is new
        q = draw(window(0, j - 2))
        for dr in range(3):
            sim[dr][j] = (q >> (dr + 6)) & 1
    for i in range(3, N):
        q = draw(window(i - 2, 0))                 # left end: row i of the window is new
        for dc in range(3):
            sim[i][dc] = (q >> (2 + 3 * dc)) & 1
        row, up1, up2 = sim[i], sim[i - 1], sim[i - 2]
        for j in range(3, N):                      # one new pixel, the window's bottom right
            v = (up2[j - 2] | (up1[j - 2] << 1) | (row[j - 2] << 2) | (up2[j - 1] << 3) | (up1[j - 1] << 4)
                 | (row[j - 1] << 5) | (up2[j] << 6) | (up1[j] << 7))
            row[j] = 0 if cut[v] > next(rnd) else 1
    return np.array(sim, np.float64)


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
    """phantomgallery.m:79-151 dispatch (shepplogan, smooth, binary, threephases, grains, tectonic)."""
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
    if name == "binary":
        return binary(N, P1)
    if name in ("threephasessmooth", "fourphases", "ppower"):
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
