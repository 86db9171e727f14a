"""FP64 CPU oracle for the AIR Tools II SIRT path -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``airtoolsii_amd``) never does.

It wraps ``build/libair_oracle.so`` (built from ``air_oracle.c`` by the
Makefile beside it) and restates the MATLAB host logic around the C loop:

* ``check_inputs``      auxil/check_inputs.m:59-259  (SIRT fields)
* ``calc_relaxpar``     auxil/calc_relaxpar.m:89-171 (rho via a converged
                        symmetric Lanczos from SciPy instead of eigs/ARPACK
                        under rng(0), which only a MATLAB run reproduces)
* ``sirt``              sirt/sirt.m:132-281
* ``train_relaxpar``    train/train_relaxpar.m:43-271 (SIRT branch)
* ``train_dpme``        train/train_dpme.m:1-137 (noise columns passed in)
* ``fbp``               auxil/fbp.m:31-80 (numpy complex FFT)
* ``art``, ``cart``     art/art.m:227-330, cart/cart.m:211-323 (sweeps in C)
* the test problems of testprobs/ (paralleltomo, fanlineartomo,
  fancurvedtomo, seismictomo, sphericaltomo, seismicwavetomo) and the
  phantoms of testprobs/phantomgallery.m, in C or numpy
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libair_oracle.so")
_lib = None

METHODS = {"landweber": 0, "cimmino": 1, "cav": 2, "drop": 3, "sart": 4}
STOPRULES = {"NONE": 0, "NCP": 1, "DP": 2, "ME": 3}

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


class _SirtOpts(C.Structure):
    _fields_ = [("stoprule_type", C.c_int), ("taudelta", C.c_double),
                ("res_dims", C.c_int * 2), ("ncp_smooth", C.c_int),
                ("lb_kind", C.c_int), ("lb", C.c_double), ("lb_vec", C.c_void_p),
                ("ub_kind", C.c_int), ("ub", C.c_double), ("ub_vec", C.c_void_p),
                ("line_search", C.c_int), ("lam", C.c_void_p), ("nthreads", C.c_int)]


class _SirtInfo(C.Structure):
    _fields_ = [("stoprule", C.c_int), ("finaliter", C.c_int),
                ("relaxpar", C.c_double), ("nsaved", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.ora_sind.restype = C.c_double; L.ora_sind.argtypes = [C.c_double]
        L.ora_cosd.restype = C.c_double; L.ora_cosd.argtypes = [C.c_double]
        L.ora_atand.restype = C.c_double; L.ora_atand.argtypes = [C.c_double]
        L.ora_round.restype = C.c_double; L.ora_round.argtypes = [C.c_double]
        L.ora_linspace.argtypes = [C.c_double, C.c_double, C.c_int64, _f64p]
        L.ora_calczeta.restype = C.c_double; L.ora_calczeta.argtypes = [C.c_int]
        L.ora_mt_seed.argtypes = [C.c_void_p, C.c_uint32]
        L.ora_mt_rand.restype = C.c_double; L.ora_mt_rand.argtypes = [C.c_void_p]
        L.ora_phantom_shepplogan.argtypes = [C.c_int, _f64p]
        L.ora_phantom_smooth.argtypes = [C.c_int, C.c_int, _f64p]
        L.ora_phantom_threephases.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, _f64p]
        L.ora_phantom_grains.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, _f64p]
        L.ora_clique_hist.argtypes = [C.c_void_p, C.c_int, C.c_int, _f64p]
        L.ora_phantom_binary.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_int, C.c_int, _f64p]
        L.ora_phantom_binary.restype = C.c_int64
        L.ora_paralleltomo.restype = C.c_void_p
        L.ora_paralleltomo.argtypes = [C.c_int, _f64p, C.c_int, C.c_int, C.c_double]
        L.ora_fanlineartomo.restype = C.c_void_p
        L.ora_fanlineartomo.argtypes = [C.c_int, _f64p, C.c_int, C.c_int,
                                        C.c_double, C.c_double, C.c_double]
        L.ora_fancurvedtomo.restype = C.c_void_p
        L.ora_fancurvedtomo.argtypes = [C.c_int, _f64p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.ora_seismictomo.restype = C.c_void_p
        L.ora_seismictomo.argtypes = [C.c_int, C.c_int, C.c_int]
        L.ora_sphericaltomo.restype = C.c_void_p
        L.ora_sphericaltomo.argtypes = [C.c_int, _f64p, C.c_int, C.c_int]
        L.ora_seismicwavetomo.restype = C.c_void_p
        L.ora_splitmix64.restype = C.c_uint64
        L.ora_splitmix64.argtypes = [C.c_uint64]
        L.ora_rng_u01.restype = C.c_double
        L.ora_rng_u01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64]
        L.ora_art_sweep.restype = None
        L.ora_art_sweep.argtypes = ([C.c_void_p] * 4 + [C.c_int64, C.c_void_p, C.c_double] + [C.c_void_p] * 3
                                    + [C.c_int64, C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_double, C.c_void_p])
        L.ora_cart_sweep.restype = None
        L.ora_cart_sweep.argtypes = ([C.c_void_p] * 4 + [C.c_int64] + [C.c_void_p] * 3 + [C.c_int64]
                                     + [C.c_void_p] * 2 + [C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_double,
                                                           C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int,
                                                           C.c_int, C.c_uint64])
        L.ora_seismicwavetomo.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        for f in ("ora_csr_m", "ora_csr_n", "ora_csr_nnz"):
            getattr(L, f).restype = C.c_int64
            getattr(L, f).argtypes = [C.c_void_p]
        L.ora_csr_error.restype = C.c_int; L.ora_csr_error.argtypes = [C.c_void_p]
        L.ora_csr_copy.argtypes = [C.c_void_p, _i64p, _i32p, _f64p]
        L.ora_csr_free.argtypes = [C.c_void_p]
        L.ora_weights.restype = C.c_int
        L.ora_weights.argtypes = [C.c_int, C.c_int64, C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p]
        L.ora_spmv.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.ora_sirt.restype = C.c_int
        L.ora_sirt.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p, _i64p, _i32p, _f64p,
                               C.c_void_p, C.c_void_p, _f64p, _f64p, _i32p, C.c_int,
                               C.POINTER(_SirtOpts), _f64p, _i32p, _f64p, C.POINTER(_SirtInfo)]
        _lib = L
    return _lib


# ---------------------------------------------------------------- built-ins
def sind(x):
    return lib().ora_sind(float(x))


def cosd(x):
    return lib().ora_cosd(float(x))


def linspace(d1, d2, n):
    out = np.empty(int(n), np.float64)
    lib().ora_linspace(float(d1), float(d2), int(n), out)
    return out


def calczeta(k):
    """calc_relaxpar.m:179-215 for a vector k (error for k(1) < 2, :202-204)."""
    k = np.atleast_1d(np.asarray(k, dtype=np.int64))
    if k.size == 0 or k[0] < 2:
        raise ValueError("The polynomial is not defined for k < 2")
    return np.array([lib().ora_calczeta(int(v)) for v in k])


def ncp_stat(rk, res_dims):
    """NCP statistic ncc of a residual, check_stoprules.m:96-133 (res_dims as in options.stoprule)."""
    rk = np.ascontiguousarray(np.ravel(rk), np.float64)
    rd = list(np.atleast_1d(res_dims))
    dims = (C.c_int * 2)(int(rd[0]) if len(rd) > 1 else rk.size, int(rd[1]) if len(rd) > 1 else 0)
    f = lib().ora_ncp_stat
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p]
    return f(rk.ctypes.data, rk.size, dims)


def mt_rand(seed, count):
    st = (C.c_uint32 * 625)()
    lib().ora_mt_seed(C.byref(st), int(seed))
    return np.array([lib().ora_mt_rand(C.byref(st)) for _ in range(int(count))])


# ---------------------------------------------------------------- phantoms
_TI_TXT = os.path.join(_HERE, "..", "airtoolsii_amd", "data", "binary_training_image.txt")


def binary_training_image():
    """getTI, phantomgallery.m:677-2641, read from the data file (one image row
    per line of '0'/'1' characters) as an 81 x 84 uint8 array."""
    rows = [line.strip() for line in open(_TI_TXT) if line.strip()]
    return (np.frombuffer("".join(rows).encode(), np.uint8) - ord("0")).reshape(len(rows), len(rows[0]))


def clique_histogram(Z):
    """CompHistTrain(Z, ones(3), 1), phantomgallery.m:458-509, in C: H[id-1]."""
    Z = np.asfortranarray(np.asarray(Z, np.uint8))
    H = np.zeros(512)
    lib().ora_clique_hist(Z.ctypes.data, Z.shape[0], Z.shape[1], H)
    return H


def binary_draws(N, seed):
    """Number of rand draws binary(N, seed) takes (the C loop's count)."""
    ti = np.asfortranarray(binary_training_image())
    out = np.zeros(N * N)
    return int(lib().ora_phantom_binary(N, 1, int(seed), ti.ctypes.data, ti.shape[0], ti.shape[1], out))


def phantomgallery(name, N, P1=None, P2=None):
    """phantomgallery.m:79-151 for shepplogan/smooth/threephases/grains/binary.
    Returns the N x N image (column-major semantics: use ``im.ravel('F')``)."""
    L = lib()
    out = np.zeros(N * N, np.float64)
    if name == "shepplogan":
        L.ora_phantom_shepplogan(N, out)
    elif name == "smooth":
        L.ora_phantom_smooth(N, 4 if P1 is None else int(P1), out)
    elif name == "threephases":
        L.ora_phantom_threephases(N, 100 if P1 is None else int(P1),
                                  0 if P2 is None else 1, 0 if P2 is None else int(P2), out)
    elif name == "grains":
        L.ora_phantom_grains(N, 0 if P1 is None else int(P1),
                             0 if P2 is None else 1, 0 if P2 is None else int(P2), out)
    elif name == "binary":
        ti = np.asfortranarray(binary_training_image())
        nd = L.ora_phantom_binary(N, 0 if P1 is None else 1, 0 if P1 is None else int(P1),
                                  ti.ctypes.data, ti.shape[0], ti.shape[1], out)
        if nd < 0:
            raise ValueError("binary needs N >= 3")
    else:
        raise ValueError("Illegal phantom name")
    return out.reshape((N, N), order="F")


# ---------------------------------------------------------------- tracers
def _take_csr(h):
    L = lib()
    if not h:
        raise MemoryError("oracle tracer failed")
    try:
        if L.ora_csr_error(h):
            raise IndexError("Index exceeds matrix dimensions (column outside 1..N^2)")
        m, n, nnz = L.ora_csr_m(h), L.ora_csr_n(h), L.ora_csr_nnz(h)
        rp = np.empty(m + 1, np.int64); ci = np.empty(nnz, np.int32); v = np.empty(nnz, np.float64)
        L.ora_csr_copy(h, rp, ci, v)
    finally:
        L.ora_csr_free(h)
    return sp.csr_matrix((v, ci, rp), shape=(m, n))


def paralleltomo(N, theta=None, p=None, d=None):
    """testprobs/paralleltomo.m:62-89 defaults + :112-301 matrix build."""
    theta = np.arange(0.0, 180.0) if theta is None else np.atleast_1d(np.asarray(theta, np.float64))
    p = int(lib().ora_round(np.sqrt(2) * N)) if p is None else int(p)
    d = float(p - 1) if d is None else float(d)
    return _take_csr(lib().ora_paralleltomo(int(N), np.ascontiguousarray(theta), len(theta), p, d))


def fanlineartomo(N, theta=None, p=None, R=2.0, dw=2.5, sd=3.0):
    """testprobs/fanlineartomo.m:70-114 defaults + :141-358 matrix build."""
    theta = np.arange(0.0, 359.0, 2.0) if theta is None else np.atleast_1d(np.asarray(theta, np.float64))
    p = int(lib().ora_round(np.sqrt(2) * N)) if p is None else int(p)
    Rn, dwn, sdn = R * N, dw * N, sd * N
    if Rn < np.sqrt(2) / 2 * N:
        raise ValueError("R must be greater than half squareroot 2")
    return _take_csr(lib().ora_fanlineartomo(int(N), np.ascontiguousarray(theta), len(theta), p,
                                             float(Rn), float(dwn), float(sdn)))


def fancurvedtomo(N, theta=None, p=None, R=2.0, d=None):
    """testprobs/fancurvedtomo.m:69-113 defaults + :140-339 matrix build."""
    theta = np.arange(0.0, 359.0, 2.0) if theta is None else np.atleast_1d(np.asarray(theta, np.float64))
    p = int(lib().ora_round(np.sqrt(2) * N)) if p is None else int(p)
    if R < math.sqrt(2) / 2:
        raise ValueError("R must be greater than half squareroot 2")
    if d is None:
        d = 2.0 * ((180.0 / math.pi) * math.atan(1.0 / (2.0 * R - 1.0)))
    if d < 0 or d > 180:
        raise ValueError("The angle of the source must be in the interval [0 180]")
    return _take_csr(lib().ora_fancurvedtomo(int(N), np.ascontiguousarray(theta), len(theta), p,
                                             float(R * N), float(d)))


def seismictomo(N, s=None, p=None):
    """testprobs/seismictomo.m:65-74 defaults + :96-287 matrix build."""
    s = int(N) if s is None else int(s)
    p = 2 * int(N) if p is None else int(p)
    return _take_csr(lib().ora_seismictomo(int(N), s, p))


def sphericaltomo(N, angles=None, numCircles=None):
    """testprobs/sphericaltomo.m:63-75 defaults + :108-270 matrix build."""
    angles = np.arange(0.0, 359.0, 2.0) if angles is None else np.atleast_1d(np.asarray(angles, np.float64))
    nC = int(lib().ora_round(np.sqrt(2) * N)) if numCircles is None else int(numCircles)
    return _take_csr(lib().ora_sphericaltomo(int(N), np.ascontiguousarray(angles), len(angles), nC))


def seismicwavetomo(N, s=None, p=None, omega=10.0):
    """testprobs/seismicwavetomo.m:70-83 defaults + :114-288 matrix build."""
    s = int(N) if s is None else int(s)
    p = 2 * int(N) if p is None else int(p)
    return _take_csr(lib().ora_seismicwavetomo(int(N), s, p, float(omega)))


def tectonic(N):
    """phantomgallery.m:422-454, element by element in 1-based indices."""
    rnd = lambda v: int(math.floor(v + 0.5))
    N5, N13, N7, N20 = rnd(N / 5), rnd(N / 13), rnd(N / 7), rnd(N / 20)
    x = np.zeros((N, N))

    def setv(i, j, v):                      # x(i,j) = v with 1-based (i,j)
        if not (1 <= i <= N and 1 <= j <= N):
            raise IndexError("tectonic index outside the image")
        x[i - 1, j - 1] = v

    for i in range(N5, N5 + N7 + 1):
        for j in range(5 * N13, N + 1):
            setv(i, j, 0.75)
    i = N5
    for j in range(1, N20 + 1):
        if j % 2 != 0:
            i -= 1
            for jj in range(5 * N13 + j, N + 1):
                setv(i, jj, 0.75)
    for i in range(N5, 2 * N5 + 1):
        for j in range(1, 5 * N13 + 1):
            setv(i, j, 1.0)
    vec = list(range(N5, 2 * N5 + 1))
    for j in range(5 * N13, min(12 * N13, N) + 1):
        if j % 2 != 0:
            vec = [v + 1 for v in vec]
        for v in vec:
            setv(v, j, 1.0)
    return x


# ---------------------------------------------------------------- weights
def _csr_parts(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int32),
            np.ascontiguousarray(A.data, np.float64))


def weights(method, A):
    """get_mfun_dfun.m matrix branch; returns (M, D) diagonals."""
    m, n = A.shape
    rp, ci, v = _csr_parts(A)
    M = np.empty(m); D = np.empty(n)
    if lib().ora_weights(METHODS[method], m, n, rp, ci, v, M, D) != 0:
        raise ValueError("SIRT method not defined.")
    return M, D


def weights_handle(method, afun, m, n):
    """get_mfun_dfun.m's function-handle branches (the handle called on unit
    vectors and on ones, as the reference does): Cimmino :68-73, CAV :108-121,
    DROP :150-155 and :179-184, SART :208 and :229.  Returns (M, D)."""
    M, D = np.ones(m), np.ones(n)
    if method == "landweber":
        return M, D
    with np.errstate(divide="ignore", invalid="ignore"):
        if method in ("cimmino", "drop"):
            normAi = np.zeros(m)
            for i in range(m):
                e = np.zeros(m); e[i] = 1
                normAi[i] = np.linalg.norm(afun(e, "transp")) ** 2
            M = (1.0 / m) * (1.0 / normAi) if method == "cimmino" else 1.0 / normAi
            M[M == np.inf] = 0
        if method in ("cav", "drop"):
            s = np.zeros(n)
            for j in range(n):
                e = np.zeros(n); e[j] = 1
                s[j] = 1.0 / np.sum(afun(e, "notransp") != 0)
            if method == "drop":
                s[s == np.inf] = 0
                D = s
            else:
                normAs = np.zeros(m)
                for i in range(m):
                    e = np.zeros(m); e[i] = 1
                    normAs[i] = s @ (afun(e, "transp") ** 2)
                M = 1.0 / normAs
                M[M == np.inf] = 0
        if method == "sart":
            M = 1.0 / np.abs(afun(np.ones(n), "notransp"))
            M[M == np.inf] = 0
            D = 1.0 / np.abs(afun(np.ones(m), "transp"))
            D[D == np.inf] = 0
    return M, D


def spmv(A, x, nthreads=0):
    rp, ci, v = _csr_parts(A)
    y = np.empty(A.shape[0])
    lib().ora_spmv(A.shape[0], rp, ci, v, np.ascontiguousarray(x, np.float64), y, nthreads)
    return y


def spectral_radius(A, M, D):
    """Largest eigenvalue of D A' M A (calc_relaxpar.m:91-103), computed as the
    top eigenvalue of the symmetric D^(1/2) A' M A D^(1/2) with SciPy's Lanczos
    to tolerance 1e-12 (a library routine, not reference code)."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    A = sp.csr_matrix(A)
    AT = A.T.tocsr()
    sD = np.sqrt(D)

    def mv(v):
        v = np.ravel(v)
        return sD * (AT @ (M * (A @ (sD * v))))

    op = LinearOperator((A.shape[1], A.shape[1]), matvec=mv, dtype=np.float64)
    vals = eigsh(op, k=1, which="LA", tol=1e-12, v0=np.ones(A.shape[1]), return_eigenvectors=False)
    return float(vals[0])


# ---------------------------------------------------------------- SIRT
def relaxpar_vector(relaxpar, rho, kmax):
    """calc_relaxpar.m:105-171 -> (casel, per-iteration lambda vector or None, relaxpar)."""
    if relaxpar is None:
        lam = 1.9 / rho
        return 1, np.full(kmax, lam), lam
    if not isinstance(relaxpar, str):
        arr = np.ravel(np.asarray(relaxpar, np.float64))
        if arr.size > 1:        # a per-iteration sequence, applied as relaxpar(k) like the psi vectors (:160-164)
            return 3, arr[:kmax], arr
        lam = float(arr[0])
        return 1, np.full(kmax, lam), lam
    s = relaxpar.lower()
    if s.startswith("line"):
        return 2, None, None
    if s.startswith("psi1"):
        z = calczeta(np.arange(2, kmax))
        nu = 2.0 if s.startswith("psi1mod") else 1.0
        v = np.concatenate([[np.sqrt(2), np.sqrt(2)], nu * 2 * (1 - z)]) / rho
        return 3, v[:kmax], v
    if s.startswith("psi2"):
        kk = np.arange(2, kmax)
        z = calczeta(kk)
        nu = 1.5 if s.startswith("psi2mod") else 1.0
        v = np.concatenate([[np.sqrt(2), np.sqrt(2)], nu * 2 * (1 - z) / ((1 - z ** kk) ** 2)]) / rho
        return 3, v[:kmax], v
    raise ValueError("The chosen relaxation strategy is not valid for this method.")


def sirt(method, A, b, K, x0=None, options=None, nthreads=0, return_resnorm=False):
    """sirt/sirt.m:132-281 on a SciPy sparse A (FP64).  ``method`` is a name
    or a dict with optional 'M'/'D' vectors (custom method)."""
    options = dict(options or {})
    A = sp.csr_matrix(A, dtype=np.float64)
    m, n = A.shape
    b = np.ascontiguousarray(np.ravel(b), np.float64)
    if b.shape[0] != m:
        raise ValueError("The size of A and b do not match.")
    K = np.atleast_1d(np.asarray(K, np.int64))
    if K.size == 0:
        raise ValueError("Max no. of iterations must ALWAYS be specified.")
    if not np.all(np.diff(K) > 0):
        raise ValueError("Elements of K must be strictly increasing.")
    kmax = int(K[-1])
    x0 = np.zeros(n) if x0 is None else np.ascontiguousarray(np.ravel(x0), np.float64)
    if x0.shape[0] != n:
        raise ValueError("The size of x0 does not match the problem.")
    if method is None:
        method = "sart"
    if isinstance(method, str):
        M, D = weights(method, A)
    else:
        M = np.ascontiguousarray(method.get("M"), np.float64) if method.get("M") is not None else np.ones(m)
        D = np.ascontiguousarray(method.get("D"), np.float64) if method.get("D") is not None else np.ones(n)
    rho = options.get("rho", np.nan)
    if isinstance(method, str) and method == "sart":
        rho = 1.0
    rp_in = options.get("relaxpar")
    needs_rho = not (isinstance(rp_in, str) and rp_in.lower().startswith("line"))
    if np.isnan(rho) and needs_rho:
        rho = spectral_radius(A, M, D)
    casel, lam, relaxpar = relaxpar_vector(rp_in, rho, kmax)
    st = options.get("stoprule", {}) or {}
    stype = str(st.get("type", "none")).upper()
    o = _SirtOpts()
    o.stoprule_type = STOPRULES[stype]
    o.taudelta = float(st.get("taudelta", np.nan))
    rd = st.get("res_dims", None)
    if rd is not None:
        rd = list(np.atleast_1d(rd))
        o.res_dims[0] = int(rd[0]); o.res_dims[1] = int(rd[1]) if len(rd) > 1 else 0
    o.ncp_smooth = int(st.get("ncp_smooth", 2))
    keep = []
    for nm, kind, val, vec in (("lbound", "lb_kind", "lb", "lb_vec"), ("ubound", "ub_kind", "ub", "ub_vec")):
        if nm in options and options[nm] is not None:
            bv = np.ravel(np.asarray(options[nm], np.float64))
            if bv.size == 1:
                setattr(o, kind, 1); setattr(o, val, float(bv[0]))
            elif bv.size == n:
                arr = np.ascontiguousarray(bv); keep.append(arr)
                setattr(o, kind, 2); setattr(o, vec, arr.ctypes.data)
            else:
                raise ValueError(f"{nm} must either be scalar or same length as vector of unknowns x.")
    o.line_search = 1 if casel == 2 else 0
    if lam is not None:
        lam = np.ascontiguousarray(np.resize(lam, kmax), np.float64); keep.append(lam)
        o.lam = lam.ctypes.data
    o.nthreads = int(nthreads)
    rp, ci, v = _csr_parts(A)
    AT = A.T.tocsr(); AT.sort_indices()
    cp, ri, vc = (np.ascontiguousarray(AT.indptr, np.int64), np.ascontiguousarray(AT.indices, np.int32),
                  np.ascontiguousarray(AT.data, np.float64))
    Mc = np.ascontiguousarray(M); Dc = np.ascontiguousarray(D)
    X = np.zeros(n * len(K)); its = np.zeros(len(K) + 1, np.int32); rn = np.zeros(kmax + 1)
    info = _SirtInfo()
    lib().ora_sirt(m, n, rp, ci, v, cp, ri, vc, Mc.ctypes.data, Dc.ctypes.data, b, x0,
                   np.ascontiguousarray(K, np.int32), len(K), C.byref(o), X, its, rn, C.byref(info))
    X = X.reshape((len(K), n)).T[:, :info.nsaved]
    nsv = int(np.sum(K < info.finaliter)) + 1
    out_info = {"stoprule": info.stoprule, "finaliter": info.finaliter,
                "relaxpar": relaxpar if info.finaliter == 0 or casel != 1 else float(lam[0]),
                "rho": rho, "itersaved": its[:nsv].astype(np.int64)}
    if casel == 2:
        out_info["relaxpar"] = info.relaxpar
    if return_resnorm:
        return X, out_info, {"M": M, "D": D}, rn[:info.finaliter + 1]
    return X, out_info, {"M": M, "D": D}


def train_dpme(A, b_exact, x_exact, method, type, delta, s, kmax, options=None, noise=None, run=None):
    """train/train_dpme.m:1-137 with the unscaled noise columns given as
    ``noise`` (m x s; the reference draws them with randn).  ``run(A,b,K,x0,
    opts)`` defaults to this oracle's ``sirt``; ``resid(x, b)`` is b - A x."""
    if run is None:
        def run(A_, b_, K_, x0_, opts_):
            return _run_method(method, A_, b_, K_, x0_, opts_)
    if not str(type)[:2].upper() in ("DP", "ME"):
        raise ValueError("The chosen type is invalid")
    opts = {} if options is None else dict(options)
    opts["stoprule"] = {"type": "none"}
    m = A.shape[0]
    e = np.array(noise, np.float64).reshape(m, s)
    e = delta * e / (np.ones((m, 1)) @ np.sqrt(np.sum(e * e, axis=0))[None, :])
    b = b_exact[:, None] @ np.ones((1, s)) + e
    tauh = np.zeros(s)
    kstep = kmax if kmax < 10 else 10
    for i in range(s):
        bi = b[:, i]
        X = np.zeros((x_exact.shape[0], 1))
        x0 = np.zeros(x_exact.shape[0])
        k = kstep
        K = np.arange(1, k + 1)
        Ek = np.zeros(0)
        while True:
            Xny, _ = run(A, bi, K, x0, opts)
            xd = Xny - x_exact[:, None]
            Ek = np.concatenate([Ek, np.sqrt(np.sum(xd * xd, axis=0))])
            X = np.concatenate([X, Xny], axis=1)
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
        rk = bi - A @ X[:, kopt - 1]
        if str(type)[:2].upper() == "DP":
            tauk = np.linalg.norm(rk) / delta
        else:
            if kopt - 2 < 0:
                raise IndexError("Index in position 2 is invalid")
            rkm1 = bi - A @ X[:, kopt - 2]
            tauk = 0.5 * (rkm1 @ (rkm1 + rk)) / (delta * np.linalg.norm(rkm1))
        tauh[i] = tauk
    return float(np.mean(tauh))


# ---------------------------------------------------------------- ART / CART
def _bounds(opts, n):
    out = []
    for key in ("lbound", "ubound"):
        v = opts.get(key)
        if v is None or np.size(v) == 0:
            out += [0, 0.0, None]
        elif np.size(v) == 1:
            out += [1, float(np.ravel(v)[0]), None]
        else:
            out += [2, 0.0, np.ascontiguousarray(np.ravel(v), np.float64)]
    return out


def _nanmax(v):
    """MATLAB's max: NaN entries are ignored unless every entry is NaN."""
    v = np.asarray(v, np.float64)
    return np.nan if np.all(np.isnan(v)) else np.nanmax(v)


def _stop(opts, k, kmax, r, rprev, dk=None):
    """check_stoprules.m:53-161 for NCP, DP, ME, NONE -> code or None.  ``dk``
    is the NCP filter window (a list, updated in place), Inf at the start
    (check_inputs.m:207-209)."""
    st = opts.get("stoprule") or {}
    t = str(st.get("type", "none")).upper()
    nrk = np.linalg.norm(r)
    if t == "NCP":
        ncc = ncp_stat(r, st["res_dims"])
        if _nanmax(dk) < _nanmax(list(dk[1:]) + [ncc]) or k >= kmax:
            return 1 if k != kmax else 0
        dk[:] = list(dk[1:]) + [ncc]
        return None
    if t == "DP":
        if nrk <= st["taudelta"] or k >= kmax:
            return 2 if k != kmax else 0
        return None
    if t == "ME":
        if rprev is not None and (rprev @ (rprev + r) / 2) / np.linalg.norm(rprev) <= st["taudelta"]:
            return 3
        if k >= kmax:
            return 0
        return None
    return 0 if k >= kmax else None


def _ptr(a):
    return None if a is None else a.ctypes.data


def art(art_method, A, b, K, x0=None, options=None):
    """art/art.m:1-330 with the inner row loop in C (ora_art_sweep).  The
    random choices of randkaczmarz use the SplitMix64 stream (options['seed'])."""
    opts = dict(options or {})
    A = sp.csr_matrix(A, dtype=np.float64)
    A.sort_indices()
    m, n = A.shape
    K = np.atleast_1d(np.asarray(K, np.int64))
    method = "kaczmarz" if art_method is None else art_method
    sym = isinstance(method, str) and method.lower().startswith("sym")
    if sym:
        if np.any(K % 2):
            raise ValueError("For symkaczmarz only even iteration numbers can be requested.")
        K = K // 2
    kmax = int(K[-1])
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    nrm = np.asarray(A.multiply(A).sum(axis=1)).ravel()
    if isinstance(method, str):
        I = np.nonzero(nrm > 0)[0]
        if method.lower() == "symkaczmarz":
            I = np.concatenate([I, I[::-1]])
        elif method.lower() not in ("kaczmarz", "randkaczmarz"):
            raise ValueError("Unknown ART method specified")
    else:
        I = np.ravel(np.asarray(method, np.int64)) - 1
        I = I[nrm[I] > 0]
    nrm = nrm + float(opts.get("damp", 0.0)) * nrm.max()
    rnd = isinstance(method, str) and method.lower() == "randkaczmarz"
    seed = int(opts.get("seed", 0))
    L = lib()
    if rnd:
        cumul = np.cumsum(nrm / nrm.sum())
        fast = np.max(np.abs(cumul - np.arange(1, m + 1) / m)) < 0.05
    rp = opts.get("relaxpar")
    lam = 1.0 if rp is None else (float(rp) if not callable(rp) else 0.0)
    lbk, lb, lbv, ubk, ub, ubv = _bounds(opts, n)
    rk = b - A @ x
    dk = [np.inf] * int((opts.get("stoprule") or {}).get("ncp_smooth", 2))
    code = _stop(opts, 0, kmax, rk, None, dk)
    rprev = rk if code is None else None
    X, k = [], 0
    ip, ic, iv = (np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int32),
                  np.ascontiguousarray(A.data))
    li = 0
    while code is None:
        k += 1
        if rnd:
            perm = I.copy()
            for t in range(len(perm) - 1, 0, -1):
                j = int(L.ora_rng_u01(seed, 1, k, t) * (t + 1))
                perm[t], perm[j] = perm[j], perm[t]
            if fast:
                steps = perm
            else:
                u = np.array([L.ora_rng_u01(seed, 2, k, t) for t in range(len(perm))])
                steps = np.minimum(np.array([np.sum(cumul < uu) for uu in u]), m - 1)
            ivals = perm
        else:
            steps = ivals = I
        lamv = None
        if callable(rp):
            lamv = np.ascontiguousarray([float(rp((k - 1) * m + int(i) + 1)) for i in ivals], np.float64)
        st32 = np.ascontiguousarray(steps, np.int32)
        L.ora_art_sweep(ip.ctypes.data, ic.ctypes.data, iv.ctypes.data, st32.ctypes.data, len(st32), _ptr(lamv), lam,
                        b.ctypes.data, nrm.ctypes.data, x.ctypes.data, n, lbk, lb, _ptr(lbv), ubk, ub, _ptr(ubv))
        rk = b - A @ x
        code = _stop(opts, k, kmax, rk, rprev, dk)
        if code is None and str((opts.get("stoprule") or {}).get("type", "none")).upper() == "ME":
            rprev = rk
        if (li < len(K) and k == K[li]) or code is not None:
            X.append(x.copy())
            li += 1
    fin = 2 * k if sym else k
    return np.array(X).T, {"stoprule": code, "finaliter": fin}


def cart(cart_method, A, b, K, x0=None, options=None):
    """cart/cart.m:1-323 with the inner column loop in C (ora_cart_sweep)."""
    opts = dict(options or {})
    A = sp.csc_matrix(A, dtype=np.float64)
    A.sort_indices()
    m, n = A.shape
    K = np.atleast_1d(np.asarray(K, np.int64))
    kmax = int(K[-1])
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    nrm = np.asarray(A.multiply(A).sum(axis=0)).ravel()
    if cart_method is None or (isinstance(cart_method, str) and cart_method.lower() == "columnaction"):
        J = np.nonzero(nrm > 0)[0]
    elif isinstance(cart_method, str):
        raise ValueError("Unknown CART method specified")
    else:
        J = np.ravel(np.asarray(cart_method, np.int64)) - 1
        J = J[nrm[J] > 0]
    nrm = nrm + float(opts.get("damp", 0.0)) * nrm.max()
    lam = float(opts.get("relaxpar", 0.25))
    thr = float(opts.get("THR", 1e-4))
    kbegin = int(opts.get("Kbegin", 10))
    nunflag = int(opts.get("Nunflag", 0) or 0) or max(1, int(math.floor(kmax / 4 + 0.5)))
    seed = int(opts.get("seed", 0))
    lbk, lb, lbv, ubk, ub, ubv = _bounds(opts, n)
    flag = np.ones(n, np.uint8)
    nflag = np.zeros(n, np.int32)
    r = b - A @ x
    dk = [np.inf] * int((opts.get("stoprule") or {}).get("ncp_smooth", 2))
    code = _stop(opts, 0, kmax, r, None, dk)
    rprev = r.copy() if code is None else None
    cp, ri, cv = (np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int32),
                  np.ascontiguousarray(A.data))
    J32 = np.ascontiguousarray(J, np.int32)
    X, k, li = [], 0, 0
    while code is None:
        k += 1
        lib().ora_cart_sweep(cp.ctypes.data, ri.ctypes.data, cv.ctypes.data, J32.ctypes.data, len(J32),
                             nrm.ctypes.data, x.ctypes.data, r.ctypes.data, n, flag.ctypes.data, nflag.ctypes.data,
                             lbk, lb, _ptr(lbv), ubk, ub, _ptr(ubv), lam, thr, k, kbegin, nunflag, seed)
        code = _stop(opts, k, kmax, r, rprev, dk)
        if code is None and str((opts.get("stoprule") or {}).get("type", "none")).upper() == "ME":
            rprev = r.copy()
        if (li < len(K) and k == K[li]) or code is not None:
            X.append(x.copy())
            li += 1
    return np.array(X).T, {"stoprule": code, "finaliter": k, "flags": flag.copy()}


def fbp(A, b, theta, filter="ram-lak"):
    """auxil/fbp.m:31-80 with numpy's complex FFT over the full symmetric filter."""
    filter = filter.lower()
    A = sp.csr_matrix(A)
    if filter == "none":
        return A.T @ b
    m = A.shape[0]
    na = len(np.atleast_1d(theta))
    nr = m // na
    if nr * na != m:
        raise ValueError("reshape: size(A,1) is not a multiple of length(theta)")
    B = np.reshape(np.asarray(b, np.float64), (nr, na), order="F")
    order = max(64, 2 ** int(math.ceil(math.log2(2 * nr))))
    n = np.arange(order // 2 + 1, dtype=np.float64)
    f = n / n.max()
    f[0] = 0.4 / order
    f[-1] = 1 - f[0]
    w = 2 * np.pi * n / order
    if filter == "shepp-logan":
        t = w[1:] / (2 * np.pi)
        f[1:] = f[1:] * (np.sin(np.pi * t) / (np.pi * t))
    elif filter == "cosine":
        f[1:] = f[1:] * np.cos(w[1:] / 2)
    elif filter == "hamming":
        f[1:] = f[1:] * (.54 + .46 * np.cos(w[1:]))
    elif filter == "hann":
        f[1:] = f[1:] * (1 + np.cos(w[1:])) / 2
    elif filter != "ram-lak":
        raise ValueError("Iinvalid filter")
    f = np.concatenate([f, f[-2:0:-1]])
    P = np.zeros((len(f), na))
    P[:nr] = B
    F = np.fft.fft(P, axis=0) * f[:, None]
    P = np.real(np.fft.ifft(F, axis=0))[:nr]
    return A.T @ P.ravel(order="F") * np.pi / (2 * na)


ROW_ACTION = ("kaczmarz", "symkaczmarz", "randkaczmarz", "columnaction")


def _run_method(method, A, b, K, x0, opts):
    """[X, info] of the named SIRT, ART or CART method of this oracle."""
    if method == "columnaction":
        X, info = cart(method, A, b, K, x0, opts)
        return X, info
    if method in ROW_ACTION:
        return art(method, A, b, K, x0, opts)
    X, info, _ = sirt(method, A, b, K, x0, opts)
    return X, info


def train_relaxpar(A, b, x_ex, method, kmax, options=None, run=None):
    """train/train_relaxpar.m:43-271.  SIRT methods start from the default
    relaxpar with relaxparmax = 2/rho; the ART methods and columnaction
    (:90-96) start from relaxpar 0.25 with relaxparmax = 2, and their golden
    section search tests the upper probe first (:213-268).  ``run(A,b,K,x0,
    opts)`` defaults to this oracle's sirt/art/cart with the named method."""
    is_sirt = method not in ROW_ACTION
    if run is None:
        def run(A_, b_, K_, x0_, opts_):
            return _run_method(method, A_, b_, K_, x0_, opts_)
    opts = {} if options is None else dict(options)
    if options is not None:
        opts["stoprule"] = {"type": "none"}
        opts.pop("relaxpar", None)
    fudge = 0.02
    stepk = 10
    K = np.arange(1, stepk + 1)
    k = 0
    Er = []
    x0 = None
    while True:
        if k + stepk > kmax:
            K = np.arange(1, kmax - k + 1)
            k = kmax
        else:
            k = k + stepk
        if k <= stepk and not is_sirt:
            opts["relaxpar"] = 0.25
        Xnew, info = run(A, b, K, x0, opts)
        if k <= stepk:
            if is_sirt:
                opts["rho"] = info["rho"]
                relaxparmax = 2.0 / info["rho"]
            else:
                relaxparmax = 2.0
        Er.extend(np.sqrt(np.sum((Xnew - x_ex[:, None]) ** 2, axis=0)).tolist())
        kopt = int(np.argmin(Er)) + 1
        minE = Er[kopt - 1]
        if kopt == k and k < kmax:
            x0 = Xnew[:, -1]
        else:
            break
    meub = (1 + fudge) * minE

    def probe(lam):
        o2 = dict(opts); o2["relaxpar"] = lam
        X, _ = run(A, b, np.arange(1, kopt + 1), None, o2)
        ee = np.sqrt(np.sum((X - x_ex[:, None]) ** 2, axis=0))
        kk = int(np.sum((ee[:-1] > meub) & (np.diff(ee) < 0))) + 1
        return kk, ee[kk - 1]

    r = (3 - np.sqrt(5)) / 2
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
    return (alphap + betap) / 2
