"""SIRT with many right-hand sides on one operator (SpMM passes) against one
run per right-hand side and the FP64 oracle.

Two product paths are covered: the LDS-staged chunk kernels on the tiled
layout (default) and the compressed-row/column kernels (AIR_SPMM=0, also the
path of operators without a tiled layout)."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu
TOL = 1e-5
PATHS = ["chunks", "rows"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def use_path(monkeypatch, path):
    if path == "rows":
        monkeypatch.setenv("AIR_SPMM", "0")
    else:
        monkeypatch.delenv("AIR_SPMM", raising=False)


@pytest.fixture(scope="module")
def prob(oracle):
    N, theta, p = 40, np.arange(0.0, 180.0, 3.0), 57
    A = air.paralleltomo(N, theta, p)
    Ao = oracle.paralleltomo(N, theta, p)
    rng = np.random.default_rng(2)
    xs = [air.phantomgallery("grains", N, None, s).ravel(order="F") for s in range(1, 65)]
    B = np.stack([Ao @ x for x in xs], 1)
    B = B + 0.02 * np.linalg.norm(B, axis=0) * rng.standard_normal(B.shape) / np.sqrt(B.shape[0])
    return A, Ao, B


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("method,nrhs", [("cimmino", 64), ("sart", 5), ("drop", 1), ("landweber", 17)])
def test_multi_equals_single_runs(prob, oracle, monkeypatch, path, method, nrhs):
    use_path(monkeypatch, path)
    A, Ao, B = prob
    opts = {"lbound": 0, "ubound": 1}
    X, info = air.sirt_multi(method, A, B[:, :nrhs], 30, None, opts)
    assert X.shape == (A.shape[1], nrhs)
    for j in sorted({0, nrhs // 2, nrhs - 1}):
        Xs, _, _ = air.sirt(method, A, B[:, j], 30, None, opts)
        assert rel(X[:, j], Xs[:, -1]) <= TOL, (method, j)
        rho = 1.0 if method == "sart" else oracle.spectral_radius(Ao, *oracle.weights(method, Ao))
        Xo, _, _ = oracle.sirt(method, Ao, B[:, j], 30, None, dict(opts, rho=rho))
        assert rel(X[:, j], Xo[:, -1]) <= 2 * TOL, (method, j)
        assert abs(info["resnorm"][j] - np.linalg.norm(B[:, j] - Ao @ X[:, j])) <= 1e-4 * info["resnorm"][j]


@pytest.mark.parametrize("path", PATHS)
def test_multi_warm_start_and_scalar_relaxpar(prob, monkeypatch, path):
    use_path(monkeypatch, path)
    A, _, B = prob
    X1, _ = air.sirt_multi("sart", A, B[:, :8], 10, None, {"relaxpar": 1.2})
    X2, _ = air.sirt_multi("sart", A, B[:, :8], 10, X1, {"relaxpar": 1.2})
    X3, _ = air.sirt_multi("sart", A, B[:, :8], 20, None, {"relaxpar": 1.2})
    assert rel(X2, X3) <= 1e-6
    with pytest.raises(Exception):
        air.sirt_multi("sart", A, B[:, :8], 10, None, {"stoprule": {"type": "DP", "taudelta": 1.0}})


def test_chunk_kernels_several_tiles_and_chunks(oracle, monkeypatch):
    """N=160 (a 3 x 3 grid of 64 x 64 tiles, the last row and column partly
    outside the image), 180 angles: several batches of 1024 pieces per tile,
    32 x 32 blocks with dozens of 256-row chunks, a warm start outside the box.
    Every column against its own single-RHS run and against the row kernels."""
    N, theta, p = 160, np.arange(0.0, 180.0, 1.0), 227
    A = air.paralleltomo(N, theta, p)
    Ao = oracle.paralleltomo(N, theta, p)
    rng = np.random.default_rng(5)
    xs = np.stack([air.phantomgallery("threephases", N, None, s).ravel(order="F") for s in range(64)], 1)
    B = Ao @ xs
    B = B + 0.01 * np.linalg.norm(B, axis=0) * rng.standard_normal(B.shape) / np.sqrt(B.shape[0])
    X0 = rng.uniform(-0.5, 1.5, xs.shape)
    opts = {"lbound": 0, "ubound": 1}
    monkeypatch.delenv("AIR_SPMM", raising=False)
    Xc, ic = air.sirt_multi("cimmino", A, B, 12, X0, opts)
    monkeypatch.setenv("AIR_SPMM", "0")
    Xr, ir = air.sirt_multi("cimmino", A, B, 12, X0, opts)
    monkeypatch.delenv("AIR_SPMM", raising=False)
    for j in range(64):
        assert rel(Xc[:, j], Xr[:, j]) <= TOL, j
    for j in (0, 31, 63):
        Xs, _, _ = air.sirt("cimmino", A, B[:, j], 12, X0[:, j], opts)
        assert rel(Xc[:, j], Xs[:, -1]) <= TOL, j
    np.testing.assert_allclose(ic["resnorm"], ir["resnorm"], rtol=1e-5)
    # SART (column weights D in the update) with a scalar relaxation parameter
    Xs2, _ = air.sirt_multi("sart", A, B[:, :16], 8, None, {"relaxpar": 1.5, "lbound": 0})
    for j in (0, 15):
        Xo, _, _ = oracle.sirt("sart", Ao, B[:, j], 8, None, {"relaxpar": 1.5, "lbound": 0})
        assert rel(Xs2[:, j], Xo[:, -1]) <= 2 * TOL, j


def test_chunk_kernels_general_matrix(oracle, monkeypatch):
    """A random sparse matrix that is not an image operator (columns in plain
    blocks of 4096) with 20 right-hand sides, against the oracle per column."""
    import scipy.sparse as sp
    rng = np.random.default_rng(11)
    m, n = 3000, 9000
    Ao = sp.random(m, n, density=0.004, random_state=12, format="csr") + sp.eye(m, n, format="csr")
    Ao = sp.csr_matrix(Ao)
    A = air.as_operator(Ao)
    xs = rng.uniform(0, 1, (n, 20))
    B = Ao @ xs
    monkeypatch.delenv("AIR_SPMM", raising=False)
    X, info = air.sirt_multi("cimmino", A, B, 15, None, {"lbound": 0})
    M, D = oracle.weights("cimmino", Ao)
    rho = oracle.spectral_radius(Ao, M, D)
    for j in (0, 7, 19):
        Xo, _, _ = oracle.sirt("cimmino", Ao, B[:, j], 15, None, {"lbound": 0, "rho": rho})
        assert rel(X[:, j], Xo[:, -1]) <= 2 * TOL, j
        assert abs(info["resnorm"][j] - np.linalg.norm(B[:, j] - Ao @ X[:, j])) <= 1e-4 * info["resnorm"][j]
