"""SIRT with many right-hand sides on one operator (SpMM passes) against one
run per right-hand side and the FP64 oracle."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


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


@pytest.mark.parametrize("method,nrhs", [("cimmino", 64), ("sart", 5), ("drop", 1), ("landweber", 17)])
def test_multi_equals_single_runs(prob, oracle, method, nrhs):
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


def test_multi_warm_start_and_scalar_relaxpar(prob):
    A, _, B = prob
    X1, _ = air.sirt_multi("sart", A, B[:, :8], 10, None, {"relaxpar": 1.2})
    X2, _ = air.sirt_multi("sart", A, B[:, :8], 10, X1, {"relaxpar": 1.2})
    X3, _ = air.sirt_multi("sart", A, B[:, :8], 20, None, {"relaxpar": 1.2})
    assert rel(X2, X3) <= 1e-6
    with pytest.raises(Exception):
        air.sirt_multi("sart", A, B[:, :8], 10, None, {"stoprule": {"type": "DP", "taudelta": 1.0}})
