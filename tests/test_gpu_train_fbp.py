"""train_dpme (train/train_dpme.m) and fbp (auxil/fbp.m) on the engine against
the FP64 oracle: the same noise columns, the same chunked search for the
iteration of smallest error, the same filters."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problem(oracle):
    N, theta, p = 32, np.arange(0.0, 180.0, 3.0), 45
    A = air.paralleltomo(N, theta, p)
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    return A, Ao, x, theta


@pytest.mark.parametrize("kind", ["DP", "ME"])
def test_train_dpme_matches_oracle(oracle, problem, kind):
    A, Ao, x, _ = problem
    b = Ao @ x
    delta = 0.05 * np.linalg.norm(b)
    noise = np.random.default_rng(7).standard_normal((b.size, 3))
    tau = air.train_dpme(A, b, x, "sart", kind, delta, 3, 200, noise=noise)
    tau_o = oracle.train_dpme(Ao, b, x, "sart", kind, delta, 3, 200, None, noise)
    print(kind, tau, tau_o, flush=True)
    assert abs(tau - tau_o) <= 1e-4 * abs(tau_o)
    assert 0.5 < tau < 2.0


def test_train_dpme_tau_stops_near_the_best_iterate(oracle, problem):
    # the trained DP tau, used as the stop rule on a fresh noisy right-hand side,
    # stops within a few iterations of the iterate of smallest error
    A, Ao, x, _ = problem
    b = Ao @ x
    delta = 0.05 * np.linalg.norm(b)
    tau = air.train_dpme(A, b, x, "sart", "DP", delta, 4, 200, seed=3)
    e = np.random.default_rng(11).standard_normal(b.size)
    bn = b + delta * e / np.linalg.norm(e)
    X, _, _ = air.sirt("sart", A, bn, np.arange(1, 201))
    err = np.linalg.norm(X - x[:, None], axis=0)
    kbest = int(np.argmin(err)) + 1
    Xs, info, _ = air.sirt("sart", A, bn, 200, None, {"stoprule": {"type": "DP", "taudelta": tau * delta}})
    assert abs(info["finaliter"] - kbest) <= max(5, 0.25 * kbest)
    assert np.linalg.norm(Xs[:, -1] - x) <= 1.05 * err.min()


@pytest.mark.parametrize("filt", ["ram-lak", "shepp-logan", "cosine", "hamming", "hann", "none"])
def test_fbp_matches_oracle(oracle, problem, filt):
    A, Ao, x, theta = problem
    b = Ao @ x
    xf = air.fbp(A, b, theta, filt)
    xo = oracle.fbp(Ao, b, theta, filt)
    assert np.linalg.norm(xf - xo) <= 1e-5 * np.linalg.norm(xo)


def test_fbp_argument_errors(problem):
    A, Ao, x, theta = problem
    b = Ao @ x
    with pytest.raises(ValueError):
        air.fbp(A, b, theta, "gaussian")
    with pytest.raises(ValueError):
        air.fbp(A, b, theta[:-1])          # 60*45 rows are not a multiple of 59 angles
