"""SIRT at the headline size against the FP64 oracle: paralleltomo N=512,
angles 0:0.5:179.5, p=725 (m = 261,000, n = 262,144, about 1.2e8 entries),
grains phantom, 2 % noise, box [0, 1], K = 200 iterations of Cimmino, DROP
and SART, FP32 on the GPU (stored matrix and matrix-free projector) against
the oracle's FP64 iterates from the same inputs and the same rho.  The bar is
a relative 2-norm error of 1e-5 for x_K."""
import os

import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu

N, P, K = 512, 725, 200
THETA = np.arange(0.0, 180.0, 0.5)
TOL = 1e-5
THREADS = min(os.cpu_count() or 1, 16)


@pytest.fixture(scope="module")
def c2(oracle):
    print("c2 parity: building the oracle matrix", flush=True)
    Ao = oracle.paralleltomo(N, THETA, P)
    A = air.paralleltomo(N, THETA, P)
    x_true = air.phantomgallery("grains", N, None, 1).ravel(order="F")
    b, _ = air.add_noise(Ao @ x_true, 0.02, seed=1)
    return {"Ao": Ao, "A": A, "b": b}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("method", ["cimmino", "drop", "sart"])
def test_c2_sirt_iterates_match_oracle(oracle, c2, method):
    opts = {"lbound": 0, "ubound": 1}
    print(f"c2 parity: {method} K={K} on the GPU", flush=True)
    X, info, _ = air.sirt(method, c2["A"], c2["b"], K, None, opts)
    assert info["finaliter"] == K
    print(f"c2 parity: {method} K={K} on the oracle, rho {info['rho']:.9e}", flush=True)
    Xo, io, _ = oracle.sirt(method, c2["Ao"], c2["b"], K, None, dict(opts, rho=info["rho"]), nthreads=THREADS)
    err = rel(X[:, -1], Xo[:, -1])
    print(f"c2 parity: {method} rel err {err:.3e}", flush=True)
    assert err <= TOL


def test_c2_matrix_free_iterates_match_oracle(oracle, c2):
    opts = {"lbound": 0, "ubound": 1}
    Af = air.paralleltomo(N, THETA, P, None, 0, 0)
    X, info, _ = air.sirt("cimmino", Af, c2["b"], K, None, opts)
    Xo, _, _ = oracle.sirt("cimmino", c2["Ao"], c2["b"], K, None, dict(opts, rho=info["rho"]), nthreads=THREADS)
    err = rel(X[:, -1], Xo[:, -1])
    print(f"c2 parity: matrix-free cimmino rel err {err:.3e}", flush=True)
    assert err <= TOL
