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


def test_c2_matrix_free_products_match_matrix(c2):
    """The matrix-free pair at full size: the DDA forward against A32 x (A's
    values rounded to FP32) row by row, the pixel-driven adjoint against A32' y
    pixel by pixel, and the adjoint identity <A x, y> = <x, A' y>, which holds
    to rounding because both products use the same entries.  The floors of
    1e-9 max|x| and 1e-9 max|y| cover segments below the 1e-10 merge
    tolerance, where the DDA and the tracer may differ."""
    Ao = c2["Ao"]
    Af = air.paralleltomo(N, THETA, P, None, 0, 0)
    A32 = Ao.copy()
    A32.data = A32.data.astype(np.float32).astype(np.float64)
    Aabs = abs(Ao)
    rng = np.random.default_rng(11)
    x = rng.uniform(0, 1, Ao.shape[1]).astype(np.float32).astype(np.float64)
    y = rng.uniform(-1, 1, Ao.shape[0]).astype(np.float32).astype(np.float64)
    print("c2 parity: matrix-free products", flush=True)
    fx, ty = Af @ x, Af.T @ y
    assert np.all(np.abs(fx - A32 @ x) <= 1e-5 * (Aabs @ np.abs(x)) + 1e-9 * np.abs(x).max())
    assert np.all(np.abs(ty - A32.T @ y) <= 2e-7 * (Aabs.T @ np.abs(y)) + 1e-9 * np.abs(y).max())
    scale = float((Aabs @ np.abs(x)) @ np.abs(y))
    gap = abs(float(fx @ y) - float(x @ ty))
    print(f"c2 parity: adjoint identity gap {gap / scale:.3e}", flush=True)
    assert gap <= 1e-6 * scale
