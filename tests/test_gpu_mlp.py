"""The learned relaxation network's device kernels (BF16 MFMA forward and
backward, Adam) against the FP64 reference in oracle/mlp.py."""
import numpy as np
import pytest

from airtoolsii_amd.mlp import MLP, NPARAM
from oracle import mlp as ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    return MLP()


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 16)).astype(np.float32)
    T = np.stack([rng.uniform(0.1, 1.9, n), (rng.uniform(size=n) > 0.6)], 1).astype(np.float32)
    T[::7, 0] = -1
    return X, T


@pytest.mark.parametrize("n", [1, 16, 37, 1000])
def test_forward_matches_reference(net, n):
    X, _ = _data(n, 1)
    p = ref.init_params(2)
    out = net.forward(p, X)
    o_emu, _ = ref.forward(p, X, emulate_bf16=True)
    o_fp64, _ = ref.forward(p, X)
    assert np.max(np.abs(out - o_emu)) <= 1e-4 * max(1.0, np.max(np.abs(o_emu)))
    assert np.max(np.abs(out - o_fp64)) <= 3e-2 * max(1.0, np.max(np.abs(o_fp64)))


@pytest.mark.parametrize("n", [16, 100, 5000])
def test_gradient_matches_reference(net, n):
    """Gradients of the device kernels against the CPU reference: within 1e-3 of
    the BF16-emulating reference, within 3e-2 of the exact FP64 gradient."""
    X, T = _data(n, 3)
    p = ref.init_params(4)
    loss, g = net.grad(p, X, T, stop_weight=0.5)
    L_emu, g_emu = ref.loss_grad(p, X, T, 0.5, emulate_bf16=True)
    L_64, g_64 = ref.loss_grad(p, X, T, 0.5)
    assert abs(loss - L_emu) <= 1e-5 * abs(L_emu)
    assert np.linalg.norm(g - g_emu) <= 1e-3 * np.linalg.norm(g_emu)
    assert np.linalg.norm(g - g_64) <= 3e-2 * np.linalg.norm(g_64)
    assert abs(loss - L_64) <= 1e-2 * abs(L_64)


def test_training_fits_a_smooth_target(net):
    """Adam on the device fits lambda targets that depend smoothly on the
    features and stop labels given by a threshold, well beyond chance."""
    rng = np.random.default_rng(7)
    n = 20000
    X = rng.standard_normal((n, 16)).astype(np.float32)
    t0 = 1.0 + 0.8 * np.tanh(0.7 * X[:, 0] - 0.5 * X[:, 3])
    t1 = (X[:, 1] + 0.5 * X[:, 2] > 0.3).astype(np.float32)
    T = np.stack([t0, t1], 1).astype(np.float32)
    p0 = ref.init_params(5)
    p0[-2] = np.log(1.0 / (ref.LAMBDA_MAX - 1.0))     # start at lambda rho = 1, the targets' centre
    L0, _ = net.grad(p0, X, T)
    p, hist = net.train(p0, X, T, epochs=15, batch=1024, lr=3e-3, seed=1)
    assert hist.shape == (15,) and np.all(np.isfinite(hist))
    L1, _ = net.grad(p, X, T)
    assert L1 < 0.15 * L0 and hist[-1] < hist[0]
    o = net.forward(p, X)
    lam = ref.LAMBDA_MAX / (1 + np.exp(-o[:, 0]))
    assert np.sqrt(np.mean((lam - t0) ** 2)) < 0.3 * np.std(t0)
    assert np.mean((o[:, 1] > 0) == (t1 > 0.5)) > 0.95
    # deterministic: the same seed gives the same weights
    p2, _ = net.train(p0, X, T, epochs=15, batch=1024, lr=3e-3, seed=1)
    assert np.array_equal(p, p2)
