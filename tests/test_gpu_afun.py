"""A caller's function handle as the operator (sirt/sirt.m:104-115,
auxil/check_inputs.m:67-75): the engine calls it for every product and runs
the rest of each iteration on the GPU; the weights follow get_mfun_dfun.m's
function-handle branches.  Iterates against the FP64 oracle with the same
handle-branch weights and rho."""
import numpy as np
import pytest

import airtoolsii_amd as air
from airtoolsii_amd._lib import AirError

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def prob(oracle):
    N, theta, p = 20, np.arange(0.0, 180.0, 5.0), 29
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    e = np.random.default_rng(5).standard_normal(b.size)
    b = b + 0.02 * np.linalg.norm(b) * e / np.linalg.norm(e)
    calls = {"n": 0}

    def afun(v, flag):
        if flag == "size":
            return Ao.shape
        calls["n"] += 1
        return Ao @ v if flag == "notransp" else Ao.T @ v

    return dict(Ao=Ao, b=b, x=x, afun=afun, calls=calls, m=Ao.shape[0], n=Ao.shape[1])


@pytest.mark.parametrize("method", ["landweber", "cimmino", "cav", "drop", "sart"])
def test_handle_weights_match_oracle(prob, oracle, method):
    A = air.as_operator(prob["afun"])
    assert isinstance(A, air.operators.FunctionOperator) and A.shape == prob["Ao"].shape
    M, D = air.weights(method, A)
    Mo, Do = oracle.weights_handle(method, prob["afun"], prob["m"], prob["n"])
    np.testing.assert_allclose(M, Mo, rtol=1e-13, atol=0)
    np.testing.assert_allclose(D, Do, rtol=1e-13, atol=0)


@pytest.mark.parametrize("method", ["landweber", "cimmino", "cav", "drop", "sart"])
def test_handle_sirt_matches_oracle(prob, oracle, method):
    Mo, Do = oracle.weights_handle(method, prob["afun"], prob["m"], prob["n"])
    rho = 1.0 if method == "sart" else oracle.spectral_radius(prob["Ao"], Mo, Do)
    opts = {"rho": rho, "lbound": 0}
    before = prob["calls"]["n"]
    X, info, _ = air.sirt(method, prob["afun"], prob["b"], [4, 25], None, opts)
    assert prob["calls"]["n"] - before >= 2 * 25            # every product went through the handle
    Xo, info_o, _ = oracle.sirt({"M": Mo, "D": Do}, prob["Ao"], prob["b"], [4, 25], None, opts)
    for j in range(2):
        assert rel(X[:, j], Xo[:, j]) <= TOL, (method, j, rel(X[:, j], Xo[:, j]))


def test_handle_default_rho_and_dp_stop(prob, oracle):
    """Default lambda = 1.9/rho with rho from the engine's Lanczos over handle
    products; the discrepancy principle stops where the oracle's does."""
    Mo, Do = oracle.weights_handle("cimmino", prob["afun"], prob["m"], prob["n"])
    rho_o = oracle.spectral_radius(prob["Ao"], Mo, Do)
    rho = air.spectral_radius("cimmino", prob["afun"])
    assert abs(rho - rho_o) <= 1e-5 * rho_o
    delta = 0.02 * np.linalg.norm(prob["Ao"] @ prob["x"])
    opts = {"rho": rho_o, "stoprule": {"type": "DP", "taudelta": 1.02 * delta}}
    X, info, _ = air.cimmino(prob["afun"], prob["b"], 400, None, opts)
    Xo, info_o, _ = oracle.sirt({"M": Mo, "D": Do}, prob["Ao"], prob["b"], 400, None, opts)
    assert info["stoprule"] == info_o["stoprule"] == 2
    assert info["finaliter"] == info_o["finaliter"]
    assert rel(X[:, -1], Xo[:, -1]) <= TOL


def test_handle_errors(prob):
    def bad_len(v, flag):
        if flag == "size":
            return prob["Ao"].shape
        return np.zeros(3)

    def raises(v, flag):
        if flag == "size":
            return prob["Ao"].shape
        raise RuntimeError("projector failed")

    for fn in (bad_len, raises):
        A = air.as_operator(fn)
        with pytest.raises(AirError):
            air.sirt("landweber", A, prob["b"], 3, None, {"rho": 1.0})
        assert A.error is not None
    # the no-allocation entry point has no host staging for a caller's function
    import ctypes as C
    from airtoolsii_amd import _lib
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    A = air.as_operator(prob["afun"])
    pin, pout = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(pin), 4 * A.shape[1]) == 0 and h.hipMalloc(C.byref(pout), 4 * A.shape[0]) == 0
    try:
        assert _lib.load().air_apply_device(A.engine.h, 0, pin, pout) == -4      # AIR_E_UNSUPPORTED
    finally:
        h.hipFree(pin)
        h.hipFree(pout)
