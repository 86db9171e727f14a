"""ART (art/art.m) and CART (cart/cart.m) on the engine against the FP64
oracle's literal sweeps: same step orders, same random streams, same stop
rules.  With FP64 storage the iterates agree to rounding; with FP32 storage
they agree with the oracle run on the FP32-rounded matrix."""
import numpy as np
import pytest
import scipy.sparse as sp

import airtoolsii_amd as air
from airtoolsii_amd._lib import AirError

pytestmark = pytest.mark.gpu

N, THETA, P = 24, np.arange(0.0, 180.0, 6.0), 34


@pytest.fixture(scope="module")
def prob(oracle):
    Ao = oracle.paralleltomo(N, THETA, P)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    e = np.random.default_rng(1).standard_normal(b.size)
    bn = b + 0.02 * np.linalg.norm(b) * e / np.linalg.norm(e)
    A64 = air.paralleltomo(N, THETA, P, dtype="f64")
    A32 = air.paralleltomo(N, THETA, P, dtype="f32")
    Ao32 = Ao.copy()
    Ao32.data = Ao32.data.astype(np.float32).astype(np.float64)
    return {"Ao": Ao, "Ao32": Ao32, "A64": A64, "A32": A32, "x": x, "b": bn}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


ART_CASES = {
    "kaczmarz_plain": ("kaczmarz", [3, 10], {}),
    "kaczmarz_bounds_damp": ("kaczmarz", [5], {"lbound": 0.0, "ubound": 1.0, "damp": 0.05, "relaxpar": 0.6}),
    "symkaczmarz": ("symkaczmarz", [4, 10], {"lbound": 0.0}),
    "randkaczmarz": ("randkaczmarz", [2, 6], {"seed": 17}),
    "custom_order": (list(range(1020, 0, -3)), [4], {"relaxpar": 1.2}),
    "relax_function": ("kaczmarz", [4], {"relaxpar": lambda j: 0.5 + 0.4 * np.cos(np.asarray(j) / 500.0)}),
    "dp_stop": ("kaczmarz", [50], {"stoprule": {"type": "DP", "taudelta": 1.5}}),
    "me_stop": ("kaczmarz", [50], {"stoprule": {"type": "ME", "taudelta": 1.2}}),
    "ncp_stop_2d": ("kaczmarz", [5, 50], {"stoprule": {"type": "NCP", "res_dims": [P, len(THETA)]}}),
    "ncp_stop_1d": ("kaczmarz", [50], {"stoprule": {"type": "NCP", "res_dims": [1020], "ncp_smooth": 1}}),
    "ncp_stop_rand": ("randkaczmarz", [50], {"seed": 3, "stoprule": {"type": "NCP", "res_dims": [P, len(THETA)],
                                                                     "ncp_smooth": 3}}),
}


@pytest.mark.parametrize("name", list(ART_CASES))
def test_art_matches_oracle_fp64(oracle, prob, name):
    meth, K, opts = ART_CASES[name]
    print(f"art case {name}", flush=True)
    X, info = air.art(meth, prob["A64"], prob["b"], K, None, opts)
    Xo, io = oracle.art(meth, prob["Ao"], prob["b"], K, None, opts)
    assert info["finaliter"] == io["finaliter"] and info["stoprule"] == io["stoprule"]
    assert X.shape == Xo.shape
    for q in range(X.shape[1]):
        assert rel(X[:, q], Xo[:, q]) <= 1e-10, (q, rel(X[:, q], Xo[:, q]))


def test_art_fp32_storage_matches_oracle_on_rounded_matrix(oracle, prob):
    X, info = air.kaczmarz(prob["A32"], prob["b"], [5], None, {"lbound": 0.0})
    Xo, _ = oracle.art("kaczmarz", prob["Ao32"], prob["b"], [5], None, {"lbound": 0.0})
    assert rel(X[:, -1], Xo[:, -1]) <= 1e-10
    Xf, _ = oracle.art("kaczmarz", prob["Ao"], prob["b"], [5], None, {"lbound": 0.0})
    assert rel(X[:, -1], Xf[:, -1]) <= 1e-5


@pytest.mark.parametrize("form", ["global", "lds_one_wave", "lds_chunked"])
@pytest.mark.parametrize("name", ["kaczmarz_bounds_damp", "randkaczmarz", "relax_function", "symkaczmarz",
                                  "custom_order"])
def test_art_global_and_lds_forms(oracle, prob, monkeypatch, form, name):
    """x in global memory with a fence per step, x in LDS with one wave, or x
    in LDS with the rows staged chunk by chunk by the rest of the work-group:
    the same steps, so the same iterates."""
    monkeypatch.setenv("AIR_ART_LDS", "0" if form == "global" else "1")
    monkeypatch.setenv("AIR_ART_CHUNK", "1" if form == "lds_chunked" else "0")
    meth, K, opts = ART_CASES[name]
    X, info = air.art(meth, prob["A64"], prob["b"], K, None, opts)
    Xo, io = oracle.art(meth, prob["Ao"], prob["b"], K, None, opts)
    assert info["finaliter"] == io["finaliter"]
    for q in range(X.shape[1]):
        assert rel(X[:, q], Xo[:, q]) <= 1e-10


def test_art_image_beyond_64k_lds(oracle):
    # N = 100: x takes 80 KB, above 64 KB of LDS (dynamic LDS attribute, or the
    # global form where the device allows less); rows up to 200 entries
    N, theta, p = 100, np.arange(0.0, 180.0, 10.0), 141
    A = air.paralleltomo(N, theta, p, dtype="f64")
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    X, info = air.kaczmarz(A, b, [3], None, {"lbound": 0.0})
    Xo, _ = oracle.art("kaczmarz", Ao, b, [3], None, {"lbound": 0.0})
    assert rel(X[:, -1], Xo[:, -1]) <= 1e-10


CART_CASES = {
    "columnaction_default": ("columnaction", [3, 6], {}),
    "flagging": ("columnaction", [2, 8], {"relaxpar": 1.0, "Kbegin": 3, "THR": 1e-3, "Nunflag": 3, "seed": 4}),
    "bounds_custom_order": (list(range(576, 0, -1)), [5], {"lbound": 0.0, "ubound": np.full(576, 0.9),
                                                          "relaxpar": 0.8}),
    "me_stop": ("columnaction", [300], {"relaxpar": 1.0, "stoprule": {"type": "ME", "taudelta": 1.2}}),
    "ncp_stop": ("columnaction", [300], {"stoprule": {"type": "NCP", "res_dims": [P, len(THETA)], "ncp_smooth": 4}}),
}


@pytest.mark.parametrize("name", list(CART_CASES))
def test_cart_matches_oracle_fp64(oracle, prob, name):
    meth, K, opts = CART_CASES[name]
    print(f"cart case {name}", flush=True)
    X, info = air.cart(meth, prob["A64"], prob["b"], K, None, opts)
    Xo, io = oracle.cart(meth, prob["Ao"], prob["b"], K, None, opts)
    assert info["finaliter"] == io["finaliter"] and info["stoprule"] == io["stoprule"]
    for q in range(X.shape[1]):
        assert rel(X[:, q], Xo[:, q]) <= 1e-10, (q, rel(X[:, q], Xo[:, q]))


@pytest.mark.parametrize("form", ["one_wave", "chunked"])
@pytest.mark.parametrize("name", ["columnaction_default", "flagging", "bounds_custom_order", "repeated_columns"])
def test_cart_one_wave_and_chunked_forms(oracle, prob, monkeypatch, form, name):
    """CART with the residual in global memory (one wave) or in LDS with the
    columns staged chunk by chunk; an order that repeats columns always runs
    the one-wave form.  Same steps, so the same iterates."""
    monkeypatch.setenv("AIR_ART_CHUNK", "1" if form == "chunked" else "0")
    if name == "repeated_columns":
        meth, K, opts = list(range(576, 0, -1)) + [5, 300, 5, 17], [4], {"relaxpar": 0.9, "lbound": 0.0}
    else:
        meth, K, opts = CART_CASES[name]
    X, info = air.cart(meth, prob["A64"], prob["b"], K, None, opts)
    Xo, io = oracle.cart(meth, prob["Ao"], prob["b"], K, None, opts)
    assert info["finaliter"] == io["finaliter"] and info["stoprule"] == io["stoprule"]
    for q in range(X.shape[1]):
        assert rel(X[:, q], Xo[:, q]) <= 1e-10, (q, rel(X[:, q], Xo[:, q]))


def test_art_info_and_errors(prob):
    X, info = air.symkaczmarz(prob["A32"], prob["b"], [2, 6])
    assert info["finaliter"] == 6 and list(info["itersaved"]) == [2, 6] and X.shape[1] == 2
    assert len(info["resnorm"]) == 4 and np.all(np.diff(info["resnorm"]) < 0)
    with pytest.raises(ValueError, match="even"):
        air.symkaczmarz(prob["A32"], prob["b"], [3])
    with pytest.raises(ValueError, match="Unknown ART"):
        air.art("bogus", prob["A32"], prob["b"], 2)
    with pytest.raises(ValueError, match="res_dims must be specified"):
        air.kaczmarz(prob["A32"], prob["b"], 2, None, {"stoprule": {"type": "NCP"}})
    with pytest.raises(AirError, match="prod\\(res_dims\\) == m"):
        air.columnaction(prob["A32"], prob["b"], 2, None, {"stoprule": {"type": "NCP", "res_dims": [P, 7]}})


def test_art_ncp_stop_is_where_the_statistic_first_rises(oracle, prob):
    """The engine's NCP decisions against the oracle's statistic of the
    engine's own iterates: with smooth 2 the run stops at the first sweep whose
    ncc exceeds both of the two before it."""
    dims = [P, len(THETA)]
    X, info = air.kaczmarz(prob["A64"], prob["b"], np.arange(1, 51), None,
                           {"stoprule": {"type": "NCP", "res_dims": dims}})
    k = info["finaliter"]
    assert info["stoprule"] == 1 and 2 <= k < 50 and X.shape[1] == k
    R = [prob["b"]] + [prob["b"] - prob["Ao"] @ X[:, j] for j in range(k)]
    ncc = [oracle.ncp_stat(r, dims) for r in R]
    padded = [np.inf, np.inf] + ncc                  # dk starts as Inf
    assert ncc[k] > max(padded[k:k + 2])
    assert all(ncc[j] <= max(padded[j:j + 2]) for j in range(k))
    with pytest.raises(ValueError, match="scalar"):
        air.cart("columnaction", prob["A32"], prob["b"], 2, None, {"relaxpar": lambda j: 1.0})
    Af = air.paralleltomo(N, THETA, P, None, 0, 0)
    with pytest.raises(Exception, match="stored matrix"):
        air.kaczmarz(Af, prob["b"], 1)


@pytest.mark.parametrize("meth", ["kaczmarz", "columnaction"])
def test_train_relaxpar_row_action_matches_oracle(oracle, prob, meth):
    """train_relaxpar.m's ART/CART branch (relaxpar 0.25 first, relaxparmax 2,
    upper probe tested first) on the engine's sweeps and on the oracle's."""
    print(f"train_relaxpar {meth}", flush=True)
    lam = air.train_relaxpar(prob["A64"], prob["b"], prob["x"], meth, 30)
    lam_o = oracle.train_relaxpar(prob["Ao"], prob["b"], prob["x"], meth, 30)
    assert 0 < lam < 2 and abs(lam - lam_o) <= 1e-6, (lam, lam_o)


def test_train_dpme_with_kaczmarz_matches_oracle(oracle, prob):
    b0 = prob["Ao"] @ prob["x"]
    E = np.random.default_rng(9).standard_normal((b0.size, 2))
    delta = 0.02 * np.linalg.norm(b0)
    for kind in ("DP", "ME"):
        tau = air.train_dpme(prob["A64"], b0, prob["x"], "kaczmarz", kind, delta, 2, 20, noise=E)
        tau_o = oracle.train_dpme(prob["Ao"], b0, prob["x"], "kaczmarz", kind, delta, 2, 20, noise=E)
        assert abs(tau - tau_o) <= 1e-8 * abs(tau_o), (kind, tau, tau_o)
