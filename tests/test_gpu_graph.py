"""Graph mode of the SIRT loop (iterations captured once into a HIP graph and
replayed) against direct launches: the same kernels with the iteration number
and lambda read on the device, so the iterates must be bit-identical, for
constant and per-iteration lambda, save lists, stop rules inside a graph, warm
continuation in odd chunks and a batched operator."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def prob(oracle):
    N, theta, p = 48, np.arange(0.0, 180.0, 2.0), 68
    A = air.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = A @ x
    e = np.random.default_rng(3).standard_normal(b.size)
    b = b + 0.02 * np.linalg.norm(b) * e / np.linalg.norm(e)
    return A, b, x


CASES = {
    "default_lambda": ("cimmino", [100], {}),
    "psi1_schedule": ("sart", [60], {"relaxpar": "psi1"}),
    "vector_lambda_bounds": ("drop", [70], {"relaxpar": np.linspace(0.5, 1.5, 70), "lbound": 0.0, "ubound": 1.0}),
    "save_list": ("cimmino", [5, 17, 40, 100], {"lbound": 0.0}),
    "dp_stop_inside_graph": ("sart", [300], {"stoprule": {"type": "DP", "taudelta": None}}),
    "me_stop": ("cimmino", [400], {"stoprule": {"type": "ME", "taudelta": None}}),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("G", [8, 32])
def test_graph_matches_direct_launches(prob, name, G):
    A, b, x = prob
    meth, K, opts = CASES[name]
    opts = dict(opts)
    if "stoprule" in opts:
        X0, info0, _ = air.sirt(meth, A, b, K[-1])
        rn = info0["resnorm"]
        # a threshold met around iteration 37 or 150, inside a graph of G
        kk = 37 if name.startswith("dp") else 150
        opts["stoprule"] = {"type": opts["stoprule"]["type"], "taudelta": float(rn[kk]) * (1 + 1e-9)}
        if name.startswith("me"):
            opts["stoprule"]["taudelta"] = float(rn[kk]) * 0.999
    print(f"graph case {name} G={G}", flush=True)
    Xd, infod, _ = air.sirt(meth, A, b, K, None, opts)
    n0 = A.graph_launches()
    Xg, infog, _ = air.sirt(meth, A, b, K, None, dict(opts, graph=G))
    assert A.graph_launches() > n0             # the graph path ran
    assert infog["finaliter"] == infod["finaliter"] and infog["stoprule"] == infod["stoprule"]
    assert list(infog["itersaved"]) == list(infod["itersaved"])
    np.testing.assert_array_equal(Xg, Xd)
    np.testing.assert_array_equal(infog["resnorm"], infod["resnorm"])
    if "stoprule" in opts:
        assert infod["finaliter"] < K[-1]


def test_graph_replays_between_eager_work(prob):
    """Replays of one captured graph interleaved with host synchronisation and
    other device work, each checked against the direct-launch run."""
    A, b, x = prob
    # both runs live on the same operator (one run state per engine): the
    # direct run first, recording x after every 8 iterations
    rd = air.SirtRun("cimmino", A, b, 96, None, {"lbound": 0.0})
    xs = []
    for _ in range(12):
        rd.iterate(8)
        xs.append(rd.fetch()[0])
    rg = air.SirtRun("cimmino", A, b, 96, None, {"lbound": 0.0, "graph": 8})
    for q in range(12):
        rg.iterate(8)
        y = A @ np.ones(A.shape[1])          # eager work on the same stream between replays
        assert np.isfinite(y).all()
        xg, info = rg.fetch()
        assert info["finaliter"] == 8 * (q + 1)
        np.testing.assert_array_equal(xg, xs[q])


def test_graph_odd_chunks_and_remainders(prob):
    A, b, x = prob
    rd = air.SirtRun("sart", A, b, 100, None, {"relaxpar": "psi2"})
    rd.iterate(100)
    xd, infod = rd.fetch()
    rg = air.SirtRun("sart", A, b, 100, None, {"relaxpar": "psi2", "graph": 16})
    n0 = A.graph_launches()
    for c in (3, 29, 50, 18):
        rg.iterate(c)
    assert A.graph_launches() - n0 == 0 + 1 + 3 + 1    # 29 -> 16+13, 50 -> 3x16+2, 18 -> 16+2
    xg, infog = rg.fetch()
    assert infog["finaliter"] == 100
    np.testing.assert_array_equal(xg, xd)
    np.testing.assert_array_equal(infog["resnorm"], infod["resnorm"])


def test_graph_on_a_batched_operator():
    thetas = [np.arange(0.0, 180.0, 180.0 / na) for na in (30, 45, 20)]
    B = air.paralleltomo_batch([32, 40, 24], thetas, None, "cimmino")
    rng = np.random.default_rng(5)
    bs = [rng.uniform(0, 1, int(B.row_off[i + 1] - B.row_off[i])) for i in range(B.nprob)]
    xd, _ = air.batch_sirt(B, bs, 50)
    n0 = B.graph_launches()
    xg, _ = air.batch_sirt(B, bs, 50, {"graph": 16})
    assert B.graph_launches() > n0
    for a, c in zip(xd, xg):
        np.testing.assert_array_equal(a, c)
