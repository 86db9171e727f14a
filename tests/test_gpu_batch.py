"""Batches of independent problems: the block-diagonal tiled operator and one
joint SIRT run against separate runs of each problem and the FP64 oracle."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu

TOL = 1e-5
SPECS = [(24, np.arange(0.0, 180.0, 4.0)), (48, np.linspace(0, 180, 31)[:-1]), (33, np.arange(0.0, 180.0, 7.5)),
         (64, np.arange(0.0, 180.0, 2.0))]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def problems():
    out = []
    for i, (N, th) in enumerate(SPECS):
        x = air.phantomgallery("shepplogan" if i % 2 == 0 else "smooth", N).ravel(order="F")
        out.append((N, th, x))
    return out


@pytest.mark.parametrize("method", ["cimmino", "sart", "drop"])
def test_batch_equals_separate_runs(problems, oracle, method):
    Ns = [N for N, _, _ in problems]
    A = air.paralleltomo_batch(Ns, [th for _, th, _ in problems], method=method)
    assert A.nprob == len(Ns) and A.layout()["tiled"] == 1
    bs, singles = [], []
    for q, (N, th, x) in enumerate(problems):
        p = int(np.floor(np.sqrt(2) * N + 0.5))
        Aq = air.paralleltomo(N, th, p)
        assert (A.row_off[q + 1] - A.row_off[q], A.col_off[q + 1] - A.col_off[q]) == Aq.shape
        b = Aq.matvec(x)
        bs.append(b + 0.01 * np.linalg.norm(b) * np.random.default_rng(q).standard_normal(b.size) / np.sqrt(b.size))
        singles.append(Aq)
    xs, info = air.batch_sirt(A, bs, 25, {"lbound": 0})
    assert info["finaliter"] == 25
    for q, Aq in enumerate(singles):
        rho = 1.0 if method == "sart" else air.spectral_radius(method, Aq)
        assert abs(A.rho[q] - rho) <= 1e-6 * rho
        Xq, _, _ = air.sirt(method, Aq, bs[q], 25, None, {"lbound": 0, "rho": A.rho[q]})
        assert rel(xs[q], Xq[:, -1]) <= TOL, (method, q)
        # the FP64 oracle on the same problem
        Ao = oracle.paralleltomo(problems[q][0], problems[q][1], int(np.floor(np.sqrt(2) * problems[q][0] + 0.5)))
        Xo, _, _ = oracle.sirt(method, Ao, bs[q], 25, None, {"lbound": 0, "rho": A.rho[q]})
        assert rel(xs[q], Xo[:, -1]) <= TOL, (method, q)


def test_batch_apply_is_block_diagonal(problems):
    A = air.paralleltomo_batch([N for N, _, _ in problems], [th for _, th, _ in problems])
    x = np.concatenate([x for _, _, x in problems])
    y = A.matvec(x)
    for q, (N, th, xq) in enumerate(problems):
        Aq = air.paralleltomo(N, th, int(np.floor(np.sqrt(2) * N + 0.5)))
        yq = Aq.matvec(xq)
        assert rel(y[A.row_off[q]:A.row_off[q + 1]], yq) <= 1e-6
    w = np.random.default_rng(0).standard_normal(A.shape[0])
    g = A.rmatvec(w)
    assert abs(np.dot(y, w) - np.dot(x, g)) <= 1e-5 * abs(np.dot(y, w))
