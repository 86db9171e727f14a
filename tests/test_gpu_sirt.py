"""SIRT on the GPU (K2/K3/K4) against the FP64 oracle: iterates within 1e-5
relative after K iterations on identical inputs (same A, b, rho), stop rules,
save lists, bounds, relaxation strategies, weights, rho and train_relaxpar."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu

TOL = 1e-5          # relative 2-norm error of FP32 iterates vs the FP64 oracle


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def c1(oracle):
    """C1: paralleltomo N=64, 0:179, p=91; Shepp-Logan; 1% seeded noise."""
    N, theta, p = 64, np.arange(0.0, 180.0), 91
    Ao = oracle.paralleltomo(N, theta, p)
    A = air.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    e = np.random.default_rng(1).standard_normal(b.size)
    b = b + 0.01 * np.linalg.norm(b) * e / np.linalg.norm(e)
    rho = {mth: (1.0 if mth == "sart" else oracle.spectral_radius(Ao, *oracle.weights(mth, Ao)))
           for mth in ("landweber", "cimmino", "cav", "drop", "sart")}
    return dict(A=A, Ao=Ao, x=x, b=b, rho=rho, N=N, theta=theta, p=p)


@pytest.mark.parametrize("method", ["landweber", "cimmino", "cav", "drop", "sart"])
def test_methods_match_oracle_c1(c1, oracle, method):
    opts = {"rho": c1["rho"][method], "lbound": 0}
    X, info, _ = air.sirt(method, c1["A"], c1["b"], 50, None, opts)
    Xo, info_o, _ = oracle.sirt(method, c1["Ao"], c1["b"], 50, None, opts)
    err = rel(X[:, -1], Xo[:, -1])
    print(f"{method}: rel err {err:.2e}")
    assert err <= TOL
    assert info["finaliter"] == 50 and info["stoprule"] == 0
    assert abs(info["relaxpar"] - 1.9 / c1["rho"][method]) <= 1e-15 * info["relaxpar"]


@pytest.mark.parametrize("method", ["cimmino", "sart", "drop"])
def test_wrappers_and_save_list(c1, oracle, method):
    fn = getattr(air, method)
    K = [5, 17, 40]
    opts = {"rho": c1["rho"][method], "lbound": 0, "ubound": 1}
    X, info, _ = fn(c1["A"], c1["b"], K, None, opts)
    Xo, info_o, _ = oracle.sirt(method, c1["Ao"], c1["b"], K, None, opts)
    assert X.shape == Xo.shape == (c1["A"].shape[1], 3)
    assert list(info["itersaved"]) == list(info_o["itersaved"]) == K
    for j in range(3):
        assert rel(X[:, j], Xo[:, j]) <= TOL
    assert X.min() >= 0 and X.max() <= 1


@pytest.mark.parametrize("strategy", ["line", "psi1", "psi1mod", "psi2", "psi2mod", 0.8])
def test_relaxation_strategies(c1, oracle, strategy):
    method = "sart"
    opts = {"relaxpar": strategy, "rho": 1.0, "lbound": 0}
    X, info, _ = air.sirt(method, c1["A"], c1["b"], 30, None, opts)
    Xo, info_o, _ = oracle.sirt(method, c1["Ao"], c1["b"], 30, None, opts)
    assert rel(X[:, -1], Xo[:, -1]) <= TOL
    if strategy == "line":
        assert abs(info["relaxpar"] - info_o["relaxpar"]) <= 1e-4 * abs(info_o["relaxpar"])


def test_custom_method_and_landweber_identity(c1, oracle):
    A, Ao, b = c1["A"], c1["Ao"], c1["b"]
    rn2 = np.asarray(Ao.multiply(Ao).sum(1)).ravel()
    Mc = np.where(rn2 > 0, 1 / np.where(rn2 > 0, rn2, 1), 0)
    Dc = np.linspace(0.5, 1.5, A.shape[1])
    opts = {"rho": oracle.spectral_radius(Ao, Mc, Dc)}
    X, _, ext = air.sirt({"M": Mc, "D": Dc}, A, b, 25, None, opts)
    Xo, _, _ = oracle.sirt({"M": Mc, "D": Dc}, Ao, b, 25, None, opts)
    assert rel(X[:, -1], Xo[:, -1]) <= TOL
    assert np.array_equal(ext.M, Mc) and np.array_equal(ext.D, Dc)
    X1, _, _ = air.sirt({}, A, b, 10, None, {"rho": c1["rho"]["landweber"]})
    X2, _, _ = air.landweber(A, b, 10, None, {"rho": c1["rho"]["landweber"]})
    assert np.array_equal(X1, X2)


def test_warm_start_and_vector_bounds(c1, oracle):
    n = c1["A"].shape[1]
    rng = np.random.default_rng(3)
    x0 = rng.uniform(0, 0.5, n)
    lb = np.where(np.arange(n) % 3 == 0, 0.1, -np.inf)
    ub = np.where(np.arange(n) % 5 == 0, 0.3, np.inf)
    opts = {"rho": c1["rho"]["cimmino"], "lbound": lb, "ubound": ub}
    X, _, _ = air.cimmino(c1["A"], c1["b"], 20, x0, opts)
    Xo, _, _ = oracle.sirt("cimmino", c1["Ao"], c1["b"], 20, x0, opts)
    assert rel(X[:, -1], Xo[:, -1]) <= TOL
    assert np.all(X[::3, -1] >= 0.1 - 1e-7) and np.all(X[::5, -1] <= 0.3 + 1e-7)


def test_discrepancy_principle(c1, oracle):
    opts = {"rho": 1.0, "lbound": 0}
    _, _, _, rn = oracle.sirt("sart", c1["Ao"], c1["b"], 60, None, opts, return_resnorm=True)
    k = 23
    tau = float(np.sqrt(rn[k] * rn[k + 1]))          # between two iterations: robust to FP32 norms
    o2 = dict(opts, stoprule={"type": "DP", "taudelta": tau})
    X, info, _ = air.sart(c1["A"], c1["b"], [10, 20, 30, 60], None, o2)
    Xo, info_o, _ = oracle.sirt("sart", c1["Ao"], c1["b"], [10, 20, 30, 60], None, o2)
    assert info["stoprule"] == info_o["stoprule"] == 2
    assert info["finaliter"] == info_o["finaliter"] == k + 1
    assert list(info["itersaved"]) == list(info_o["itersaved"]) == [10, 20, k + 1]
    assert X.shape[1] == 3 and rel(X[:, -1], Xo[:, -1]) <= TOL
    np.testing.assert_allclose(info["resnorm"], rn[: k + 2], rtol=1e-5)
    # satisfied at x0: empty X, itersaved = [0]
    X, info, _ = air.sart(c1["A"], c1["b"], 10, None, dict(opts, stoprule={"type": "DP", "taudelta": 1e12}))
    assert X.shape[1] == 0 and info["finaliter"] == 0 and list(info["itersaved"]) == [0]


def test_monotone_error_rule(c1, oracle):
    opts = {"rho": c1["rho"]["cimmino"]}
    kmax = 60
    Xall, _, _ = oracle.sirt("cimmino", c1["Ao"], c1["b"], np.arange(1, kmax + 1), None, opts)
    R = [c1["b"]] + [c1["b"] - c1["Ao"] @ Xall[:, k] for k in range(kmax)]
    crit = np.array([(R[k - 1] @ (R[k - 1] + R[k])) / 2 / np.linalg.norm(R[k - 1]) for k in range(1, kmax + 1)])
    k = 30
    tau = float(0.5 * (crit[k - 1] + crit[k]))       # between two iterations
    o2 = dict(opts, stoprule={"type": "ME", "taudelta": tau})
    X, info, _ = air.cimmino(c1["A"], c1["b"], kmax, None, o2)
    Xo, info_o, _ = oracle.sirt("cimmino", c1["Ao"], c1["b"], kmax, None, o2)
    assert info["stoprule"] == info_o["stoprule"] == 3
    assert info["finaliter"] == info_o["finaliter"]
    assert rel(X[:, -1], Xo[:, -1]) <= TOL


@pytest.mark.parametrize("method", ["cimmino", "cav", "drop", "sart"])
def test_weights_match_oracle(c1, oracle, method):
    Mo, Do = oracle.weights(method, c1["Ao"])
    A64 = air.paralleltomo(c1["N"], c1["theta"], c1["p"], dtype="f64")
    M, D = air.weights(method, A64)
    np.testing.assert_allclose(M, Mo, rtol=1e-12)
    np.testing.assert_allclose(D, Do, rtol=1e-12)
    M, D = air.weights(method, c1["A"])
    np.testing.assert_allclose(M, Mo, rtol=2e-6)
    np.testing.assert_allclose(D, Do, rtol=2e-6)


@pytest.mark.parametrize("walk", ["dda", "exact"])
@pytest.mark.parametrize("method", ["cimmino", "cav", "drop", "sart"])
def test_matrix_free_weights_match_oracle(c1, oracle, monkeypatch, method, walk):
    """Weights of the matrix-free operator: row sums from the tile walk (FP64
    sums of the FP32-rounded entries), column sums and counts from the ray
    walk, against the oracle's FP64 weights of the stored matrix."""
    monkeypatch.setenv("AIR_MF_EXACT", "1" if walk == "exact" else "0")
    Mo, Do = oracle.weights(method, c1["Ao"])
    Af = air.paralleltomo(c1["N"], c1["theta"], c1["p"], None, 0, 0)
    M, D = air.weights(method, Af)
    np.testing.assert_allclose(M, Mo, rtol=2e-6)
    np.testing.assert_allclose(D, Do, rtol=2e-6)


@pytest.mark.parametrize("method", ["landweber", "cimmino", "drop"])
def test_spectral_radius(c1, method):
    rho = air.spectral_radius(method, c1["A"])
    assert abs(rho - c1["rho"][method]) <= 1e-5 * c1["rho"][method]


def test_default_rho_path(c1, oracle):
    X, info, _ = air.cimmino(c1["A"], c1["b"], 20)
    assert abs(info["rho"] - c1["rho"]["cimmino"]) <= 1e-5 * c1["rho"]["cimmino"]
    Xo, _, _ = oracle.sirt("cimmino", c1["Ao"], c1["b"], 20, None, {"rho": c1["rho"]["cimmino"]})
    assert rel(X[:, -1], Xo[:, -1]) <= 1e-4


@pytest.mark.parametrize("method", ["sart", "cimmino"])
def test_matrix_free_sirt_matches_matrix(c1, oracle, method):
    Af = air.paralleltomo(c1["N"], c1["theta"], c1["p"], None, 0, 0)
    opts = {"rho": c1["rho"][method], "lbound": 0}
    X, info, _ = air.sirt(method, Af, c1["b"], 20, None, opts)
    Xo, _, _ = oracle.sirt(method, c1["Ao"], c1["b"], 20, None, opts)
    assert rel(X[:, -1], Xo[:, -1]) <= TOL


@pytest.mark.parametrize("dtype,tol", [("bf16", 2e-2), ("f16", 3e-3), ("f64", 1e-6)])
def test_reduced_precision_storage(c1, oracle, dtype, tol):
    A = air.paralleltomo(c1["N"], c1["theta"], c1["p"], dtype=dtype)
    opts = {"rho": 1.0, "lbound": 0}
    X, _, _ = air.sart(A, c1["b"], 30, None, opts)
    Xo, _, _ = oracle.sirt("sart", c1["Ao"], c1["b"], 30, None, opts)
    err = rel(X[:, -1], Xo[:, -1])
    print(f"{dtype}: rel err vs FP64 oracle {err:.2e}")
    assert err <= tol


def test_fanlinear_sirt(oracle):
    N, theta, p = 48, np.arange(0.0, 360.0, 4.0), 68
    A = air.fanlineartomo(N, theta, p)
    Ao = oracle.fanlineartomo(N, theta, p)
    x = oracle.phantomgallery("threephases", N, None, 2).ravel(order="F")
    b = Ao @ x
    opts = {"rho": 1.0, "lbound": 0, "ubound": 1}
    X, _, _ = air.sart(A, b, 40, None, opts)
    Xo, _, _ = oracle.sirt("sart", Ao, b, 40, None, opts)
    assert rel(X[:, -1], Xo[:, -1]) <= TOL


def test_train_relaxpar_matches_oracle(oracle):
    N, theta, p = 32, np.arange(0.0, 180.0, 3.0), 45
    A = air.paralleltomo(N, theta, p)
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    e = np.random.default_rng(2).standard_normal(b.size)
    b = b + 0.03 * np.linalg.norm(b) * e / np.linalg.norm(e)
    lam = air.train_relaxpar(A, b, x, "sart", 100)
    lam_o = oracle.train_relaxpar(Ao, b, x, "sart", 100)
    assert abs(lam - lam_o) <= 0.03 * 2.0


# Every forward/backward kernel shape of the tiled layout (the autotuner picks
# one per matrix): (forward shape, backward shape, pieces per block or entries
# per wave).  Forward shapes 0-4 are per-piece lane groups, 5-8 the stream
# form (8 or 16 entries per lane, without/with prefetch), 9-10 the lane form
# (whole-lane segments, without/with prefetch), 11-12 the step-coded lane form
# (one word of unit moves per 8 entries); backward shapes 0-3 are
# per-column lane groups, 4-6 the stream form (8 entries with prefetch, 16
# without, 16 with), 7-8 the lane form.  N=96 gives 4 forward tiles with
# padded columns.  A fourth field of 16 puts 16 waves in each block of the
# step-coded forward (one staged tile of x per 16 waves); the other forms keep 4.
@pytest.mark.parametrize("cfg", ["0,0,256", "1,1,512", "2,2,1024", "3,3,256", "4,0,512",
                                 "5,1,512", "6,2,2048", "7,3,8192", "8,0,100000",
                                 "5,4,2048", "7,5,4096", "0,6,256", "8,6,512",
                                 "9,7,2048", "10,8,8192", "9,8,512", "10,7,100000", "0,7,256", "10,0,4096",
                                 "5,8,1024", "11,7,4096", "12,8,2048", "11,0,100000", "12,5,512",
                                 "11,9,4096", "12,10,2048", "0,9,256", "9,10,512", "5,9,8192",
                                 "11,10,4096,16", "12,10,2048,16", "12,9,512,16", "11,8,100000,16",
                                 "12,0,1024,16", "10,10,2048,16"])
def test_tiled_kernel_shapes(oracle, monkeypatch, cfg):
    N, theta, p = 96, np.arange(0.0, 180.0, 2.0), 137
    monkeypatch.setenv("AIR_TILED_CFG", cfg)
    A = air.paralleltomo(N, theta, p)
    L = A.layout()
    f, bcfg, par, *rest = (int(v) for v in cfg.split(","))
    wpb = rest[0] if rest else 4
    assert L["fwd_waves_per_block"] == (wpb if f >= 11 else 4)
    assert L["tiled"] == 1 and L["fwd_stream"] == (1 if f >= 5 else 0)
    assert L["bwd_lanes"] == {4: 8, 5: 16, 6: 16, 7: 0, 8: 0, 9: 0, 10: 0}.get(bcfg, L["bwd_lanes"])
    assert L["bwd_step_ok"] == 1 and L["bwd_step"] == (1 if bcfg >= 9 else 0)
    assert L["bwd_step_entries"] >= L["nnz"] and L["bwd_step_entries"] % 8 == 0
    if f >= 5:
        assert (L["fwd_lanes"], L["fwd_loads"]) == [(8, 0), (8, 1), (16, 0), (16, 1), (0, 0), (0, 1),
                                                    (0, 0), (0, 1)][f - 5]
    assert L["fwd_step_ok"] == 1 and L["fwd_step"] == (1 if f >= 11 else 0)
    # the lane form's streams: every segment padded to whole lanes of 8 entries
    assert L["fwd_entries"] % 8 == 0 and L["bwd_entries"] % 8 == 0
    assert L["nnz"] <= L["fwd_entries"] < L["nnz"] + 8 * L["pieces"]
    assert L["nnz"] <= L["bwd_entries"] < L["nnz"] + 8 * L["n_internal"]
    assert L["fwd_pieces_per_block"] == (max(par, 512) if f >= 5 else par)
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("threephases", N, None, 3).ravel(order="F")
    b = Ao @ x
    opts = {"rho": oracle.spectral_radius(Ao, *oracle.weights("cimmino", Ao)), "lbound": 0}
    X, _, _ = air.cimmino(A, b, [3, 15], None, opts)
    Xo, _, _ = oracle.sirt("cimmino", Ao, b, [3, 15], None, opts)
    for j in range(2):
        assert rel(X[:, j], Xo[:, j]) <= TOL, (cfg, j, rel(X[:, j], Xo[:, j]))


@pytest.mark.parametrize("dtype", ["bf16", "f16", "f64"])
@pytest.mark.parametrize("cfg", ["1,1,256", "6,1,2048", "8,6,4096", "7,5,512", "10,8,4096", "9,7,512",
                                 "11,8,4096", "12,7,512", "12,10,4096", "11,9,512", "12,10,4096,16",
                                 "11,9,512,16"])
def test_tiled_shapes_narrow_storage(c1, monkeypatch, dtype, cfg):
    """Per-piece and stream forward kernels agree on every storage type."""
    opts = {"rho": c1["rho"]["sart"], "lbound": 0}
    out = []
    for f in ("1,1,256", cfg):
        monkeypatch.setenv("AIR_TILED_CFG", f)
        A = air.paralleltomo(c1["N"], c1["theta"], c1["p"], dtype=dtype)
        X, _, _ = air.sart(A, c1["b"], 20, None, opts)
        out.append(X[:, -1])
    assert rel(out[1], out[0]) <= (1e-12 if dtype == "f64" else 1e-5)


@pytest.mark.parametrize("method,smooth,dims", [("cimmino", 2, "2d"), ("sart", 1, "2d"), ("cimmino", 4, "1d")])
def test_ncp_stopping_rule(c1, oracle, method, smooth, dims):
    """NCP on the device (hipFFT of the residual columns, check_stoprules.m:96-151)
    stops at the oracle's iteration with the oracle's iterate."""
    res_dims = [c1["p"], len(c1["theta"])] if dims == "2d" else [c1["A"].shape[0]]
    opts = {"rho": c1["rho"][method], "lbound": 0,
            "stoprule": {"type": "NCP", "res_dims": res_dims, "ncp_smooth": smooth}}
    X, info, _ = air.sirt(method, c1["A"], c1["b"], [5, 200], None, opts)
    Xo, info_o, _ = oracle.sirt(method, c1["Ao"], c1["b"], [5, 200], None, opts)
    print(method, smooth, dims, info["finaliter"], info_o["finaliter"])
    assert info["stoprule"] == info_o["stoprule"] == 1
    assert info["finaliter"] == info_o["finaliter"] < 200
    assert list(info["itersaved"]) == list(info_o["itersaved"])
    assert rel(X[:, -1], Xo[:, -1]) <= TOL


def test_ncp_rejects_bad_res_dims(c1):
    with pytest.raises(Exception):
        air.cimmino(c1["A"], c1["b"], 5, None, {"stoprule": {"type": "NCP", "res_dims": [90, 180]}})


def test_train_relaxpar_search_exact_on_engine_trajectories(oracle):
    """train_relaxpar.m:43-271's search on the engine: the oracle's search driven
    by the engine's own SIRT runs (deterministic: the same call returns the same
    iterates) must return the product's lambda exactly, for the SART branch
    (rho = 1) and for Cimmino (lambda range 2/rho from the first run)."""
    N, theta, p = 32, np.arange(0.0, 180.0, 3.0), 45
    A = air.paralleltomo(N, theta, p)
    Ao = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    e = np.random.default_rng(2).standard_normal(b.size)
    b = b + 0.03 * np.linalg.norm(b) * e / np.linalg.norm(e)
    for method, opts in (("sart", None), ("cimmino", {"lbound": 0})):
        lam = air.train_relaxpar(A, b, x, method, 100, opts)

        def run(A_, b_, K_, x0_, o_, method=method):
            X, info, _ = air.sirt(method, A, b_, K_, x0_, o_)
            return X, info
        lam_o = oracle.train_relaxpar(Ao, b, x, method, 100, opts, run=run)
        print(method, lam, lam_o)
        assert abs(lam - lam_o) <= 1e-12 * abs(lam_o), (method, lam, lam_o)


@pytest.mark.parametrize("method,strategy", [("landweber", None), ("cimmino", None), ("cav", None), ("drop", None),
                                             ("sart", None), ("sart", "line"), ("cimmino", "psi2")])
def test_compressed_row_fallback_c1(c1, oracle, monkeypatch, method, strategy):
    """With the tiled layout switched off (AIR_DISABLE_TILED=1) the SIRT loop runs
    on the CSR/CSC kernels (k_resid, k_update, the line-search pass)."""
    monkeypatch.setenv("AIR_DISABLE_TILED", "1")
    A = air.paralleltomo(c1["N"], c1["theta"], c1["p"])
    assert A.layout()["tiled"] == 0
    opts = {"rho": c1["rho"][method], "lbound": 0}
    if strategy is not None:
        opts["relaxpar"] = strategy
    X, info, _ = air.sirt(method, A, c1["b"], [7, 40], None, opts)
    Xo, info_o, _ = oracle.sirt(method, c1["Ao"], c1["b"], [7, 40], None, opts)
    for j in range(2):
        assert rel(X[:, j], Xo[:, j]) <= TOL, (method, strategy, j, rel(X[:, j], Xo[:, j]))
    # the ME stop rule's partial sums on the same kernels
    if method == "cimmino" and strategy is None:
        o2 = dict(opts, stoprule={"type": "ME", "taudelta": 0.0})
        X2, info2, _ = air.sirt(method, A, c1["b"], 25, None, o2)
        X2o, info2o, _ = oracle.sirt(method, c1["Ao"], c1["b"], 25, None, o2)
        assert info2["finaliter"] == info2o["finaliter"] and rel(X2[:, -1], X2o[:, -1]) <= TOL


def test_fallback_when_a_tile_exceeds_lds(oracle):
    """A matrix whose 256-column tiles hold more distinct rows than the
    backward kernel can stage in LDS: the build declines the tiled layout by
    itself and SIRT runs on the CSR/CSC kernels."""
    import scipy.sparse as sp
    rng = np.random.default_rng(5)
    m, n = 24000, 512
    Ao = sp.random(m, n, density=0.02, random_state=rng, format="csr")
    Ao.data = rng.uniform(0.1, 1.0, Ao.nnz)
    Ao.sort_indices()
    A = air.as_operator(Ao)
    assert A.layout()["tiled"] == 0
    x = rng.uniform(0, 1, n)
    b = Ao @ x
    for method in ("sart", "cimmino"):
        M, D = oracle.weights(method, Ao)
        opts = {"rho": 1.0 if method == "sart" else oracle.spectral_radius(Ao, M, D), "lbound": 0}
        X, _, _ = air.sirt(method, A, b, 30, None, opts)
        Xo, _, _ = oracle.sirt(method, Ao, b, 30, None, opts)
        assert rel(X[:, -1], Xo[:, -1]) <= TOL, method


def test_deferred_norms_equal_direct_finalize(c1, oracle):
    """With the stop rule 'none' the tiled loop defers the residual norms (block
    partials kept in a ring of 32 iterations and reduced afterwards).  The run
    in uneven chunks that cross the ring's wrap must give the iterates of the
    same run with a per-iteration finalize (DP with taudelta 0, which never
    fires), and the oracle's residual history."""
    from airtoolsii_amd.sirt import SirtRun
    opts = {"rho": c1["rho"]["cimmino"], "lbound": 0}
    K = 90
    run = SirtRun("cimmino", c1["A"], c1["b"], K, None, opts)
    for n in (7, 40, 1, 25, 30):                       # 103 > K: the last call stops at K
        run.iterate(n)
    x1, info1 = run.fetch()
    run2 = SirtRun("cimmino", c1["A"], c1["b"], K, None, dict(opts, stoprule={"type": "DP", "taudelta": 0.0}))
    run2.iterate(K)
    x2, info2 = run2.fetch()
    assert info1["finaliter"] == info2["finaliter"] == K and info1["stoprule"] == info2["stoprule"] == 0
    assert np.array_equal(x1, x2)
    np.testing.assert_allclose(info1["resnorm"], info2["resnorm"], rtol=1e-12)
    _, _, _, rn = oracle.sirt("cimmino", c1["Ao"], c1["b"], K, None, opts, return_resnorm=True)
    np.testing.assert_allclose(info1["resnorm"], rn[: K + 1], rtol=1e-5)
    # the whole call (save list, chunked loop) on the deferred path
    X, info, _ = air.cimmino(c1["A"], c1["b"], [5, 33, 64, K], None, opts)
    assert info["finaliter"] == K and np.array_equal(X[:, -1], x1)
    np.testing.assert_allclose(info["resnorm"], info1["resnorm"], rtol=1e-12)


@pytest.mark.parametrize("geom", ["parallel", "fanlinear", "fancurved"])
def test_step_coded_forward_bit_identical_to_lane_form(oracle, monkeypatch, geom):
    """The step-coded forward decodes each lane's 8 tile positions from one
    word; the lane form reads them from 16-bit codes.  With the same wave
    partition (entries per wave) and the same backward shape, both sum the
    same products in the same order, so the iterates must agree bit for bit
    (a single mis-decoded position would change them), and so must the step
    form with and without the next round prefetched, in either direction."""
    N = 80
    if geom == "fanlinear":
        theta = np.arange(0.0, 360.0, 4.0)
        make = lambda: air.fanlineartomo(N, theta, 113)
        Ao = oracle.fanlineartomo(N, theta, 113)
    elif geom == "fancurved":
        theta = np.arange(0.0, 360.0, 4.0)
        make = lambda: air.fancurvedtomo(N, theta, 113)
        Ao = oracle.fancurvedtomo(N, theta, 113)
    else:
        theta = np.arange(0.0, 180.0, 2.0)
        make = lambda: air.paralleltomo(N, theta, 113)
        Ao = oracle.paralleltomo(N, theta, 113)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    opts = {"rho": oracle.spectral_radius(Ao, *oracle.weights("cimmino", Ao)), "lbound": 0}
    out = {}
    for cfg in ("9,7,2048", "11,7,2048", "10,7,4096", "12,7,4096", "11,7,4096", "11,9,4096", "11,10,4096"):
        monkeypatch.setenv("AIR_TILED_CFG", cfg)
        A = make()
        L = A.layout()
        assert L["fwd_step_ok"] == 1 and L["fwd_step"] == (1 if int(cfg.split(",")[0]) >= 11 else 0)
        X, _, _ = air.cimmino(A, b, [1, 6], None, opts)
        out[cfg] = X
    assert np.array_equal(out["9,7,2048"], out["11,7,2048"]), geom          # lane form vs step form
    assert np.array_equal(out["10,7,4096"], out["12,7,4096"]), geom         # the same, next round prefetched
    assert np.array_equal(out["11,7,4096"], out["12,7,4096"]), geom         # step form without / with prefetch
    assert np.array_equal(out["11,9,4096"], out["11,10,4096"]), geom        # step backward without / with prefetch
    Xo, _, _ = oracle.sirt("cimmino", Ao, b, [1, 6], None, opts)
    assert rel(out["12,7,4096"][:, 1], Xo[:, 1]) <= TOL


@pytest.mark.parametrize("geom", ["fanlinear", "fancurved", "seismic"])
def test_step_coded_forward_other_geometries(oracle, monkeypatch, geom):
    """The step-coded forward and backward on other traced geometries: straight
    rays from a point source cross every tile as unit-step staircases (both y
    directions, corner crossings), and every pixel's rays sweep through the
    runs of its backward tile; a geometry the build cannot encode keeps the
    column codes.  Either way the iterates match the oracle's."""
    monkeypatch.setenv("AIR_TILED_CFG", "12,10,4096")
    N = 80
    if geom == "fanlinear":
        theta = np.arange(0.0, 360.0, 4.0)
        A, Ao = air.fanlineartomo(N, theta, 113), oracle.fanlineartomo(N, theta, 113)
    elif geom == "fancurved":
        theta = np.arange(0.0, 360.0, 4.0)
        A, Ao = air.fancurvedtomo(N, theta, 113), oracle.fancurvedtomo(N, theta, 113)
    else:
        A, Ao = air.seismictomo(N, 60, 90), oracle.seismictomo(N, 60, 90)
    L = A.layout()
    assert L["tiled"] == 1
    if geom != "seismic":
        assert L["fwd_step_ok"] == 1 and L["fwd_step"] == 1
    if geom == "fanlinear":
        assert L["bwd_step_ok"] == 1 and L["bwd_step"] == 1
    assert L["bwd_step"] == L["bwd_step_ok"]
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = Ao @ x
    opts = {"rho": oracle.spectral_radius(Ao, *oracle.weights("sart", Ao)), "lbound": 0}
    X, _, _ = air.sart(A, b, [4, 20], None, opts)
    Xo, _, _ = oracle.sirt("sart", Ao, b, [4, 20], None, opts)
    for j in range(2):
        assert rel(X[:, j], Xo[:, j]) <= TOL, (geom, j, rel(X[:, j], Xo[:, j]))


@pytest.mark.parametrize("stride", [1, 4])
def test_sampled_kernel_timing(c1, stride):
    """Event timing with stride s brackets the kernels of iterations 0, s, 2s, ...
    counted over the calls since it was enabled; the events change no iterate."""
    from airtoolsii_amd.sirt import SirtRun
    opts = {"rho": c1["rho"]["cimmino"], "lbound": 0}
    run = SirtRun("cimmino", c1["A"], c1["b"], 40, None, opts)
    run.iterate(3)
    run.kernel_timing(stride)
    for n in (10, 7, 13):
        run.iterate(n)
    kt = run.kernel_times()
    assert kt["launches"] == -(-30 // stride)
    assert kt["cols_ms"] > 0 and kt["rows_ms"] > 0 and kt["reduce_ms"] > 0
    run.kernel_timing(False)
    run.iterate(10)
    assert run.kernel_times()["launches"] == 0
    x1, info1 = run.fetch()
    ref = SirtRun("cimmino", c1["A"], c1["b"], 40, None, opts)
    ref.iterate(40)
    x2, _ = ref.fetch()
    assert info1["finaliter"] == 40 and np.array_equal(x1, x2)
    from airtoolsii_amd._lib import AirError
    with pytest.raises(AirError):
        run.kernel_timing(-1)


@pytest.mark.parametrize("dtype,tol", [("f32", 1e-5), ("bf16", 2e-2)])
def test_step_coded_backward_wide_staging(oracle, monkeypatch, dtype, tol):
    """Fan beams over 720 angles stage up to 720 runs of 28 rows per 16 x 16
    tile: more than 16384 slots (the lane words address them with 15 bits)
    and more than 76 KB of LDS (16 waves per tile, one block per CU).  SART
    on the step-coded backward with that staging (the update pass, and the
    line search's norm pass) gives the oracle's iterates, as the lane form does."""
    N, theta, p, K = 128, np.arange(0.0, 360.0, 0.5), 181, 20
    Ao = oracle.fanlineartomo(N, theta, p)
    x = oracle.phantomgallery("threephases", N, None, 3).ravel(order="F")
    b = Ao @ x
    opts = {"rho": 1.0, "lbound": 0}
    Xo, _, _ = oracle.sirt("sart", Ao, b, K, None, opts)
    out = {}
    for cfg in ("11,10,4096,16", "11,8,4096,16"):          # step-coded backward, then the lane form
        monkeypatch.setenv("AIR_TILED_CFG", cfg)
        A = air.fanlineartomo(N, theta, p, dtype=dtype)
        L = A.layout()
        if cfg.startswith("11,10"):
            assert L["bwd_step"] == 1 and L["bwd_step_base_bits"] == 15, L
            assert L["bwd_step_lds_bytes"] > 76 * 1024, L
        else:
            assert L["bwd_step"] == 0, L
        X, _, _ = air.sart(A, b, K, None, opts)
        out[cfg] = X[:, -1]
        del A
    for cfg, xk in out.items():
        print(cfg, rel(xk, Xo[:, -1]))
        assert rel(xk, Xo[:, -1]) <= tol, cfg
    if dtype == "f32":
        # the line search's pass (the backward's norm mode) on the wide staging
        monkeypatch.setenv("AIR_TILED_CFG", "11,10,4096,16")
        A = air.fanlineartomo(N, theta, p, dtype=dtype)
        lopts = {"relaxpar": "line", "rho": 1.0, "lbound": 0}
        X, info, _ = air.sart(A, b, 10, None, lopts)
        Xo2, info_o, _ = oracle.sirt("sart", Ao, b, 10, None, lopts)
        assert rel(X[:, -1], Xo2[:, -1]) <= tol
        assert abs(info["relaxpar"] - info_o["relaxpar"]) <= 1e-4 * abs(info_o["relaxpar"])
