"""GPU ray tracer (K1) against the FP64 oracle: CSR row pointers and column
indices bit-exact, FP64 values bit-exact, narrowed values equal to the
narrowing of the oracle's values; matrix-free products against the matrix."""
import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu


def bf16_round(v64):
    f = np.asarray(v64, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


PAR_CASES = {
    "c1_N64_180ang_p91": dict(N=64, theta=np.arange(0.0, 180.0), p=91, d=None),
    "oddN_evenp_fullcircle": dict(N=33, theta=np.arange(0.0, 360.0, 7.5), p=40, d=50.0),
    "grid_aligned_and_corners": dict(N=8, theta=np.arange(0.0, 360.0, 45.0), p=9, d=None),
    "rays_missing_grid": dict(N=16, theta=np.random.default_rng(4).uniform(0, 360, 40), p=23, d=60.0),
    "tiny_N1": dict(N=1, theta=np.array([0.0, 30.0, 90.0]), p=3, d=1.0),
    "c2_angle_subset": dict(N=512, theta=np.arange(0.0, 180.0, 0.5)[::24], p=725, d=None),
    "N100_partial_tiles": dict(N=100, theta=np.arange(0.0, 180.0, 7.0), p=141, d=None),
    "near_axis_angles": dict(N=32, theta=np.array([1e-3, 0.5, 44.9999999, 45.0, 89.999, 90.0005, 134.99999,
                                                   179.9999, 270.00001, 359.999]), p=47, d=None),
    "two_rays": dict(N=12, theta=np.arange(0.0, 180.0, 10.0), p=2, d=3.0),
}
FAN_CASES = {
    "fan_N64_even_p": dict(N=64, theta=np.arange(0.0, 360.0, 2.0), p=90),
    "fan_N31_odd_p": dict(N=31, theta=np.arange(0.0, 360.0, 11.0), p=45),
    "fan_c3_angle_subset": dict(N=1024, theta=np.arange(0.0, 360.0, 0.5)[::48], p=1448),
}
CURVED_CASES = {
    "curved_N64_defaults": dict(N=64, theta=np.arange(0.0, 359.0, 2.0), p=90, R=None, d=None),
    "curved_N31_R1p5_d40": dict(N=31, theta=np.arange(0.0, 360.0, 11.0), p=45, R=1.5, d=40.0),
    "curved_N16_full_half_plane": dict(N=16, theta=np.arange(0.0, 360.0, 45.0), p=31, R=2.0, d=180.0),
    "curved_N512_angle_subset": dict(N=512, theta=np.arange(0.0, 360.0, 1.0)[::24], p=725, R=None, d=None),
}
SEISMIC_CASES = {
    "seismic_N64_defaults": dict(N=64, s=None, p=None),
    "seismic_N33_s7_p9": dict(N=33, s=7, p=9),
    "seismic_N8_horizontal_rays": dict(N=8, s=8, p=16),
    "seismic_N16_one_receiver": dict(N=16, s=5, p=1),
    "seismic_N256_s32_p512": dict(N=256, s=32, p=512),
}


def build_pair(oracle, kind, cfg, dtype, is_matrix=1):
    if kind == "par":
        A = air.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"], 0, is_matrix, dtype=dtype)
        Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"])
    elif kind == "fan":
        A = air.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"], None, None, None, 0, is_matrix, dtype=dtype)
        Ao = oracle.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"])
    elif kind == "seismic":
        A = air.seismictomo(cfg["N"], cfg["s"], cfg["p"], 0, is_matrix, dtype=dtype)
        Ao = oracle.seismictomo(cfg["N"], cfg["s"], cfg["p"])
    else:
        R = 2.0 if cfg["R"] is None else cfg["R"]
        A = air.fancurvedtomo(cfg["N"], cfg["theta"], cfg["p"], cfg["R"], cfg["d"], 0, is_matrix, dtype=dtype)
        Ao = oracle.fancurvedtomo(cfg["N"], cfg["theta"], cfg["p"], R, cfg["d"])
    return A, Ao


TABLES = {"par": PAR_CASES, "fan": FAN_CASES, "curved": CURVED_CASES, "seismic": SEISMIC_CASES}
CASES = [(kind, k) for kind, tab in TABLES.items() for k in tab]


@pytest.mark.parametrize("kind,name", CASES)
def test_csr_bit_exact_fp64(oracle, kind, name):
    cfg = TABLES[kind][name]
    print(f"case {kind}/{name}", flush=True)
    A, Ao = build_pair(oracle, kind, cfg, "f64")
    rp, ci, v = A.csr_arrays()
    assert A.shape == Ao.shape and A.nnz == Ao.nnz
    assert np.array_equal(rp, Ao.indptr)
    assert np.array_equal(ci, Ao.indices)
    assert np.array_equal(v.view(np.uint64), Ao.data.view(np.uint64)), \
        f"{np.count_nonzero(v != Ao.data)} FP64 values differ"


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("kind,name", [("par", "c1_N64_180ang_p91"), ("fan", "fan_N64_even_p"),
                                       ("curved", "curved_N64_defaults"), ("seismic", "seismic_N64_defaults")])
def test_csr_narrowed_values(oracle, kind, name, dtype):
    cfg = TABLES[kind][name]
    A, Ao = build_pair(oracle, kind, cfg, dtype)
    rp, ci, v = A.csr_arrays()
    assert np.array_equal(rp, Ao.indptr) and np.array_equal(ci, Ao.indices)
    if dtype == "f32":
        ref = Ao.data.astype(np.float32).astype(np.float64)
    elif dtype == "bf16":
        ref = bf16_round(Ao.data)
    else:
        ref = Ao.data.astype(np.float32).astype(np.float16).astype(np.float64)
    assert np.array_equal(v, ref)


def test_c2_full_structure_and_row_sums(oracle):
    """The headline matrix: paralleltomo(512, 0:0.5:179.5, 725), every entry."""
    theta = np.arange(0.0, 180.0, 0.5)
    A = air.paralleltomo(512, theta, 725, dtype="f64")
    Ao = oracle.paralleltomo(512, theta, 725)
    rp, ci, v = A.csr_arrays()
    assert A.nnz == Ao.nnz == 120155712
    assert np.array_equal(rp, Ao.indptr)
    assert np.array_equal(ci, Ao.indices)
    assert np.array_equal(v, Ao.data)


def test_transpose_and_products(oracle):
    cfg = PAR_CASES["oddN_evenp_fullcircle"]
    A, Ao = build_pair(oracle, "par", cfg, "f64")
    rng = np.random.default_rng(0)
    x = rng.standard_normal(A.shape[1])
    y = rng.standard_normal(A.shape[0])
    np.testing.assert_allclose(A(x, "notransp"), Ao @ x, rtol=2e-6, atol=2e-6 * np.abs(Ao @ x).max())
    np.testing.assert_allclose(A(y, "transp"), Ao.T @ y, rtol=2e-6, atol=2e-6 * np.abs(Ao.T @ y).max())
    assert A(None, "size") == Ao.shape
    with pytest.raises(ValueError):
        A(x, "bogus")


@pytest.mark.parametrize("kind", ["par", "fan", "curved", "seismic"])
def test_matrix_free_products_match_matrix(oracle, kind):
    cfg = {"par": PAR_CASES["c1_N64_180ang_p91"], "fan": FAN_CASES["fan_N64_even_p"],
           "curved": CURVED_CASES["curved_N64_defaults"], "seismic": SEISMIC_CASES["seismic_N64_defaults"]}[kind]
    if kind == "par":
        Af = air.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"], 0, 0)
        Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"])
    elif kind == "curved":
        Af = air.fancurvedtomo(cfg["N"], cfg["theta"], cfg["p"], None, None, 0, 0)
        Ao = oracle.fancurvedtomo(cfg["N"], cfg["theta"], cfg["p"])
    elif kind == "seismic":
        Af = air.seismictomo(cfg["N"], cfg["s"], cfg["p"], 0, 0)
        Ao = oracle.seismictomo(cfg["N"], cfg["s"], cfg["p"])
    else:
        Af = air.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"], None, None, None, 0, 0)
        Ao = oracle.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"])
    assert Af.matrix_free and Af.shape == Ao.shape
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 1, Ao.shape[1])
    y = rng.uniform(0, 1, Ao.shape[0])
    fx, fy = Ao @ x, Ao.T @ y
    assert np.linalg.norm(Af @ x - fx) <= 1e-6 * np.linalg.norm(fx)
    assert np.linalg.norm(Af.T @ y - fy) <= 1e-6 * np.linalg.norm(fy)
    # adjoint identity <A x, y> = <x, A' y> (demos/demo_matrixfree.m:97-138)
    assert abs((Af @ x) @ y - x @ (Af.T @ y)) <= 1e-5 * abs(fx @ y)


@pytest.mark.parametrize("walk", ["dda", "dda_tile_adjoint", "exact"])
@pytest.mark.parametrize("kind,name", CASES)
def test_matrix_free_tiles_reproduce_every_entry(oracle, monkeypatch, kind, name, walk):
    """The tile-major projector walks each ray once per image tile it passes
    and keeps the segments whose pixel lies in the tile: every entry of A must
    come out exactly once.  The crossing walk (AIR_MF_EXACT=1) forms the
    tracer's values bit for bit; the DDA walk forms them to rounding and drops
    segments below the tracer's 1e-10 merge tolerance, so its bar adds
    1e-9 max|x| (a segment of that size is the only place the two rules can
    differ).  For parallel beams the DDA's adjoint is pixel-driven (each pixel
    clips its candidate rays with the DDA's expressions), so its bar adds
    1e-9 max|y| likewise; AIR_MF_PIXEL=0 keeps the tile adjoint with the
    crossing walk.  The adjoint sums in FP64, so each pixel agrees with A32' y
    (A's values rounded to FP32) to the FP32 rounding of the result; a lost or
    doubled entry would be off by the size of an entry.  The forward sums in
    FP32 in another order than a row of A."""
    print(f"matrix-free tiles {kind} {name} {walk}", flush=True)
    monkeypatch.setenv("AIR_MF_EXACT", "1" if walk == "exact" else "0")
    monkeypatch.setenv("AIR_MF_PIXEL", "0" if walk == "dda_tile_adjoint" else "1")
    Af, Ao = build_pair(oracle, kind, TABLES[kind][name], "f32", is_matrix=0)
    assert Af.matrix_free and Af.shape == Ao.shape
    A32 = Ao.copy()
    A32.data = A32.data.astype(np.float32).astype(np.float64)
    Aabs = abs(Ao)
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, Ao.shape[1])
    y = rng.uniform(-1, 1, Ao.shape[0])
    x32 = x.astype(np.float32).astype(np.float64)
    y32 = y.astype(np.float32).astype(np.float64)
    floor_f = 1e-30 if walk == "exact" else 1e-9 * np.abs(x32).max()
    floor_t = 1e-9 * np.abs(y32).max() if walk == "dda" else 1e-30
    rel_t = 2e-7
    err_f = np.abs(Af @ x - A32 @ x32)
    assert np.all(err_f <= 1e-5 * (Aabs @ np.abs(x32)) + floor_f), err_f.max()
    err_t = np.abs(Af.T @ y - A32.T @ y32)
    assert np.all(err_t <= rel_t * (Aabs.T @ np.abs(y32)) + floor_t), err_t.max()


def test_matlab_csc_upload_matches_tracer(oracle):
    cfg = PAR_CASES["oddN_evenp_fullcircle"]
    Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"]).tocsc()
    A = air.from_matlab_csc(Ao.shape[0], Ao.shape[1], Ao.indptr.astype(np.int64), Ao.indices.astype(np.int64),
                            Ao.data, dtype="f64")
    rp, ci, v = A.csr_arrays()
    Ar = Ao.tocsr()
    Ar.sort_indices()
    assert np.array_equal(rp, Ar.indptr) and np.array_equal(ci, Ar.indices) and np.array_equal(v, Ar.data)


def test_curved_return_all_and_errors():
    A, b, x, theta, p, R, d = air.fancurvedtomo(32, return_all=True)
    assert A.shape == (180 * 45, 32 * 32) and p == 45 and R == 2.0
    assert abs(d - 2 * np.degrees(np.arctan(1 / 3))) < 1e-12
    np.testing.assert_allclose(b, A @ x, rtol=1e-6)
    with pytest.raises(ValueError, match="half squareroot"):
        air.fancurvedtomo(8, [0.0], 5, 0.6)
    with pytest.raises(ValueError, match="interval"):
        air.fancurvedtomo(8, [0.0], 5, 2.0, 200.0)


def test_seismic_return_all(oracle):
    A, b, x, s, p = air.seismictomo(40, return_all=True)
    assert A.shape == (40 * 80, 1600) and (s, p) == (40, 80)
    np.testing.assert_array_equal(x, oracle.tectonic(40).ravel(order="F"))
    np.testing.assert_allclose(b, oracle.seismictomo(40) @ x, rtol=1e-6)


SPHERICAL_CASES = {
    "spherical_N16_nC23": dict(N=16, angles=np.arange(0.0, 360.0, 37.0), nC=23),
    "spherical_N33_odd": dict(N=33, angles=np.arange(0.0, 360.0, 10.0), nC=20),
    "spherical_N64_defaults": dict(N=64, angles=None, nC=None),
    "spherical_N256_angle_subset": dict(N=256, angles=np.arange(0.0, 360.0, 2.0)[::15], nC=None),
}


@pytest.mark.parametrize("name", list(SPHERICAL_CASES))
def test_spherical_csr_bit_exact_fp64(oracle, name):
    cfg = SPHERICAL_CASES[name]
    print(f"case spherical/{name}", flush=True)
    A = air.sphericaltomo(cfg["N"], cfg["angles"], cfg["nC"], dtype="f64")
    Ao = oracle.sphericaltomo(cfg["N"], cfg["angles"], cfg["nC"])
    rp, ci, v = A.csr_arrays()
    assert A.shape == Ao.shape and A.nnz == Ao.nnz
    assert np.array_equal(rp, Ao.indptr)
    assert np.array_equal(ci, Ao.indices)
    assert np.array_equal(v.view(np.uint64), Ao.data.view(np.uint64))


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_spherical_narrowed_values_and_products(oracle, dtype):
    A = air.sphericaltomo(48, np.arange(0.0, 360.0, 5.0), 30, dtype=dtype)
    Ao = oracle.sphericaltomo(48, np.arange(0.0, 360.0, 5.0), 30)
    rp, ci, v = A.csr_arrays()
    assert np.array_equal(rp, Ao.indptr) and np.array_equal(ci, Ao.indices)
    ref = Ao.data.astype(np.float32).astype(np.float64) if dtype == "f32" else bf16_round(Ao.data)
    assert np.array_equal(v, ref)
    x = np.random.default_rng(2).uniform(0, 1, Ao.shape[1])
    tol = 1e-6 if dtype == "f32" else 1e-2
    assert np.linalg.norm(A @ x - Ao @ x) <= tol * np.linalg.norm(Ao @ x)


def test_spherical_return_all():
    A, b, x, angles, nC = air.sphericaltomo(24, return_all=True)
    assert A.shape == (180 * 34, 576) and nC == 34 and len(angles) == 180
    np.testing.assert_allclose(b, A @ x, rtol=1e-6)
    Af, bf, xf, _, _ = air.sphericaltomo(24, None, None, 0, 0, return_all=True)
    assert Af.matrix_free and Af.shape == A.shape
    assert np.array_equal(xf, x)
    np.testing.assert_allclose(bf, b, rtol=1e-5, atol=1e-6 * np.abs(b).max())


@pytest.mark.parametrize("N,angles,nC", [(16, np.arange(0.0, 360.0, 37.0), 23), (33, np.arange(0.0, 360.0, 10.0), 20),
                                         (64, None, None)])
def test_spherical_matrix_free_products(oracle, N, angles, nC):
    """sphericaltomo(..., isMatrix=0) re-forms the circle samples on every
    product in FP64 (the reference's function handle, sphericaltomo.m:243-254):
    A x and A' y equal the FP64 matrix's products up to the FP32 rounding of
    the result, in every row and every pixel (including the samples outside
    the image, rows that miss it and circles that leave and re-enter a pixel)."""
    print(f"case spherical matrix-free N={N}", flush=True)
    Af = air.sphericaltomo(N, angles, nC, 0, 0)
    Ao = oracle.sphericaltomo(N, angles, Af.geometry["numCircles"] if nC is None else nC).tocsr()
    assert Af.matrix_free and Af.shape == Ao.shape
    Aabs = abs(Ao)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, Ao.shape[1]).astype(np.float32).astype(np.float64)
    y = rng.uniform(-1, 1, Ao.shape[0]).astype(np.float32).astype(np.float64)
    assert np.all(np.abs(Af @ x - Ao @ x) <= 2e-7 * (Aabs @ np.abs(x)) + 1e-30)
    assert np.all(np.abs(Af.T @ y - Ao.T @ y) <= 2e-7 * (Aabs.T @ np.abs(y)) + 1e-30)


@pytest.mark.parametrize("method", ["cimmino", "cav", "drop", "sart"])
def test_spherical_matrix_free_sirt_matches_matrix(method):
    """Weights of the matrix-free operator (row and column statistics of the
    binned circle samples) equal the stored matrix's; 20 SIRT iterations on
    both operators agree."""
    A = air.sphericaltomo(32, np.arange(0.0, 360.0, 4.0), 40)
    Af = air.sphericaltomo(32, np.arange(0.0, 360.0, 4.0), 40, 0, 0)
    x = air.phantomgallery("shepplogan", 32).ravel(order="F")
    b = A @ x
    opts = {"lbound": 0}
    X, info, _ = air.sirt(method, A, b, 20, None, opts)
    Xf, _, _ = air.sirt(method, Af, b, 20, None, dict(opts, rho=info["rho"]))
    assert np.linalg.norm(Xf[:, -1] - X[:, -1]) <= 1e-5 * np.linalg.norm(X[:, -1])


WAVE_CASES = {
    "wave_N16_defaults": dict(N=16, s=None, p=None, omega=None),
    "wave_N33_s7_p9_omega5": dict(N=33, s=7, p=9, omega=5.0),
    "wave_N64_defaults": dict(N=64, s=None, p=None, omega=None),
    "wave_N128_s16_p64": dict(N=128, s=16, p=64, omega=20.0),
}


@pytest.mark.parametrize("name", list(WAVE_CASES))
def test_seismicwave_csr_matches_oracle(oracle, name):
    """Columns equal and FP64 values within 1e-12 relative: cos and exp on the
    device may differ from the C library's by an ulp (the threshold 1e-6 is not
    met within such a margin by any entry of these cases)."""
    cfg = WAVE_CASES[name]
    print(f"case seismicwave/{name}", flush=True)
    A = air.seismicwavetomo(cfg["N"], cfg["s"], cfg["p"], cfg["omega"], dtype="f64")
    Ao = oracle.seismicwavetomo(cfg["N"], cfg["s"], cfg["p"], 10.0 if cfg["omega"] is None else cfg["omega"])
    rp, ci, v = A.csr_arrays()
    assert A.shape == Ao.shape and A.nnz == Ao.nnz
    assert np.array_equal(rp, Ao.indptr)
    assert np.array_equal(ci, Ao.indices)
    np.testing.assert_allclose(v, Ao.data, rtol=1e-12, atol=0)


def test_seismicwave_f32_and_return_all(oracle):
    A, b, x, s, p, omega = air.seismicwavetomo(24, return_all=True)
    Ao = oracle.seismicwavetomo(24)
    assert (s, p, omega) == (24, 48, 10.0) and A.shape == Ao.shape
    rp, ci, v = A.csr_arrays()
    assert np.array_equal(ci, Ao.indices)
    np.testing.assert_allclose(v, Ao.data.astype(np.float32), rtol=1e-6)
    np.testing.assert_array_equal(x, oracle.tectonic(24).ravel(order="F"))
    np.testing.assert_allclose(b, Ao @ x, rtol=1e-5)


@pytest.mark.parametrize("N,s,p,omega", [(24, None, None, None), (33, 5, 9, 20.0), (40, 8, 13, 4.0)])
def test_seismicwave_matrix_free_products(oracle, N, s, p, omega):
    """seismicwavetomo(..., isMatrix=0) re-evaluates the Fresnel kernels on
    every product: its entries are the stored matrix's, so A x agrees with
    A32 x to FP32 summation and A' y with A32' y to the result's FP32
    rounding in every pixel (FP64 sums over the rows in order)."""
    print(f"case seismicwave matrix-free N={N} s={s} p={p}", flush=True)
    Af = air.seismicwavetomo(N, s, p, omega, 0, 0)
    Ao = oracle.seismicwavetomo(N, s, p, 10.0 if omega is None else omega)
    assert Af.matrix_free and Af.shape == Ao.shape
    A32 = Ao.copy()
    A32.data = A32.data.astype(np.float32).astype(np.float64)
    Aabs = abs(Ao)
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 1, Ao.shape[1]).astype(np.float32).astype(np.float64)
    y = rng.uniform(-1, 1, Ao.shape[0]).astype(np.float32).astype(np.float64)
    assert np.all(np.abs(Af @ x - A32 @ x) <= 1e-5 * (Aabs @ np.abs(x)) + 1e-30)
    assert np.all(np.abs(Af.T @ y - A32.T @ y) <= 2e-7 * (Aabs.T @ np.abs(y)) + 1e-30)


@pytest.mark.parametrize("method", ["cimmino", "cav", "drop", "sart"])
def test_seismicwave_matrix_free_sirt_matches_matrix(method):
    """Weights from the matrix-free row and column statistics, then 20 SIRT
    iterations: the iterates of the stored and the matrix-free operator agree."""
    A = air.seismicwavetomo(24)
    Af = air.seismicwavetomo(24, None, None, None, 0, 0)
    x = air.phantomgallery("tectonic", 24).ravel(order="F")
    b = A @ x
    opts = {"lbound": 0}
    X, info, _ = air.sirt(method, A, b, 20, None, opts)
    Xf, _, _ = air.sirt(method, Af, b, 20, None, dict(opts, rho=info["rho"]))
    assert np.linalg.norm(Xf[:, -1] - X[:, -1]) <= 1e-5 * np.linalg.norm(X[:, -1])
