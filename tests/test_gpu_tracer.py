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
}
FAN_CASES = {
    "fan_N64_even_p": dict(N=64, theta=np.arange(0.0, 360.0, 2.0), p=90),
    "fan_N31_odd_p": dict(N=31, theta=np.arange(0.0, 360.0, 11.0), p=45),
    "fan_c3_angle_subset": dict(N=1024, theta=np.arange(0.0, 360.0, 0.5)[::48], p=1448),
}


def build_pair(oracle, kind, cfg, dtype):
    if kind == "par":
        A = air.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"], dtype=dtype)
        Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"])
    else:
        A = air.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"], dtype=dtype)
        Ao = oracle.fanlineartomo(cfg["N"], cfg["theta"], cfg["p"])
    return A, Ao


CASES = [("par", k) for k in PAR_CASES] + [("fan", k) for k in FAN_CASES]


@pytest.mark.parametrize("kind,name", CASES)
def test_csr_bit_exact_fp64(oracle, kind, name):
    cfg = (PAR_CASES if kind == "par" else FAN_CASES)[name]
    print(f"case {kind}/{name}", flush=True)
    A, Ao = build_pair(oracle, kind, cfg, "f64")
    rp, ci, v = A.csr_arrays()
    assert A.shape == Ao.shape and A.nnz == Ao.nnz
    assert np.array_equal(rp, Ao.indptr)
    assert np.array_equal(ci, Ao.indices)
    assert np.array_equal(v.view(np.uint64), Ao.data.view(np.uint64)), \
        f"{np.count_nonzero(v != Ao.data)} FP64 values differ"


@pytest.mark.parametrize("dtype", ["f32", "bf16", "f16"])
@pytest.mark.parametrize("kind,name", [("par", "c1_N64_180ang_p91"), ("fan", "fan_N64_even_p")])
def test_csr_narrowed_values(oracle, kind, name, dtype):
    cfg = (PAR_CASES if kind == "par" else FAN_CASES)[name]
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


@pytest.mark.parametrize("kind", ["par", "fan"])
def test_matrix_free_products_match_matrix(oracle, kind):
    cfg = PAR_CASES["c1_N64_180ang_p91"] if kind == "par" else FAN_CASES["fan_N64_even_p"]
    if kind == "par":
        Af = air.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"], 0, 0)
        Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"])
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


def test_matlab_csc_upload_matches_tracer(oracle):
    cfg = PAR_CASES["oddN_evenp_fullcircle"]
    Ao = oracle.paralleltomo(cfg["N"], cfg["theta"], cfg["p"], cfg["d"]).tocsc()
    A = air.from_matlab_csc(Ao.shape[0], Ao.shape[1], Ao.indptr.astype(np.int64), Ao.indices.astype(np.int64),
                            Ao.data, dtype="f64")
    rp, ci, v = A.csr_arrays()
    Ar = Ao.tocsr()
    Ar.sort_indices()
    assert np.array_equal(rp, Ar.indptr) and np.array_equal(ci, Ar.indices) and np.array_equal(v, Ar.data)
