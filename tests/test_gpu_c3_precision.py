"""Precision study at the fan-beam scale: fanlineartomo N=1024, 720 angles over
360 degrees, p=1448 (R=2N, dw=2.5N, sd=3N), threephases(1024) phantom, 1%
seeded noise, SART K=100.  A's values are stored in FP32, BF16 and FP16 (FP32
accumulation; FP64 weights and norms) and each reconstruction is compared with
the FP64 oracle on identical inputs.  The engine's FP64 trace of the same
geometry is compared with the oracle's CSR first: row pointers, column indices
and values bit for bit (fanlineartomo.m:141-358).

The oracle side needs about 40 GB of host memory and about half a minute of
CPU time."""
import json
import os
import time

import numpy as np
import pytest

import airtoolsii_amd as air

pytestmark = pytest.mark.gpu

# relative 2-norm error of the K=100 iterate against the FP64 oracle, per storage type
TOL = {"f32": 1e-5, "f16": 5e-3, "bf16": 3e-2}


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_fanlinear_n1024_precision(oracle):
    N, theta, p, K = 1024, np.arange(0.0, 360.0, 0.5), 1448, 100
    t0 = time.perf_counter()
    Ao = oracle.fanlineartomo(N, theta, p)
    t_trace = time.perf_counter() - t0
    print(f"C3: oracle CSR m={Ao.shape[0]} nnz={Ao.nnz} in {t_trace:.1f} s", flush=True)
    # the full 720-angle matrix, traced on the GPU with FP64 values
    A64 = air.fanlineartomo(N, theta, p, dtype="f64")
    rp, ci, v = A64.csr_arrays()
    del A64
    assert np.array_equal(rp, Ao.indptr), "row pointers differ"
    assert np.array_equal(ci, Ao.indices), "column indices differ"
    assert np.array_equal(v, Ao.data), "values differ"
    del rp, ci, v
    print("C3: CSR bit-exact", flush=True)
    x_true = oracle.phantomgallery("threephases", N, None, 1).ravel(order="F")
    b0 = Ao @ x_true
    e = np.random.default_rng(1).standard_normal(b0.size)
    b = b0 + 0.01 * np.linalg.norm(b0) * e / np.linalg.norm(e)
    opts = {"lbound": 0}
    t0 = time.perf_counter()
    Xo, info_o, _ = oracle.sirt("sart", Ao, b, K, None, opts, nthreads=min(os.cpu_count() or 1, 16))
    t_oracle = time.perf_counter() - t0
    xo = Xo[:, -1]
    out = {"workload": f"fanlineartomo N={N}, theta=0:0.5:359.5 ({len(theta)} angles), p={p}, SART K={K}, lbound 0",
           "m": Ao.shape[0], "n": Ao.shape[1], "nnz": int(Ao.nnz), "oracle_trace_s": round(t_trace, 1),
           "oracle_sirt_s": round(t_oracle, 1), "oracle_rel_err_to_phantom": rel(xo, x_true), "dtypes": {}}
    del Ao
    for dt in ("f32", "bf16", "f16"):
        A = air.fanlineartomo(N, theta, p, dtype=dt)
        assert A.nnz == out["nnz"]
        L = A.layout()
        run = air.SirtRun("sart", A, b, K, None, opts)
        t0 = time.perf_counter()
        run.iterate(K)
        t_iter = time.perf_counter() - t0
        x, info = run.fetch()
        err = rel(x, xo)
        out["dtypes"][dt] = {"rel_err_to_oracle": err, "rel_err_to_phantom": rel(x, x_true),
                             "iter_per_s": round(K / t_iter, 1), "tiled": L["tiled"], "lds_bytes": L["lds_bytes"],
                             "bwd_step": L["bwd_step"], "bwd_step_base_bits": L["bwd_step_base_bits"]}
        del run, A
    print("C3_PRECISION " + json.dumps(out), flush=True)
    for dt, r in out["dtypes"].items():
        assert r["rel_err_to_oracle"] <= TOL[dt], (dt, r)
