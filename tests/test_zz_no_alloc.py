"""air_apply_device, the entry point that allocates nothing: its products on
device memory against air_apply's on the same operators and the same FP32
inputs (the same kernels, so the results must agree bit for bit), on calls in a
row that share the GPU's scratch array, the later ones on smaller operators;
and the refusal of an operator that needs more scratch than the fixed array."""
import ctypes as C

import numpy as np
import pytest

import airtoolsii_amd as air
from airtoolsii_amd import _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    # the HIP runtime libairhip.so itself links against, so device pointers are shared
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DevVec:
    def __init__(self, h, n):
        self.h, self.n, self.p = h, n, C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), max(4 * n, 4)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.n and self.h.hipMemcpy(self.p, a.ctypes.data, 4 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n, np.float32)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, 4 * self.n, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def device_product(h, A, transp, v):
    m, n = A.shape
    nin, nout = (m, n) if transp else (n, m)
    din, dout = DevVec(h, nin), DevVec(h, nout)
    try:
        din.put(v)
        dout.put(np.full(nout, np.nan, np.float32))     # every output entry must be written
        L = _lib.load()
        _lib.check(L.air_apply_device(A.engine.h, transp, din.p, dout.p))
        _lib.check(L.air_synchronize(A.engine.h))
        return dout.get()
    finally:
        din.free()
        dout.free()


def operators():
    """Largest first, so later calls run on smaller operators over scratch the
    earlier ones filled.  The last flag says whether the products are
    deterministic (fixed summation order): the fan-beam matrix-free adjoint
    accumulates in LDS with FP64 atomics, so two calls may differ in the last
    FP64 bits and, after rounding, by one FP32 ulp."""
    t96, t40 = np.arange(0.0, 180.0, 3.0), np.arange(0.0, 180.0, 4.5)
    return [
        ("matrix-free N=96", air.paralleltomo(96, t96, 136, None, 0, 0), True),
        ("batch N=56,40", air.paralleltomo_batch([56, 40], [np.arange(0.0, 180.0, 5.0), t40]), True),
        ("matrix-free N=40", air.paralleltomo(40, t40, 57, None, 0, 0), True),
        ("stored N=32", air.paralleltomo(32, t40, 45), True),
        ("matrix-free fan N=24", air.fanlineartomo(24, np.arange(0.0, 360.0, 10.0), 34, None, None, None, 0, 0),
         False),
    ]


def test_device_products_equal_air_apply():
    h = hip()
    rng = np.random.default_rng(11)
    ops = operators()
    for rep in range(2):
        for name, A, exact in ops:
            m, n = A.shape
            lo = -1.0 if exact else 0.0                  # no cancellation where the order may vary
            x = rng.uniform(lo, 1, n).astype(np.float32)
            y = rng.uniform(lo, 1, m).astype(np.float32)
            fx = device_product(h, A, 0, x)
            gy = device_product(h, A, 1, y)
            ax, ay = A.matvec(x.astype(np.float64)), A.rmatvec(y.astype(np.float64))
            assert np.array_equal(fx.astype(np.float64), ax), (name, rep, "A x")
            if exact:
                assert np.array_equal(gy.astype(np.float64), ay), (name, rep, "A' y")
            else:
                np.testing.assert_allclose(gy, ay, rtol=2.5e-7, atol=0)
            print(f"{name} (call {rep + 1}): A x and A' y equal air_apply", flush=True)


def test_operator_beyond_the_scratch_is_refused():
    h = hip()
    N = 2100                                           # 8 N^2 bytes of FP64 accumulators > 32 MiB
    A = air.paralleltomo(N, np.array([0.0, 45.0]), 2 * N, None, 0, 0)
    assert 8 * N * N > 32 << 20
    m, n = A.shape
    with pytest.raises(_lib.AirError) as ei:
        device_product(h, A, 1, np.ones(m, np.float32))
    assert ei.value.code == -1 and "scratch" in str(ei.value)
    # the forward product needs no scratch and still runs
    x = np.random.default_rng(2).uniform(0, 1, n).astype(np.float32)
    assert np.array_equal(device_product(h, A, 0, x).astype(np.float64), A.matvec(x.astype(np.float64)))
