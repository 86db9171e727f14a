"""Engine-backed operators with the AIR Tools II ``afun`` convention.

An :class:`AirOperator` is what ``sirt``/``cimmino``/... accept as ``A``.  It
is callable exactly like a MATLAB afun handle (auxil/afun_matrix.m:44-53,
testprobs/paralleltomo.m:150-160):

    A(None, 'size')      -> (m, n)
    A(v, 'notransp')     -> A @ v   (m-vector, float64)
    A(w, 'transp')       -> A' @ w  (n-vector, float64)

and it also supports ``A @ v``, ``A.T @ w`` and ``A.shape``.  The matrix (CSR
and CSC) or the matrix-free ray geometry lives on the GPU; products run as HIP
kernels.  Host vectors are FP64 at this boundary, as in MATLAB.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import DTYPES, Engine, check, f64, load, ptr


class AirOperator:
    def __init__(self, engine: Engine, dtype: str, matrix_free: bool, geometry: dict | None = None):
        self.engine = engine
        self.dtype = dtype
        self.matrix_free = matrix_free
        self.geometry = geometry or {}
        m, n, nnz = engine.size()
        self.shape = (m, n)
        self.nnz = nnz if nnz >= 0 else None

    # ---- afun convention
    def __call__(self, v, transp_flag):
        if transp_flag == "size":
            return self.shape
        if transp_flag == "notransp":
            return self.matvec(v)
        if transp_flag == "transp":
            return self.rmatvec(v)
        raise ValueError("transp_flag must be 'size', 'notransp' or 'transp'")

    def matvec(self, v):
        v = f64(v)
        if v.shape[0] != self.shape[1]:
            raise ValueError("Incorrect length of input vector")
        out = np.empty(self.shape[0])
        check(load().air_apply(self.engine.h, 0, ptr(v), ptr(out)))
        return out

    def rmatvec(self, w):
        w = f64(w)
        if w.shape[0] != self.shape[0]:
            raise ValueError("Incorrect length of input vector")
        out = np.empty(self.shape[1])
        check(load().air_apply(self.engine.h, 1, ptr(w), ptr(out)))
        return out

    def graph_launches(self):
        """HIP graphs launched on this operator's engine (graph mode of the SIRT loop)."""
        n = _lib.C.c_int64()
        check(load().air_graph_launches(self.engine.h, _lib.C.byref(n)))
        return n.value

    def __matmul__(self, v):
        return self.matvec(v)

    @property
    def T(self):
        return _Transposed(self)

    def layout(self):
        """Storage used by the SIRT loop (tiled or compressed rows) and its sizes."""
        L = _lib.Layout()
        check(load().air_layout_info(self.engine.h, L))
        out = {}
        for f, _ in L._fields_:
            v = getattr(L, f)
            out[f] = [round(float(a), 2) for a in v] if hasattr(v, "__len__") else v
        return out

    # ---- stored matrix access
    def csr_arrays(self):
        """(rowptr int64, colind int32, values float64 widened from storage)."""
        if self.matrix_free:
            raise ValueError("matrix-free operator has no stored CSR")
        m, n = self.shape
        rp = np.empty(m + 1, np.int64)
        ci = np.empty(max(self.nnz, 1), np.int32)
        v = np.empty(max(self.nnz, 1), np.float64)
        check(load().air_get_csr(self.engine.h, ptr(rp), ptr(ci), ptr(v)))
        return rp, ci[: self.nnz], v[: self.nnz]

    def to_scipy(self):
        import scipy.sparse as sp
        rp, ci, v = self.csr_arrays()
        return sp.csr_matrix((v, ci, rp), shape=self.shape)


class FunctionOperator(AirOperator):
    """A caller's function handle as the operator (sirt/sirt.m:104-115,
    auxil/check_inputs.m:67-75): ``fn(None, 'size')`` gives (m, n),
    ``fn(v, 'notransp')`` A v and ``fn(w, 'transp')`` A' w, all FP64.  The
    engine calls it for every product (air_set_afun) and runs the rest of the
    SIRT loop on the GPU; the weights are those of get_mfun_dfun.m's
    function-handle branches.  An exception inside ``fn`` fails the engine
    call; the exception is kept in ``self.error``."""

    def __init__(self, fn, device: int = 0):
        mn = np.ravel(np.asarray(fn(None, "size")))
        if mn.size != 2:
            raise ValueError("A(v, 'size') must return [m n]")
        m, n = int(mn[0]), int(mn[1])
        self.fn = fn
        self.error = None

        def call(ctx, transp, inp, out):
            kin, kout = (m, n) if transp else (n, m)
            try:
                v = np.ctypeslib.as_array(inp, (kin,)).copy() if kin > 0 else np.zeros(0)
                y = np.ravel(np.asarray(fn(v, "transp" if transp else "notransp"), np.float64))
                if y.shape[0] != kout:
                    raise ValueError(f"the operator function returned {y.shape[0]} values, expected {kout}")
                if kout > 0:
                    np.ctypeslib.as_array(out, (kout,))[:] = y
                return 0
            except Exception as ex:          # reported through AIR_E_CALLBACK
                self.error = ex
                return -1

        self._cb = _lib.AFUN(call)           # kept alive with the operator
        e = Engine(device)
        check(load().air_set_afun(e.h, m, n, self._cb, None))
        AirOperator.__init__(self, e, "f32", True)


class _Transposed:
    def __init__(self, op):
        self.op = op
        self.shape = (op.shape[1], op.shape[0])

    def __matmul__(self, w):
        return self.op.rmatvec(w)


def as_operator(A, dtype: str = "f32", device: int = 0) -> AirOperator:
    """Wrap ``A`` as an engine operator.  Accepts an :class:`AirOperator`
    (returned unchanged), a function handle with the afun convention (a
    :class:`FunctionOperator`), a SciPy sparse matrix or a dense 2-D array
    (uploaded as CSR; the engine builds the CSC on the GPU)."""
    if isinstance(A, AirOperator):
        return A
    if callable(A) and not hasattr(A, "shape"):
        return FunctionOperator(A, device)
    import scipy.sparse as sp
    M = sp.csr_matrix(A, dtype=np.float64)
    M.sum_duplicates()
    M.sort_indices()
    e = Engine(device)
    m, n = M.shape
    check(load().air_set_matrix_csr(e.h, m, n, ptr(np.ascontiguousarray(M.indptr, np.int64)),
                                    ptr(np.ascontiguousarray(M.indices, np.int32)),
                                    ptr(np.ascontiguousarray(M.data, np.float64)), DTYPES[dtype]))
    return AirOperator(e, dtype, False)


def from_matlab_csc(m, n, jc, ir, pr, dtype: str = "f32", device: int = 0) -> AirOperator:
    """The operand a MEX gateway receives (mxGetJc / mxGetIr / mxGetPr)."""
    e = Engine(device)
    check(load().air_set_matrix_csc(e.h, int(m), int(n), ptr(np.ascontiguousarray(jc, np.int64)),
                                    ptr(np.ascontiguousarray(ir, np.int64)),
                                    ptr(np.ascontiguousarray(pr, np.float64)), DTYPES[dtype]))
    return AirOperator(e, dtype, False)


__all__ = ["AirOperator", "FunctionOperator", "as_operator", "from_matlab_csc", "_lib"]
