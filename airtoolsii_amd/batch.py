"""Many independent paralleltomo problems solved together on the engine.

`paralleltomo_batch` builds every problem on the GPU (testprobs/paralleltomo.m,
d = p-1) and joins them into one block-diagonal operator in the LDS-tiled
layout; `batch_sirt` then runs SIRT on all of them at once, one launch per
pass.  Each problem keeps its own weights (get_mfun_dfun.m) and default
relaxation 1.9/rho_p (calc_relaxpar.m), so the result equals running `method`
on every problem separately.  No counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import DTYPES, METHODS, Engine, check, load, ptr
from .operators import AirOperator
from .sirt import sirt


class BatchOperator(AirOperator):
    """The block-diagonal operator of a batch, with its per-problem offsets."""

    def __init__(self, engine, dtype, method, nprob):
        super().__init__(engine, dtype, False)
        self.method = method
        self.nprob = nprob
        ro = np.zeros(nprob + 1, np.int64)
        co = np.zeros(nprob + 1, np.int64)
        rho = np.zeros(nprob)
        M = np.zeros(self.shape[0])
        D = np.zeros(self.shape[1])
        check(load().air_batch_info(engine.h, ptr(ro), ptr(co), ptr(rho), ptr(M), ptr(D)))
        self.row_off, self.col_off, self.rho, self.M_scaled, self.D = ro, co, rho, M, D

    def split_cols(self, x):
        return [x[self.col_off[i]:self.col_off[i + 1]] for i in range(self.nprob)]

    def join_rows(self, bs):
        return np.concatenate([np.ravel(np.asarray(b, np.float64)) for b in bs])


def paralleltomo_batch(Ns, thetas, ps=None, method="cimmino", dtype="f32", device=0):
    """Batch of paralleltomo(N_i, theta_i, p_i) problems (p_i = round(sqrt(2) N_i) by default)."""
    Ns = np.asarray(Ns, np.int32)
    thetas = [np.atleast_1d(np.asarray(t, np.float64)) for t in thetas]
    if len(thetas) != Ns.size:
        raise ValueError("one angle vector per problem")
    if ps is None:
        ps = np.floor(np.sqrt(2) * Ns + 0.5).astype(np.int32)
    ps = np.asarray(ps, np.int32)
    nA = np.array([t.size for t in thetas], np.int32)
    th = np.ascontiguousarray(np.concatenate(thetas))
    e = Engine(device)
    check(load().air_build_paralleltomo_batch(e.h, int(Ns.size), ptr(Ns), ptr(nA), ptr(th), ptr(ps),
                                              METHODS[method], DTYPES[dtype]))
    return BatchOperator(e, dtype, method, int(Ns.size))


def batch_sirt(A: BatchOperator, bs, K, options=None):
    """Run the batch's method on every problem (default lambda each); returns
    the list of per-problem iterates x_K and the info of the joint run."""
    opts = dict(options or {})
    opts.update({"relaxpar": 1.0, "rho": 1.0})
    custom = {"M": A.M_scaled}
    if A.method in ("drop", "sart"):
        custom["D"] = A.D
    X, info, _ = sirt(custom, A, A.join_rows(bs), K, None, opts)
    return A.split_cols(X[:, -1]), info
