"""Learned relaxation for SIRT: data, training and evaluation on the engine.

The network (airtoolsii_amd.mlp) sees 16 statistics of the residual at each
iteration and returns lambda_k and a stop logit.  It is trained on SIRT
trajectories the engine generates for seeded phantoms:

* along a trajectory every step uses the greedy relaxation lambda*_k, the
  value that minimises ||x_{k-1} + lambda d_k - x_true|| before projection
  (kept in (0.02, 11.98)/rho); the target of the lambda output is lambda*_k rho;
* the stop target of a state is 1 once its error is within `stop_tol` of the
  smallest error on the trajectory.

Plain SIRT, the method the network is judged against, uses the relaxation
parameter from train_relaxpar (train/train_relaxpar.m) and the discrepancy
principle (auxil/check_stoprules.m:56-72) with tau * ||e||; the learned run
keeps the same discrepancy principle as its fallback stop.  Nothing here has
a reference counterpart beyond those two.
"""
from __future__ import annotations

import time

import numpy as np

from . import phantoms
from .mlp import MLP, init_params
from .operators import as_operator
from .sirt import SirtRun, sirt
from .train import train_relaxpar

KINDS = ("shepplogan", "smooth", "threephases", "grains")


def problem(A, N, seed, noise=(0.01, 0.05), kinds=KINDS):
    """A seeded test problem: (x_true, b, ||e||, description)."""
    rng = np.random.default_rng(seed)
    kind = kinds[int(rng.integers(len(kinds)))]
    eta = float(rng.uniform(*noise))
    ps = int(rng.integers(1, 2 ** 31 - 1))
    if kind == "shepplogan":
        img = phantoms.shepplogan(N)
    elif kind == "smooth":
        img = phantoms.smooth(N, int(rng.integers(1, 5)))
    elif kind == "threephases":
        img = phantoms.threephases(N, 100, ps)
    else:
        img = phantoms.grains(N, None, ps)
    x = np.asarray(img, np.float64).ravel(order="F")
    b0 = A.matvec(x)
    b, e = phantoms.add_noise(b0, eta, seed=ps)
    return x, b, float(np.linalg.norm(e)), {"kind": kind, "noise": eta, "seed": ps}


def stop_labels(err, tol=0.01):
    """1 where the error is within (1 + tol) of the trajectory's smallest error."""
    err = np.asarray(err, np.float64)
    return (err <= (1.0 + tol) * np.nanmin(err)).astype(np.float32)


def trajectory(A, x, b, noise_norm, kmax, method="sart", bounds=(0.0, 1.0), rho=None, params=None):
    """One trajectory with greedy targets: (features kmax x 16, targets kmax x 2,
    errors kmax + 1).  With `params` the network's lambda drives the run and the
    greedy lambda is only recorded as the target (data along the network's path)."""
    opts = {"relaxpar": "greedy", "x_true": x, "noise_norm": noise_norm}
    if params is not None:
        opts["mlp"] = params
    if bounds[0] is not None:
        opts["lbound"] = bounds[0]
    if bounds[1] is not None:
        opts["ubound"] = bounds[1]
    if rho is not None:
        opts["rho"] = rho
    run = SirtRun(method, A, b, kmax, None, opts)
    run.iterate(kmax)
    xk, info = run.fetch()
    rec = run.records()
    err = np.concatenate([rec["err"][:kmax], [np.linalg.norm(xk - x)]]) / np.linalg.norm(x)
    T = np.stack([rec["target"][:kmax], stop_labels(err)[:kmax]], 1)
    return rec["features"][:kmax], T, err


def dataset(A, N, seeds, kmax, method="sart", bounds=(0.0, 1.0), rho=None, params=None):
    """Features and targets of the trajectories of the given problem seeds."""
    X, T, E = [], [], []
    for sd in seeds:
        x, b, nn, _ = problem(A, N, int(sd))
        f, t, e = trajectory(A, x, b, nn, kmax, method, bounds, rho, params)
        ok = np.all(np.isfinite(f), axis=1) & np.isfinite(t[:, 0])
        X.append(f[ok]); T.append(t[ok]); E.append(e)
    return np.concatenate(X), np.concatenate(T), E


def train_network(X, T, epochs=30, batch=4096, lr=2e-3, stop_weight=0.5, seed=0, net=None):
    """Train on the device; returns (params, per-epoch loss)."""
    net = net or MLP()
    return net.train(init_params(seed), X, T, epochs=epochs, batch=batch, lr=lr, stop_weight=stop_weight,
                     seed=seed)


def _timed_run(method, A, b, kmax, opts):
    """(x, info, setup seconds, iteration seconds) of one device-resident run."""
    t0 = time.perf_counter()
    run = SirtRun(method, A, b, kmax, None, opts)
    t1 = time.perf_counter()
    run.iterate(kmax)
    x, info = run.fetch()
    t2 = time.perf_counter()
    return x, info, t1 - t0, t2 - t1, run


def train_on_policy(A, N, seeds, kmax, rounds=3, epochs=40, method="sart", bounds=(0.0, 1.0), rho=None,
                    seed=0, log=None):
    """Dataset aggregation: round 0 trains on greedy trajectories; every later
    round adds trajectories driven by the current network (greedy targets
    recorded along its own path) and retrains on everything so far, so the
    network learns to correct the states it actually visits."""
    seeds = list(seeds)
    parts = np.array_split(np.asarray(seeds), rounds)
    Xs, Ts = [], []
    params, hist = None, None
    for r, part in enumerate(parts):
        X, T, _ = dataset(A, N, part, kmax, method, bounds, rho, params)
        Xs.append(X); Ts.append(T)
        params, hist = train_network(np.concatenate(Xs), np.concatenate(Ts), epochs=epochs, seed=seed)
        if log:
            log({"round": r, "samples": int(sum(x.shape[0] for x in Xs)), "loss": float(hist[-1])})
    return params, hist


def evaluate(A, N, seeds, params, lam_plain, kmax, method="sart", tau=1.02, bounds=(0.0, 1.0), rho=None,
             repeats=1):
    """Plain SIRT (lambda from train_relaxpar, DP with tau ||e||) against the
    learned relaxation on held-out problems.  Each run is timed in two parts:
    set-up (b to the device, weights, r0) and the iterations up to the stop
    (x back to the host included); best of `repeats`.  The error level is the
    plain run's final error.  Both methods are then timed the same way, to the
    first iteration at which each reaches that level: a run with K = that
    iteration (set-up, iterations, x to the host), so neither time includes
    the iterations a device-side stop check launches past the stop.  The
    plain method's first iteration comes from its save list, the learned
    method's from a run that records errors.  `plain_s` is the plain run with
    its discrepancy-principle stop, as a caller would run it."""
    A = as_operator(A)
    rows = []
    for sd in seeds:
        x, b, nn, desc = problem(A, N, int(sd))
        base = {}
        if bounds[0] is not None:
            base["lbound"] = bounds[0]
        if bounds[1] is not None:
            base["ubound"] = bounds[1]
        if rho is not None:
            base["rho"] = rho
        po = dict(base, relaxpar=lam_plain, stoprule={"type": "DP", "taudelta": tau * nn})
        # the discrepancy principle stays on as the fallback stop of the learned run
        lo = dict(base, relaxpar="learned", mlp=params, noise_norm=nn,
                  stoprule={"type": "DP", "taudelta": tau * nn})
        bp = bl = None
        for _ in range(max(repeats, 1)):
            xp, ip, sp_, tp_, _ = _timed_run(method, A, b, kmax, po)
            xl, il, sl_, tl_, _ = _timed_run(method, A, b, kmax, lo)
            if bp is None or sp_ + tp_ < bp[0] + bp[1]:
                bp = (sp_, tp_)
            if bl is None or sl_ + tl_ < bl[0] + bl[1]:
                bl = (sl_, tl_)
        ep = np.linalg.norm(xp - x) / np.linalg.norm(x)
        el = np.linalg.norm(xl - x) / np.linalg.norm(x)
        # the plain method's first iteration at the level ep, and its time to it
        kp = max(int(ip["finaliter"]), 1)
        Xs, _, _ = sirt(method, A, b, list(range(1, kp + 1)), None, po)
        perr = np.linalg.norm(Xs - x[:, None], axis=0) / np.linalg.norm(x)
        at = np.nonzero(perr <= ep)[0]
        kp_hit = int(at[0]) + 1 if at.size else kp
        tp_hit = None
        for _ in range(max(repeats, 1)):
            _, _, sh_, th_, _ = _timed_run(method, A, b, kp_hit, po)
            tp_hit = sh_ + th_ if tp_hit is None else min(tp_hit, sh_ + th_)
        xk, info, _, _, run = _timed_run(method, A, b, kmax, dict(lo, x_true=x))
        kl = int(info["finaliter"])
        errs = run.records()["err"] / np.linalg.norm(x)
        curve = np.concatenate([errs[:kl], [np.linalg.norm(xk - x) / np.linalg.norm(x)]])
        hit = np.nonzero(curve <= ep)[0]
        k_hit = int(hit[0]) if hit.size else None
        per_iter = bl[1] / max(int(il["finaliter"]), 1)
        t_hit = None
        if k_hit is not None:
            # the learned run that stops at k_hit, timed like the others (set-up,
            # k_hit iterations, x back to the host); best of `repeats`
            for _ in range(max(repeats, 1)):
                _, _, sh_, th_, _ = _timed_run(method, A, b, max(k_hit, 1), lo)
                t_hit = sh_ + th_ if t_hit is None else min(t_hit, sh_ + th_)
        # the myopic bound of any lambda-only controller: the greedy lambda*,
        # which knows x_true, on the same problem
        go = dict(base, relaxpar="greedy", x_true=x, noise_norm=nn)
        xg, ig, _, _, rg = _timed_run(method, A, b, kmax, go)
        kg_fin = int(ig["finaliter"])
        gcurve = np.concatenate([rg.records()["err"][:kg_fin] / np.linalg.norm(x),
                                 [np.linalg.norm(xg - x) / np.linalg.norm(x)]])
        ghit = np.nonzero(gcurve <= ep)[0]
        k_greedy = int(ghit[0]) if ghit.size else None
        rows.append({"seed": int(sd), **desc, "plain_err": float(ep), "plain_iters": int(ip["finaliter"]),
                     "plain_s": bp[0] + bp[1], "learned_err": float(el), "learned_iters": int(il["finaliter"]),
                     "learned_s": bl[0] + bl[1], "learned_stop": int(il["stoprule"]), "iters_to_plain_err": k_hit,
                     "time_to_plain_error": t_hit, "plain_iters_to_err": kp_hit, "plain_time_to_err": tp_hit,
                     "plain_setup_s": bp[0], "learned_setup_s": bl[0],
                     "greedy_iters_to_plain_err": k_greedy})
    return rows


def summarize(rows):
    """Medians over all problems.  A learned run that never reaches the plain
    method's error counts as an infinite ratio (a failure), so the medians are
    over every problem, not only the ones that succeeded; the `_reached`
    figures are the same medians over the problems that did reach it."""
    pe = np.array([r["plain_err"] for r in rows]); le = np.array([r["learned_err"] for r in rows])
    ps = np.array([r["plain_s"] for r in rows]); ls = np.array([r["learned_s"] for r in rows])
    hit = [r for r in rows if r["time_to_plain_error"] is not None]
    ratio = np.array([np.inf if r["time_to_plain_error"] is None else r["time_to_plain_error"] / r["plain_time_to_err"]
                      for r in rows])
    ih = np.array([np.inf if r["iters_to_plain_err"] is None else r["iters_to_plain_err"] / max(r["plain_iters_to_err"], 1)
                   for r in rows])
    dp = np.array([np.inf if r["time_to_plain_error"] is None else r["time_to_plain_error"] / r["plain_s"]
                   for r in rows])
    ok = np.isfinite(ratio)
    # JSON has no infinity: a median that falls on a failure is reported as null
    fin = lambda v: float(v) if np.isfinite(v) else None
    return {"problems": len(rows), "plain_err_median": float(np.median(pe)), "learned_err_median": float(np.median(le)),
            "learned_err_le_plain_frac": float(np.mean(le <= pe * 1.0)),
            "iters_to_plain_err_over_plain_median": fin(np.median(ih)),
            "iters_to_plain_err_over_plain_median_reached": fin(np.median(ih[ok])) if ok.any() else None,
            "time_to_plain_error_over_plain_median_reached": fin(np.median(ratio[ok])) if ok.any() else None,
            "time_to_plain_error_over_plain_p75": fin(np.percentile(ratio, 75, method="higher")),
            "plain_s_total": float(ps.sum()), "learned_s_total": float(ls.sum()),
            "plain_iters_median": float(np.median([r["plain_iters"] for r in rows])),
            "learned_iters_median": float(np.median([r["learned_iters"] for r in rows])),
            "reached_plain_err_frac": len(hit) / len(rows),
            "setup_s_median": float(np.median([r["plain_setup_s"] for r in rows])),
            "plain_iter_s_median": float(np.median([(r["plain_s"] - r["plain_setup_s"]) / max(r["plain_iters"], 1)
                                                    for r in rows])),
            "learned_iter_s_median": float(np.median([(r["learned_s"] - r["learned_setup_s"]) / max(r["learned_iters"], 1)
                                                      for r in rows])),
            "time_to_plain_error_over_plain_median": fin(np.median(ratio)),
            "time_to_plain_error_over_plain_dp_run_median": fin(np.median(dp)),
            "plain_time_to_err_median": float(np.median([r["plain_time_to_err"] for r in rows])),
            "plain_dp_run_s_median": float(np.median(ps)),
            "greedy_iters_to_plain_err_over_plain_median": float(np.median(
                [r["greedy_iters_to_plain_err"] / max(r["plain_iters_to_err"], 1) for r in rows
                 if r.get("greedy_iters_to_plain_err") is not None] or [np.nan]))}


def plain_relaxpar(A, N, seed, kmax, method="sart", bounds=(0.0, 1.0), rho=None):
    """train_relaxpar's lambda on one training problem (the plain method's parameter)."""
    x, b, _, _ = problem(A, N, int(seed))
    opts = {}
    if bounds[0] is not None:
        opts["lbound"] = bounds[0]
    if bounds[1] is not None:
        opts["ubound"] = bounds[1]
    if rho is not None:
        opts["rho"] = rho
    return float(train_relaxpar(A, b, x, method, kmax, opts))
