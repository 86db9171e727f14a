#!/usr/bin/env python3
"""SIRT throughput on one MI355X: paralleltomo N=512, 360 angles, p=725.

    python bench.py [--steps K] [--warmup W] [--method cimmino|sart|drop|...]
                    [--dtype f32|bf16|f16] [--matrix-free] [--no-cpu-baseline]
    python bench.py --copy-calibration     (a plain 1 GiB device copy, for
                                            scaling HBM byte counters)
    python bench.py --study batch [--batch-problems P --batch-sequential]
        256 uneven paralleltomo problems (N 64..384, 30..360 angles) in one
        block-diagonal operator, Cimmino K=100; optionally against one by one
    python bench.py --study spmm [--spmm-K K]
        64 right-hand sides on the N=512, 360-angle operator (SpMM passes)
    python bench.py --study graph [--graph-iters G]
        direct launches against graph mode at N = 64 .. 512
    python bench.py --study art
        ART/CART sweeps on the GPU against the CPU oracle at N = 64 and 256
    python bench.py --study c3 [--c3-K K]
        SART at fanlineartomo N=1024, 720 angles, p=1448: the matrix stored in
        FP32, BF16 and FP16, and matrix-free
    python bench.py --study learned [--learn-train T --learn-eval E --learn-K K]
        learned relaxation at paralleltomo N=256, 180 angles, p=362: greedy
        trajectories of T seeded problems, network trained on the GPU, then
        time to the plain method's error on E held-out problems

One step is one SIRT iteration x <- P_[0,1](x + lam D A' M (b - A x)) over the
whole problem (sirt/sirt.m:213-243): a column pass (A' M r fused with the update
and projection) and a row pass (residual, M r, FP64 norms and the stop check),
on the engine's LDS-tiled layout of A.
The matrix is traced on the GPU, the phantom is grains(512, 68 grains, seed 1)
and b = A x* + 2% Gaussian noise (seed 1); all inputs are resident in HBM
before the timed region.  Per-kernel times (the roofline's launch duration)
come from HIP events around the kernels of every --timing-stride-th timed
iteration (default 8).  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "SIRT iter/s & HBM GB/s at N=512; time to fixed error, learned vs plain λ"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_c2.json")


def kernel_bytes(L, method, s_v):
    """Algorithmic HBM bytes of one launch of each SIRT kernel (unique data moved).

    Tiled layout (entries: value + 16-bit local index):
      cols   A'(M r) + update: entries, column pointers, staging runs, M r once,
             x read+write, D (DROP/SART);
      rows   forward pieces: entries, piece pointers, x once, one partial per piece;
             the step-coded forms (both passes) store values plus one 4-byte
             word per 8 stored entries instead of the 16-bit indices;
      reduce residual epilogue: row->piece lists, partials, b, M, r and M r.
    CSR+CSC layout: entries (value + int32 index), pointers, gathered vector
    once, the epilogue vectors (the reduce figure is 0)."""
    m, n, nnz = L["m"], L["n"], L["nnz"]
    hasM = method != "landweber"
    hasD = method in ("drop", "sart")
    if L["tiled"]:
        ni, P, R = L["n_internal"], L["pieces"], L["runs"]
        idx_b = 4 * (L["bwd_step_entries"] // 8) if L.get("bwd_step") else 2 * nnz
        cols = nnz * s_v + idx_b + 8 * (ni + 1) + 12 * R + 4 * (L["tiles_bwd"] + 1) + 4 * m + 8 * ni \
            + (4 * ni if hasD else 0)
        idx_f = 4 * (L["fwd_entries"] // 8) if L.get("fwd_step") else 2 * nnz
        rows = nnz * s_v + idx_f + 8 * (P + 1) + 4 * ni + 4 * P
        red = 8 * (m + 1) + 4 * P + 4 * P + 4 * m + (4 * m if hasM else 0) + 8 * m
    else:
        cols = nnz * (s_v + 4) + 8 * (n + 1) + 4 * m + (4 * n if hasD else 0) + 8 * n
        rows = nnz * (s_v + 4) + 8 * (m + 1) + 4 * n + 4 * m + (4 * m if hasM else 0) + 8 * m
        red = 0
    return cols, rows, red


def canonical_bytes(nnz, m, n, s_v):
    """Canonical two-pass bytes of one SIRT iteration with int32 indices and FP32
    vectors: 2 nnz (s_v + 4) + 4(m+1) + 4(n+1) + 4 (4n + 4m)."""
    return 2 * nnz * (s_v + 4) + 4 * (m + 1) + 4 * (n + 1) + 4 * (4 * n + 4 * m)


def cpu_baseline(N, theta, p, b, method, rho, seconds, threads):
    """The FP64 CPU oracle (a C restatement of sirt.m, OpenMP) on the same
    problem, timed over a bounded number of iterations (about `seconds` of
    CPU work) on the host cores."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    A = O.paralleltomo(N, theta, p)
    t_trace = time.perf_counter() - t0
    opts = {"rho": rho, "lbound": 0, "ubound": 1}

    def timed(k):
        t = time.perf_counter()
        O.sirt(method, A, b, k, None, opts, nthreads=threads)
        return time.perf_counter() - t

    timed(1)                                   # warm caches
    # per-call setup (weights, transpose) is measured by a 1-iteration call and subtracted
    d1 = timed(1)
    probe = timed(6)
    per_iter = max(probe - d1, 1e-9) / 5
    samples = int(min(max(seconds / per_iter, 6), 20000))
    dt = timed(samples + 1) - d1
    per_iter = max(dt, 1e-9) / samples
    # one core as well: about a quarter of `seconds` of iterations (the setup call
    # measured at one thread too)
    t1 = time.perf_counter()
    O.sirt(method, A, b, 1, None, opts, nthreads=1)
    d1s = time.perf_counter() - t1
    k1 = int(min(max(0.25 * seconds / max(per_iter * threads, 1e-9), 3), 200))
    t1 = time.perf_counter()
    O.sirt(method, A, b, k1 + 1, None, opts, nthreads=1)
    dt1 = time.perf_counter() - t1 - d1s
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            cpu = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": 1.0 / per_iter, "unit": "iter/s", "cores": threads, "kind": "port",
            "sample": f"{samples} {method} iterations of the FP64 oracle (CSR A.x, CSC A'.y, OpenMP, "
                      f"{threads} threads) on the same {A.shape[0]}x{A.shape[1]} problem, nnz={A.nnz}, "
                      f"{dt:.1f} s; oracle tracer {t_trace:.1f} s and per-call setup not included",
            "one_core_value": round(k1 / max(dt1, 1e-9), 3),
            "one_core_sample": f"{k1} iterations at 1 thread, {dt1:.1f} s",
            "cpu_model": cpu, "host_cpus": os.cpu_count()}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size: the reference point that
    turns FETCH_SIZE / WRITE_SIZE counter readings into HBM bytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": nbytes, "bytes_written": nbytes, "reps": reps}}),
          flush=True)


def measured_traffic(kernel, workload, dtype):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the copy calibration),
    when it was taken on this workload; otherwise None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("dtype") != dtype:
        return None
    k = t.get("kernels", {}).get(kernel)
    return None if k is None else int(k["hbm_bytes"])


def learned_study(args):
    """Learned lambda_k / stop against plain SIRT (train_relaxpar lambda + DP)."""
    import airtoolsii_amd as air
    from airtoolsii_amd import learned as Lm
    N, K = 256, args.learn_K
    theta = np.arange(0.0, 180.0, 1.0)
    p = 362
    t0 = time.perf_counter()
    A = air.paralleltomo(N, theta, p)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    params, hist = Lm.train_on_policy(A, N, range(1000, 1000 + args.learn_train), K, rounds=args.learn_rounds,
                                      epochs=args.learn_epochs,
                                      log=lambda d: print(json.dumps({"progress": d}), file=sys.stderr, flush=True))
    t_train = time.perf_counter() - t0
    lam = Lm.plain_relaxpar(A, N, 999, 60)
    t0 = time.perf_counter()
    rows = Lm.evaluate(A, N, range(100000, 100000 + args.learn_eval), params, lam, 400)
    t_eval = time.perf_counter() - t0
    summ = Lm.summarize(rows)
    if args.learn_save:
        os.makedirs(os.path.dirname(os.path.abspath(args.learn_save)), exist_ok=True)
        np.save(args.learn_save, np.asarray(params, np.float32), allow_pickle=False)
    out = {"metric": "time to the plain method's error, learned vs plain lambda (median ratio)",
           "value": round(summ["time_to_plain_error_over_plain_median"], 4), "unit": "ratio",
           "higher_is_better": False, "n_gpus": 1, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"paralleltomo N={N}, theta=0:1:179 ({len(theta)} angles), p={p}, SART, box [0,1]",
                      "train_trajectories": args.learn_train, "trajectory_K": K, "eval_problems": args.learn_eval,
                      "training_rounds": args.learn_rounds,
                      "phantoms": list(Lm.KINDS), "noise": "1-5% Gaussian, seeded",
                      "plain": f"lambda={lam:.4f} from train_relaxpar, DP with tau=1.02"},
           "summary": summ, "train_loss": [round(float(v), 5) for v in hist],
           "seconds": {"build": round(t_build, 2), "trajectories_and_training": round(t_train, 1),
                       "evaluate": round(t_eval, 1)}}
    print(json.dumps(out), flush=True)


LEARNED_PARAMS = os.path.join(HERE, "airtoolsii_amd", "data", "relaxation_net_sart_n256.npy")


def learned_block(args):
    """The metric's second half in the default run: the committed network
    (trained on the GPU by `--study learned --learn-save`) against plain SIRT
    with train_relaxpar's lambda and the discrepancy principle, on held-out
    seeded problems of the training geometry.  Problems whose learned run never
    reaches the plain method's error count as failures (infinite ratio)."""
    import airtoolsii_amd as air
    from airtoolsii_amd import learned as Lm
    if not os.path.exists(LEARNED_PARAMS):
        return {"skipped": "no committed network parameters"}
    params = np.load(LEARNED_PARAMS, allow_pickle=False).astype(np.float32)
    N, theta, p = 256, np.arange(0.0, 180.0, 1.0), 362
    t0 = time.perf_counter()
    A = air.paralleltomo(N, theta, p)
    lam = Lm.plain_relaxpar(A, N, 999, 60)
    seeds = range(100000, 100000 + args.learn_bench)
    rows = Lm.evaluate(A, N, seeds, params, lam, 400)
    summ = Lm.summarize(rows)
    return {"workload": f"paralleltomo N={N}, theta=0:1:179 ({len(theta)} angles), p={p}, SART, box [0,1]",
            "held_out_problems": len(rows), "seeds": f"{seeds.start}..{seeds.stop - 1}",
            "phantoms": list(Lm.KINDS), "noise": "1-5% Gaussian, seeded",
            "plain": f"lambda={lam:.4f} from train_relaxpar, DP with tau=1.02",
            "learned": "committed network, DP with tau=1.02 as the fallback stop",
            "timing": "each method timed as one call to the first iteration at the plain run's final error "
                      "(set-up, iterations, x to the host)",
            "time_to_plain_error_over_plain_median": summ["time_to_plain_error_over_plain_median"],
            "iters_to_plain_error_over_plain_median": summ["iters_to_plain_err_over_plain_median"],
            "time_ratio_vs_plain_dp_run_median": summ["time_to_plain_error_over_plain_dp_run_median"],
            "plain_time_to_err_ms_median": round(summ["plain_time_to_err_median"] * 1e3, 3),
            "plain_dp_run_ms_median": round(summ["plain_dp_run_s_median"] * 1e3, 3),
            "reached_plain_error_frac": summ["reached_plain_err_frac"],
            "time_ratio_median_reached_only": summ["time_to_plain_error_over_plain_median_reached"],
            "time_ratio_p75": summ["time_to_plain_error_over_plain_p75"],
            "plain_err_median": summ["plain_err_median"], "learned_err_median": summ["learned_err_median"],
            "plain_iters_median": summ["plain_iters_median"], "learned_iters_median": summ["learned_iters_median"],
            "seconds": round(time.perf_counter() - t0, 1)}


def batch_study(args):
    """256 uneven paralleltomo problems in one block-diagonal operator (Cimmino,
    K=100, FP32) against solving them one by one."""
    import airtoolsii_amd as air
    from airtoolsii_amd import phantoms
    rng = np.random.default_rng(args.batch_seed)
    P = args.batch_problems
    Ns = rng.integers(64, 385, P)
    nAs = rng.integers(30, 361, P)
    thetas = [np.arange(nA) * (180.0 / nA) for nA in nAs]
    kinds = ("shepplogan", "smooth", "binary", "threephases", "grains")
    K = args.batch_K
    t0 = time.perf_counter()
    A = air.paralleltomo_batch(Ns, thetas, method="cimmino")
    t_build = time.perf_counter() - t0
    L = A.layout()
    bs, xs_true = [], []
    for q in range(P):
        N = int(Ns[q])
        kind = kinds[int(rng.integers(len(kinds)))]
        ps = int(rng.integers(1, 2 ** 31 - 1))
        img = {"shepplogan": lambda: phantoms.shepplogan(N), "smooth": lambda: phantoms.smooth(N, 4),
               "binary": lambda: phantoms.binary(N, ps),
               "threephases": lambda: phantoms.threephases(N, 100, ps),
               "grains": lambda: phantoms.grains(N, None, ps)}[kind]()
        xs_true.append(np.asarray(img, np.float64).ravel(order="F"))
    y = A.matvec(np.concatenate(xs_true))
    for q in range(P):
        b0 = y[A.row_off[q]:A.row_off[q + 1]]
        eta = float(rng.uniform(0.005, 0.05))
        e = np.random.default_rng(int(rng.integers(1, 2 ** 31 - 1))).standard_normal(b0.size)
        bs.append(b0 + eta * np.linalg.norm(b0) * e / np.linalg.norm(e))
    opts = {"relaxpar": 1.0, "rho": 1.0}
    run = air.SirtRun({"M": A.M_scaled}, A, A.join_rows(bs), K, None, opts)
    run.iterate(2)
    run.kernel_timing(True)
    t0 = time.perf_counter()
    gpu_ms = run.iterate(K - 2)
    wall = time.perf_counter() - t0
    kt = run.kernel_times()
    x, info = run.fetch()
    xs = A.split_cols(x)
    errs = [float(np.linalg.norm(xs[q] - xs_true[q]) / np.linalg.norm(xs_true[q])) for q in range(P)]
    cols_b, rows_b, red_b = kernel_bytes(L, "cimmino", 4)
    it_s = wall / (K - 2)
    out = {"metric": "batched SIRT problem-iterations/s (256 uneven problems, one launch per pass)",
           "value": round(P / it_s, 1), "unit": "problem-iterations/s", "higher_is_better": True, "n_gpus": 1,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{P} paralleltomo problems, N in 64..384, 30..360 angles over [0,180), "
                                  f"p=round(sqrt(2)N), Cimmino (default lambda per problem), K={K}",
                      "phantoms": list(kinds), "noise": "0.5-5% Gaussian, seeded", "seed": args.batch_seed,
                      "m": A.shape[0], "n": A.shape[1], "nnz": L["nnz"]},
           "batch": {"ms_per_iteration": round(it_s * 1e3, 3), "gpu_ms_per_iteration": round(gpu_ms / (K - 2), 3),
                     "cols_ms": round(kt["cols_ms"] / max(kt["launches"], 1), 3),
                     "rows_ms": round(kt["rows_ms"] / max(kt["launches"], 1), 3),
                     "bytes_per_iteration": int(cols_b + rows_b + red_b),
                     "gbs": round((cols_b + rows_b + red_b) / it_s / 1e9, 1),
                     "frac_of_8tbs": round((cols_b + rows_b + red_b) / it_s / 1e9 / HBM_PEAK_GBS, 4),
                     "build_s": round(t_build, 2), "rel_err_median": round(float(np.median(errs)), 4)}}
    if args.batch_sequential:
        # the same problems one by one, each on its own operator (iterations timed on the device)
        seq_ms = 0.0
        for q in range(P):
            p = int(np.floor(np.sqrt(2) * Ns[q] + 0.5))
            Aq = air.paralleltomo(int(Ns[q]), thetas[q], p)
            rq = air.SirtRun("cimmino", Aq, bs[q], K, None, {"rho": float(A.rho[q])})
            rq.iterate(2)
            t0 = time.perf_counter()
            rq.iterate(K - 2)
            seq_ms += (time.perf_counter() - t0) * 1e3
            del rq, Aq
        out["sequential"] = {"ms_per_iteration_all_problems": round(seq_ms / (K - 2), 3),
                             "batched_speedup": round(seq_ms / (K - 2) / (it_s * 1e3), 2)}
    print(json.dumps(out), flush=True)


def spmm_study(args):
    """64 right-hand sides on the N=512, 360-angle operator in one multi-RHS SIRT
    (SpMM passes) against the single right-hand side tiled loop."""
    import airtoolsii_amd as air
    N, theta, p, R, K = 512, np.arange(0.0, 180.0, 0.5), 725, 64, args.spmm_K
    A = air.paralleltomo(N, theta, p)
    m, n = A.shape
    xs = np.stack([air.phantomgallery("grains", N, None, s).ravel(order="F") for s in range(1, R + 1)], 1)
    B = np.empty((m, R))
    for j in range(R):
        b0 = A.matvec(xs[:, j])
        B[:, j], _ = air.add_noise(b0, 0.02, seed=j + 1)
    opts = {"lbound": 0, "ubound": 1}
    air.sirt_multi("cimmino", A, B, 2, None, opts)                 # weights, rho, warm caches
    X, info = air.sirt_multi("cimmino", A, B, K, None, opts)
    ms = info["gpu_ms"] / K
    nnz = A.nnz
    bytes_it = 2 * nnz * 8 + 4 * R * (2 * m + 2 * n) + 8 * (m + n)    # CSR + CSC entries, RHS blocks, pointers
    errs = np.linalg.norm(X - xs, axis=0) / np.linalg.norm(xs, axis=0)
    out = {"metric": "multi-RHS SIRT right-hand-side iterations/s (64 RHS on one N=512 operator)",
           "value": round(R / (ms * 1e-3), 1), "unit": "rhs-iterations/s", "higher_is_better": True, "n_gpus": 1,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"paralleltomo N={N}, theta=0:0.5:179.5 ({len(theta)} angles), p={p}, {R} "
                                  f"right-hand sides (grains phantoms, 2% noise), Cimmino, box [0,1], K={K}",
                      "m": m, "n": n, "nnz": nnz},
           "multi": {"ms_per_iteration": round(ms, 4), "bytes_per_iteration": int(bytes_it),
                     "gbs": round(bytes_it / (ms * 1e-3) / 1e9, 1),
                     "frac_of_8tbs": round(bytes_it / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "rel_err_median": round(float(np.median(errs)), 5)}}
    print(json.dumps(out), flush=True)


def graph_study(args):
    """SIRT iterations per second with direct kernel launches against graph mode
    (G iterations captured once into a HIP graph and replayed), from the
    C1-sized problem up to the N=512 operator and the 256-problem batch."""
    import airtoolsii_amd as air
    G = args.graph_iters
    cases = [(64, np.arange(0.0, 180.0), 2000), (128, np.arange(0.0, 180.0), 1000),
             (256, np.arange(0.0, 180.0), 500), (512, np.arange(0.0, 180.0, 0.5), 200)]
    rows = []
    for N, theta, K in cases:
        p = int(np.floor(np.sqrt(2) * N + 0.5))
        A = air.paralleltomo(N, theta, p)
        x = air.phantomgallery("shepplogan", N).ravel(order="F")
        b, _ = air.add_noise(A.matvec(x), 0.01, seed=1)
        row = {"workload": f"paralleltomo N={N}, {len(theta)} angles, p={p}, cimmino, lbound 0", "K": K}
        xs = {}
        for g in (0, G):
            run = air.SirtRun("cimmino", A, b, K + G, None, {"lbound": 0.0, "graph": g})
            run.iterate(G)                                  # weights cached, graph captured
            t0 = time.perf_counter()
            gpu_ms = run.iterate(K)
            wall = time.perf_counter() - t0
            xs[g] = run.fetch()[0]
            row["direct" if g == 0 else "graph"] = {"iter_per_s": round(K / wall, 1),
                                                     "us_per_iter": round(wall / K * 1e6, 2),
                                                     "gpu_us_per_iter": round(gpu_ms / K * 1e3, 2)}
        row["speedup"] = round(row["direct"]["us_per_iter"] / row["graph"]["us_per_iter"], 3)
        row["identical_iterates"] = bool(np.array_equal(xs[0], xs[G]))
        rows.append(row)
        del A
    out = {"metric": "SIRT iter/s, direct launches vs graph mode", "unit": "iter/s", "n_gpus": 1,
           "dtype": "f32", "data": "synthetic", "iters_per_graph": G, "cases": rows}
    print(json.dumps(out), flush=True)


def c3_study(args):
    """SART at the fan-beam scale: fanlineartomo N=1024, 720 angles over 360
    degrees, p=1448 (R=2N, dw=2.5N, sd=3N), with the matrix stored in FP32, BF16
    and FP16 and without a stored matrix (isMatrix=0)."""
    import airtoolsii_amd as air
    N, theta, p, K = 1024, np.arange(0.0, 360.0, 0.5), 1448, args.c3_K
    x = air.phantomgallery("threephases", N).ravel(order="F")
    rows = []
    b = None
    for dt, mf in (("f32", False), ("bf16", False), ("f16", False), ("f32", True)):
        t0 = time.perf_counter()
        A = air.fanlineartomo(N, theta, p, isMatrix=0 if mf else 1, dtype=dt)
        t_build = time.perf_counter() - t0
        if b is None:
            b, _ = air.add_noise(A.matvec(x), 0.01, seed=1)
        opts = {"lbound": 0.0, "ubound": 1.0}
        t0 = time.perf_counter()
        run = air.SirtRun("sart", A, b, K + 2, None, opts)
        run.iterate(2)                                   # weights, rho, caches
        t_setup = time.perf_counter() - t0
        gpu_ms = run.iterate(K)
        xk = run.fetch()[0]
        rows.append({"storage": "none (matrix-free)" if mf else dt, "build_s": round(t_build, 3),
                     "setup_s": round(t_setup, 3), "ms_per_iteration": round(gpu_ms / K, 4),
                     "iter_per_s": round(K / (gpu_ms * 1e-3), 1),
                     "rel_err_to_phantom": round(float(np.linalg.norm(xk - x) / np.linalg.norm(x)), 5)})
        del run, A
    out = {"metric": "SART iter/s at the fan-beam scale, stored matrix by storage type and matrix-free",
           "unit": "iter/s", "n_gpus": 1, "data": "synthetic",
           "config": {"workload": f"fanlineartomo N={N}, theta=0:0.5:359.5 ({len(theta)} angles), p={p}, "
                                  f"R=2N, dw=2.5N, sd=3N, SART, box [0,1], threephases, 1% noise, K={K}"},
           "cases": rows}
    print(json.dumps(out), flush=True)


def art_study(args):
    """ART (kaczmarz) and CART (columnaction) sweeps on the GPU against the CPU
    oracle's literal sweeps (one core), on the same matrices and inputs."""
    import airtoolsii_amd as air
    from oracle import oracle as O
    rows = []
    for N, theta, K in ((64, np.arange(0.0, 180.0), 10), (256, np.arange(0.0, 180.0), 4)):
        p = int(np.floor(np.sqrt(2) * N + 0.5))
        for dt in ("f64", "f32"):
            A = air.paralleltomo(N, theta, p, dtype=dt)
            x = air.phantomgallery("shepplogan", N).ravel(order="F")
            b = A.matvec(x)
            row = {"workload": f"paralleltomo N={N}, {len(theta)} angles, p={p}, {dt} storage", "sweeps": K}
            for name, fn in (("kaczmarz", air.kaczmarz), ("columnaction", air.columnaction)):
                fn(A, b, 1)                                       # norms, buffers warm
                t0 = time.perf_counter()
                X, info = fn(A, b, K)
                row[name + "_gpu_sweeps_per_s"] = round(K / (time.perf_counter() - t0), 2)
            if dt == "f64":
                Ao = O.paralleltomo(N, theta, p)
                t0 = time.perf_counter()
                O.art("kaczmarz", Ao, b, 2)
                row["kaczmarz_cpu_sweeps_per_s"] = round(2 / (time.perf_counter() - t0), 2)
                t0 = time.perf_counter()
                O.cart("columnaction", Ao, b, 2, None, {"Kbegin": 10 ** 6})
                row["columnaction_cpu_sweeps_per_s"] = round(2 / (time.perf_counter() - t0), 2)
            rows.append(row)
            del A
    out = {"metric": "ART/CART sweeps per second, GPU (one wave per sweep) vs CPU oracle (one core)",
           "unit": "sweeps/s", "n_gpus": 1, "data": "synthetic", "cases": rows,
           "note": "no bounds (the CPU oracle projects the whole vector after every row when bounds are set); "
                   "CART flagging off in the CPU timing (its literal max|x| per column step is O(n))"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--method", default="cimmino")
    ap.add_argument("--dtype", default="f32")
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--matrix-free", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-stride", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--study", choices=["learned", "batch", "spmm", "graph", "art", "c3"], default=None)
    ap.add_argument("--c3-K", type=int, default=20)
    ap.add_argument("--graph-iters", type=int, default=32)
    ap.add_argument("--spmm-K", type=int, default=50)
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--batch-K", type=int, default=100)
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--batch-sequential", action="store_true")
    ap.add_argument("--learn-train", type=int, default=2000)
    ap.add_argument("--learn-eval", type=int, default=512)
    ap.add_argument("--learn-K", type=int, default=150)
    ap.add_argument("--learn-epochs", type=int, default=40)
    ap.add_argument("--learn-rounds", type=int, default=3)
    ap.add_argument("--learn-save", default=None)
    ap.add_argument("--learn-bench", type=int, default=512)
    ap.add_argument("--no-learned", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this engine runs one problem per GPU and no multi-GPU path is "
              "built (independent problems or right-hand sides would go one per GPU); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return
    if args.study == "learned":
        learned_study(args)
        return
    if args.study == "batch":
        batch_study(args)
        return
    if args.study == "spmm":
        spmm_study(args)
        return
    if args.study == "graph":
        graph_study(args)
        return
    if args.study == "art":
        art_study(args)
        return
    if args.study == "c3":
        c3_study(args)
        return

    import airtoolsii_amd as air

    N = args.N
    theta = np.arange(0.0, 180.0, 0.5)
    p = int(np.floor(np.sqrt(2) * N + 0.5))
    if N == 512:
        p = 725
    t0 = time.perf_counter()
    A = air.paralleltomo(N, theta, p, None, 0, 0 if args.matrix_free else 1, dtype=args.dtype)
    t_trace = time.perf_counter() - t0
    m, n = A.shape
    x_true = air.phantomgallery("grains", N, None, 1).ravel(order="F")
    b0 = A.matvec(x_true)
    b, _ = air.add_noise(b0, 0.02, seed=1)

    kmax = args.warmup + args.steps
    opts = {"lbound": 0, "ubound": 1}
    t0 = time.perf_counter()
    run = air.SirtRun(args.method, A, b, kmax, None, opts)
    t_setup = time.perf_counter() - t0
    if args.warmup:
        run.iterate(args.warmup)
    # HIP events bracket the kernels of every timing_stride-th timed iteration:
    # each event record costs the stream a few microseconds, which the other
    # iterations are spared
    run.kernel_timing(max(args.timing_stride, 1))
    t0 = time.perf_counter()
    gpu_ms = run.iterate(args.steps)
    wall = time.perf_counter() - t0
    kt = run.kernel_times()
    x, info = run.fetch()
    rel_err = float(np.linalg.norm(x - x_true) / np.linalg.norm(x_true))

    ms_per_step = wall * 1e3 / args.steps
    value = args.steps / wall
    s_v = {"f32": 4, "bf16": 2, "f16": 2, "f64": 8}[args.dtype]
    out = {"metric": METRIC, "value": round(value, 3), "unit": "iter/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": args.dtype, "data": "synthetic",
           "config": {"workload": f"paralleltomo N={N}, theta=0:0.5:179.5 ({len(theta)} angles), p={p}, "
                                  f"{args.method}, box [0,1]" + (", matrix-free" if args.matrix_free else ""),
                      "m": m, "n": n, "nnz": A.nnz, "method": args.method, "phantom": "grains(N, 68, seed 1)",
                      "noise": "2% Gaussian, seed 1"}}
    if not args.matrix_free:
        L = A.layout()
        cols_b, rows_b, red_b = kernel_bytes(L, args.method, s_v)
        launches = max(kt["launches"], 1)
        t_cols = kt["cols_ms"] / launches * 1e-3
        t_rows = kt["rows_ms"] / launches * 1e-3
        t_red = kt["reduce_ms"] / launches * 1e-3
        names = ("tiled_backward_update", "tiled_forward") if L["tiled"] else ("csc_update", "csr_resid")
        if t_cols >= t_rows:
            dom, bytes_dom, t_dom = names[0], cols_b, t_cols
        else:
            dom, bytes_dom, t_dom = names[1], rows_b, t_rows
        achieved = bytes_dom / t_dom / 1e9
        it_bytes = cols_b + rows_b + red_b
        canon = canonical_bytes(A.nnz, m, n, s_v)
        out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": measured_traffic(dom, out["config"]["workload"], args.dtype), "kernel": dom,
                           "bytes_per_launch": int(bytes_dom), "avg_launch_us": round(t_dom * 1e6, 2),
                           "timed_launches": int(kt["launches"]),
                           "timing": f"HIP events on the engine stream around the kernels of every "
                                     f"{max(args.timing_stride, 1)}th timed iteration"}
        out["iteration"] = {"layout": "tiled" if L["tiled"] else "csr+csc", "bytes": int(it_bytes),
                            "gbs": round(it_bytes * value / 1e9, 1),
                            "frac": round(it_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
                            "canonical_bytes": int(canon), "canonical_gbs": round(canon * value / 1e9, 1),
                            "canonical_frac": round(canon * value / 1e9 / HBM_PEAK_GBS, 4),
                            "cols_us": round(t_cols * 1e6, 2), "rows_us": round(t_rows * 1e6, 2),
                            "reduce_us": round(t_red * 1e6, 2), "pieces": L["pieces"], "runs": L["runs"],
                            "fwd_stream": L["fwd_stream"], "fwd_step": L["fwd_step"], "bwd_step": L["bwd_step"],
                            "bwd_step_entries": L["bwd_step_entries"],
                            "shapes": [L["fwd_lanes"], L["fwd_loads"], L["fwd_pieces_per_block"], L["bwd_lanes"],
                                       L["bwd_loads"]],
                            "fwd_waves_per_block": L["fwd_waves_per_block"],
                            "tune_fwd_us": L["tune_fwd_us"][:55], "tune_bwd_us": L["tune_bwd_us"][:11],
                            "fwd_entries": L["fwd_entries"], "bwd_entries": L["bwd_entries"]}
    out["setup_s"] = {"trace_csr_csc": round(t_trace, 3), "weights_rho_r0": round(t_setup, 3)}
    # the whole C-ABI call as a host caller sees it: b and x0 from host memory,
    # weights, r0, the K iterations and X back over PCIe (rho given)
    t0 = time.perf_counter()
    air.sirt(args.method, A, b, args.steps, None, dict(opts, rho=info["rho"]))
    e2e = time.perf_counter() - t0
    out["end_to_end"] = {"iter_per_s": round(args.steps / e2e, 1), "seconds": round(e2e, 4),
                         "what": "one air_sirt call: host b in, weights and r0, K iterations, x_K out"}
    out["result"] = {"rel_err_to_phantom": round(rel_err, 5), "resnorm_last": float(info["resnorm"][-1]),
                     "rho": info["rho"]}
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(N, theta, p, b, args.method, info["rho"], args.cpu_seconds, threads)
        out["vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
        out["vs_cpu_one_core"] = round(value / out["cpu_baseline"]["one_core_value"], 1)
    if not args.no_learned and N == 512:
        del run
        out["learned"] = learned_block(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
