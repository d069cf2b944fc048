#!/usr/bin/env python
"""Local-energy hot path benchmark on one MI355X.

One step = Slater–Condon row assembly of every determinant of the batch
(get_local_connections layout: count pass, int64 scan, fill streamed through a
reusable HBM buffer) plus the batched FP64 slogdet of the Slater blocks.

Workload (N2/cc-pVDZ-sized): 28 orbitals, 7 alpha + 7 beta electrons,
synthetic FP64 integrals from 80 Cholesky-like vectors, 1,000,000 unique
determinants with excitation rank ~ 1/r^2, heat-bath eps1 = 1e-7, and 2,000,000
7x7 slogdets (occupied rows of a random 28x7 orbital matrix + 0.05 N(0,1)).

Prints one JSON line.  ``value`` = connected H-elements per second over the
whole step (inputs resident in HBM, outputs written to HBM).
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "connected H-elems/s + slogdets/s, served vs plain; p50/p99 E_loc latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X vector FP64
# HBM bytes moved per algorithmic byte by the fill kernel, from rocprofv3
# FETCH_SIZE / WRITE_SIZE runs calibrated on a plain 1 GiB copy
# (profiles/local_conn_fill_config1.txt): (5.689 + 14.906) GB / 14.748 GB.
FILL_TRAFFIC_RATIO = 1.391
# L2 (TCC) hit rate of the same fill launch, TCC_HIT / (TCC_HIT + TCC_MISS)
# from the counter run recorded there.
FILL_L2_HIT_RATE = 0.564


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(n_rows: int):
    from detnqs_amd import synth
    t = time.time()
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(n_rows, seed=1)
    A = synth.slater_matrices(dets, 28, 7, 7, seed=3001, noise=0.05)
    log(f"workload generated in {time.time() - t:.1f}s: {dets.shape[0]} dets, {A.shape[0]} matrices")
    return ints, dets, A


def _cpu_model() -> str:
    """lscpu's model name, with the family / model / stepping numbers (a VM
    may report only a short name)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        f = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                f.setdefault(k.strip().lower(), v.strip())
        if "model name" in f:
            ids = ", ".join(f"{k} {f[k]}" for k in ("cpu family", "model", "stepping") if k in f)
            return f"{f['model name']} ({ids})" if ids else f["model name"]
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ints, dets, budget_s: float):
    """The reference C++ core (oracle/_ref) if built, else the oracle, timed on
    the host cores on a bounded sample of the same rows (count + checksum
    streaming), in a child process started with OMP_PROC_BIND=close,
    OMP_PLACES=cores: a sample sized to about a sixth of the budget, one
    warm-up run, then the median of five."""
    import tempfile
    from detnqs_amd import synth
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as td:
        synth.write_fcidump(os.path.join(td, "w.FCIDUMP"), ints)
        np.save(os.path.join(td, "dets.npy"), np.ascontiguousarray(dets[: min(dets.shape[0], 400_000)]))
        cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", td, "--cpu-seconds", str(budget_s)]
        r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=max(120.0, 8 * budget_s))
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed: {r.stderr[-2000:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res.update({"cores": threads, "cpu_model": _cpu_model(),
                "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}})
    return res


def cpu_child(td: str, budget_s: float):
    """Runs in the child process (no GPU): the timed CPU sample."""
    from oracle.oracle import Lib, ref_available
    kind = "reference" if ref_available() else "port"
    lib = Lib("ref" if kind == "reference" else "oracle")
    ctx = lib.ctx(os.path.join(td, "w.FCIDUMP"), 28)
    ctx.hb_prepare(1e-15)
    dets = np.load(os.path.join(td, "dets.npy"))

    def run(rows):
        if kind == "reference":
            return ctx.local_conn_stream(dets[:rows], 28, True, 1e-7)[0]
        c, _ = ctx.local_conn_stats(dets[:rows], 28, True, 1e-7)
        return int(c.sum())

    rows, _, nnz = _cpu_sample(run, budget_s / 6.0, dets.shape[0])
    run(rows)  # warm-up
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        run(rows)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    print(json.dumps({"value": nnz / med, "unit": "connected H-elements/s", "kind": kind,
                      "sample": f"first {rows} rows of the workload ({nnz} elements), get_local_connections rows "
                                f"with heat-bath eps1=1e-7; median of 5 runs after a warm-up",
                      "runs_s": times, "median_s": med}), flush=True)


def bench_slogdet(args):
    """Forward-only batched FP64 slogdet, one matrix per lane (n <= 8): 2M 7x7
    (the step's Slater-block shape), 4M 4x4 and 2M 8x8 random matrices A + 3I,
    inputs resident in HBM; best of --steps calls by HIP events after warmup."""
    import torch
    from detnqs_amd.slogdet import slogdet_fwd
    out = {}
    for n, B in ((7, 2_000_000), (4, 4_000_000), (8, 2_000_000)):
        g = torch.Generator(device="cuda").manual_seed(n)
        A = torch.randn(B, n, n, dtype=torch.float64, device="cuda", generator=g)
        A += 3 * torch.eye(n, dtype=torch.float64, device="cuda")
        for _ in range(max(args.warmup, 1)):
            slogdet_fwd(A)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(args.steps):
            e0.record()
            slogdet_fwd(A)
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        alg = B * (8 * n * n + 16)
        out[f"n{n}"] = {"matrices": B, "ms": best, "slogdets_per_s": B / best * 1e3,
                        "roofline": {"bound": "hbm", "achieved": alg / best / 1e6, "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": alg / best / 1e6 / HBM_PEAK_GBS}}
        del A
    # n > 8: every tile of every kernel kind (the autotuner's trials), with
    # the fraction of HBM peak of the chosen tile; bytes per matrix: forward
    # 8n^2 + 16, forward + LU 16n^2 + 4n + 16, backward (from A) 16n^2 + 8,
    # backward from the LU 16n^2 + 4n + 8
    from detnqs_amd.tune import tune_slogdet
    rows = {}
    for n in (12, 16, 20, 24, 32, 48, 64):
        B = 262_144 if n <= 32 else 65_536
        t = tune_slogdet(n, B, ("forward", "forward_lu", "backward", "backward_lu"), reps=max(args.steps, 3),
                         log=log)
        nb = {"forward": 8 * n * n + 16, "forward_lu": 16 * n * n + 4 * n + 16, "backward": 16 * n * n + 8,
              "backward_lu": 16 * n * n + 4 * n + 8}
        rows[f"n{n}"] = {"matrices": B, "lanes": t["lanes"],
                         "best_ms": {k: min(x["ms"] for x in v) for k, v in t["trials"].items()},
                         "frac_hbm": {k: B * nb[k] / (min(x["ms"] for x in v) * 1e-3) / 1e9 / HBM_PEAK_GBS
                                      for k, v in t["trials"].items()},
                         "trials": t["trials"]}
    print(json.dumps({"metric": "FP64 slogdets/s", "form": "slogdet", "dtype": "f64",
                      "data": "synthetic", "n_gpus": 1, **out, "rows": rows}), flush=True)


def bench_diff(args):
    """Differentiation form: B = 262,144 FP64 matrices at n = 12 and n = 16,
    A = Phi[rows] + 0.05 N(0,1) (Phi 40 x n, 1% near-singular), forward that keeps
    the LU + backward that reuses it, and the reference-signature backward."""
    import torch
    from detnqs_amd import synth
    from detnqs_amd.slogdet import slogdet_bwd_from_a, slogdet_bwd_from_lu, slogdet_fwd_lu
    from detnqs_amd.tune import tune_slogdet
    B = 262_144
    out = {}
    for n in (12, 16):
        # the autotuner picks the lanes per matrix on random matrices (the
        # differentiation pair), then the measured workload runs with it
        tuned = tune_slogdet(n, B, ("forward_lu", "backward_lu", "backward"), reps=max(args.steps, 3), log=log)
        A = torch.from_numpy(synth.selection_matrices(B, n, 40, seed=400 + n, noise=0.05,
                                                      near_singular_frac=0.01)).cuda()
        cot = torch.randn(B, dtype=torch.float64, device=A.device)
        for _ in range(args.warmup):
            _, _, lu, perm = slogdet_fwd_lu(A)
            slogdet_bwd_from_lu(lu, perm, cot)
        torch.cuda.synchronize()
        # each phase as --steps launches back to back between two events, so
        # host launch gaps between single calls are not counted
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        _, _, lu, perm = slogdet_fwd_lu(A)
        outs = [slogdet_fwd_lu(A) for _ in range(2)]
        torch.cuda.synchronize()
        ev[0].record()
        for i in range(args.steps):
            outs[i % 2] = slogdet_fwd_lu(A)
        ev[1].record()
        for _ in range(args.steps):
            slogdet_bwd_from_lu(lu, perm, cot)
        ev[2].record()
        for _ in range(args.steps):
            slogdet_bwd_from_a(A, None, cot)
        ev[3].record()
        torch.cuda.synchronize()
        fwd = ev[0].elapsed_time(ev[1])
        bwd = ev[1].elapsed_time(ev[2])
        bwd_a = ev[2].elapsed_time(ev[3])
        fwd, bwd, bwd_a = fwd / args.steps, bwd / args.steps, bwd_a / args.steps
        nn = n * n
        # fwd reads A, writes sign/logabs + LU + perm; bwd reads LU + perm + cot, writes grad
        fwd_bytes = B * (8 * nn + 16 + 8 * nn + 4 * n)
        bwd_bytes = B * (8 * nn + 4 * n + 8 + 8 * nn)
        out[f"n{n}"] = {"fwd_ms": fwd, "bwd_lu_ms": bwd, "bwd_refactor_ms": bwd_a,
                        "matrices_per_s_fwd_bwd": B / ((fwd + bwd) * 1e-3),
                        "fwd_GBps": fwd_bytes / (fwd * 1e-3) / 1e9, "bwd_GBps": bwd_bytes / (bwd * 1e-3) / 1e9,
                        "bwd_refactor_GBps": B * (16 * nn + 8) / (bwd_a * 1e-3) / 1e9,
                        "fwd_frac_hbm": fwd_bytes / (fwd * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "bwd_frac_hbm": bwd_bytes / (bwd * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "bwd_refactor_frac_hbm": B * (16 * nn + 8) / (bwd_a * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "lanes_per_matrix": tuned["lanes"], "tile_trials": tuned["trials"]}
    print(json.dumps({"metric": "slogdet fwd+bwd matrices/s (differentiation form)", "form": "diff",
                      "batch": B, "dtype": "f64", "data": "synthetic", "results": out}), flush=True)


def _cpu_sample(run, budget_s, n_max, start=64):
    """Grows a row prefix until run(rows) takes ~budget_s; returns (rows, seconds, result)."""
    rows = min(start, n_max)
    while True:
        t0 = time.perf_counter()
        res = run(rows)
        dt = time.perf_counter() - t0
        if dt >= budget_s * 0.4 or rows >= n_max:
            return rows, dt, res
        rows = min(n_max, int(rows * max(2.0, budget_s / max(dt, 1e-3) * 0.9)))


def bench_spaces(args):
    """The rows next to the local-energy path on the 28-orbital workload
    (7a+7b, 80 Cholesky vectors, config-1 determinants):
      prep  prepare_det_batch of 32M determinants (tiled 1M set), every feature,
            kmax = 8, against the HF reference;
      P     gen_perturbative_dets, static heat-bath eps1 = 1e-6, V = first 20k dets;
      PT2   compute_pt2 with heat-bath eps1 = 1e-6 on the same V, normalised
            amplitudes c_i ~ N(0,1) exp(-i / 2000).
    Each is timed on torch's current stream (the stream the library launches on)
    with inputs resident in HBM; the reference C++ (oracle/_ref) on the host
    cores gives the CPU rate on a bounded row sample."""
    import tempfile
    import torch
    from detnqs_amd import core, synth
    from oracle.oracle import Lib, ref_available
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(1_000_000, seed=1)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    out = {"metric": "det features/s, P-space candidates/s, EN-PT2 rows/s", "form": "spaces", "dtype": "f64",
           "data": "synthetic", "n_gpus": 1}

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    # ---- prepare_det_batch
    big = core.dets_to_device(np.tile(dets, (32, 1)))
    B, K = big.shape[0], 8
    ref = synth.hf_det(7, 7)
    t = timed(lambda: core.prepare_det_batch_device(big, ref, 28, 7, 7, K, True, True, True, True), args.steps)
    per_det = 16 + 4 * 14 + 1 + 1 + K * (4 + 4 + 1 + 4 + 4)  # read det; occ, k, phase, 5 hp arrays
    prep = {"dets": B, "ms": t * 1e3, "dets_per_s": B / t,
            "roofline": {"bound": "hbm", "kernel": "prep_staged_kernel", "algorithmic_bytes_per_det": per_det,
                         "achieved": B * per_det / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": B * per_det / t / 1e9 / HBM_PEAK_GBS}}
    del big
    torch.cuda.empty_cache()

    # ---- P-space discovery and PT2 on V = first 20k dets
    nv = 20_000
    V = dets[:nv]
    rng = np.random.Generator(np.random.PCG64(77))
    c = rng.standard_normal(nv) * np.exp(-np.arange(nv) / 2000.0)
    c /= np.linalg.norm(c)
    dV = core.dets_to_device(V)
    dc = torch.from_numpy(c).to(dV.device)
    e_ref = core.compute_variational_energy(V, c, ctx, 28, False, 0.0) - 0.1
    # candidates of the static set (singles |H| > 1e-15 + heat-bath prefix doubles)
    off, _ = core.local_conn_count(dV, ctx, 28, True, 1e-6)
    n_cand = int(off[-1].item()) - nv  # the diagonal is not a candidate
    holder = {}

    def p_run():
        holder["P"] = core.gen_perturbative_dets_device(dV, 28, ctx, None, "static", 1e-6)

    tp = timed(p_run, 1)
    pspace = {"V": nv, "candidates": n_cand, "P": holder["P"].n, "ms": tp * 1e3, "candidates_per_s": n_cand / tp}
    del holder["P"]

    def pt2_run():
        holder["r"] = core.compute_pt2_device(dV, dc, ctx, 28, e_ref, True, 1e-6)

    tq = timed(pt2_run, 1)
    r = holder["r"]
    pt2 = {"V": nv, "ms": tq * 1e3, "rows_per_s": nv / tq, "n_ext": r["n_ext"],
           "e_pt2_internal": r["e_pt2_internal"], "e_pt2_external": r["e_pt2_external"]}

    # ---- CPU: the reference C++ on the host cores, bounded samples
    if not args.no_cpu_baseline and ref_available():
        R = Lib("ref")
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "w.FCIDUMP")
            synth.write_fcidump(p, ints)
            rctx = R.ctx(p, 28)
        rctx.hb_prepare(1e-15)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rows, dt, _ = _cpu_sample(lambda m: R.prepare_det_batch(dets[:m], ref, 28, 7, 7, K, True, True, True, True),
                                  args.cpu_seconds / 3, dets.shape[0], start=4096)
        prep["cpu_baseline"] = {"value": rows / dt, "unit": "dets/s", "cores": 1, "kind": "reference",
                                "sample": f"first {rows} dets ({dt:.1f}s); the reference loop is serial"}
        rows, dt, _ = _cpu_sample(lambda m: rctx.gen_perturbative(V[:m], 28, mode="static", eps1=1e-6),
                                  args.cpu_seconds / 3, nv)
        cc = int(core.local_conn_count(dV[:rows], ctx, 28, True, 1e-6)[0][-1].item()) - rows
        pspace["cpu_baseline"] = {"value": cc / dt, "unit": "candidates/s", "cores": cores, "kind": "reference",
                                  "sample": f"first {rows} rows of V ({cc} candidates, {dt:.1f}s), OpenMP {cores}"}
        cs = c.copy()
        rows, dt, _ = _cpu_sample(lambda m: rctx.compute_pt2(V[:m], cs[:m], 28, e_ref, True, 1e-6, threads=cores),
                                  args.cpu_seconds / 3, nv)
        pt2["cpu_baseline"] = {"value": rows / dt, "unit": "V rows/s", "cores": cores, "kind": "reference",
                               "sample": f"first {rows} rows of V ({dt:.1f}s), OpenMP {cores}"}
    out.update(prep=prep, pspace=pspace, pt2=pt2)
    print(json.dumps(out), flush=True)


def bench_project(args):
    """Projected assembly (get_ham_conn, static heat-bath eps1 = 1e-6) on the
    28-orbital workload: V = the first --project-rows config-1 determinants,
    output H_VV and H_VP as canonical COO plus the sorted P space, all left in
    HBM.  Roofline per determinant row: 12 B per kept element (4 B column +
    8 B value), 16 B per probed candidate (the key compare), 24 B per row."""
    import tempfile
    import torch
    from detnqs_amd import core, synth
    from oracle.oracle import Lib, ref_available
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(max(args.project_rows, 1024), seed=1)[:args.project_rows]
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    dV = core.dets_to_device(dets)
    nv = dV.shape[0]
    off, _ = core.local_conn_count(dV, ctx, 28, True, 1e-6)
    n_cand = int(off[-1].item())  # diagonal + kept singles + heat-bath doubles, each probed in V
    holder = {}

    def run():
        holder["B"] = core.ham_blocks_device(dV, ctx, 28, 1, True, 1e-6)

    for _ in range(args.warmup):
        run()
        holder.pop("B")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t = 0.0
    for k in range(args.steps):  # each call timed alone; its result is freed outside the timing
        if k:
            holder.pop("B")
        torch.cuda.synchronize()
        e0.record()
        run()
        e1.record()
        torch.cuda.synchronize()
        t += e0.elapsed_time(e1) * 1e-3
    t /= args.steps
    B = holder["B"]
    nvv, nvp, npp = B.nnz_vv, B.nnz_vp, B.n_p
    alg = 12 * (nvv + nvp) + 16 * n_cand + 24 * nv
    out = {"metric": "projected assembly elements/s (get_ham_conn)", "form": "project", "dtype": "f64",
           "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"28 orbitals 7a+7b, 80 Cholesky vectors, V = first {nv} config-1 dets, "
                                  "static heat-bath eps1=1e-6"},
           "V": nv, "candidates": n_cand, "nnz_vv": nvv, "nnz_vp": nvp, "n_p": npp, "ms": t * 1e3,
           "elements_per_s": (nvv + nvp) / t,
           "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": alg,
                        "note": "whole call (count, fill, radix sorts of the COO keys and of P)"}}
    if not args.no_cpu_baseline and ref_available():
        R = Lib("ref")
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "w.FCIDUMP")
            synth.write_fcidump(p, ints)
            rctx = R.ctx(p, 28)
        rctx.hb_prepare(1e-15)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rows, dt, res = _cpu_sample(lambda m: rctx.ham_conn(dets[:m], 28, True, 1e-6), args.cpu_seconds, nv)
        el = int(res["H_VV"]["val"].size + res["H_VP"]["val"].size)
        out["cpu_baseline"] = {"value": el / dt, "unit": "elements/s", "cores": cores, "kind": "reference",
                               "sample": f"first {rows} rows of V ({el} elements, {dt:.1f}s), OpenMP {cores}"}
    print(json.dumps(out), flush=True)


def bench_eff(args):
    """SpGEMM down-folding H_eff = H_VV + H_VP D^-1 H_PV (get_ham_eff, sigma
    regularisation, eps 1e-12) of the projected blocks of V = the first
    --eff-rows config-1 determinants (28 orbitals 7a+7b, static heat-bath
    eps1 = 1e-6), E_ref = min H_PP - 1.  Blocks, H_PP and the result stay in
    HBM; the call is timed alone.  Algorithmic bytes: H_VP read as rows and
    as columns (12 B x nnz(H_VP) x 2), H_PP's diagonal (8 B x n_P) and the
    result (12 B x nnz(H_eff))."""
    import ctypes
    import tempfile
    import torch
    from detnqs_amd import _lib, core, synth
    from detnqs_amd.core import _ptr, stream_handle
    from oracle.oracle import Lib, ref_available
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(max(args.eff_rows, 1024), seed=1)[:args.eff_rows]
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    nv = dets.shape[0]
    B = core.ham_blocks_device(dets, ctx, 28, 1, True, 1e-6)
    vr, vc, vv, pr, pc, pv, dp = B.to_device()
    hpp = core.get_ham_diag_device(dp, ctx)
    e_ref = float(hpp.min().item()) - 1.0

    def call():
        h = ctypes.c_void_p()
        _lib.call("dq_ham_eff", _ptr(vr), _ptr(vc), _ptr(vv), B.nnz_vv, _ptr(pr), _ptr(pc), _ptr(pv), B.nnz_vp, nv,
                  _ptr(hpp), B.n_p, e_ref, 1, 1e-12, stream_handle(), ctypes.byref(h))
        return core.Blocks(h)

    for _ in range(args.warmup):
        call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t = 0.0
    res = None
    for _ in range(args.steps):
        res = None
        torch.cuda.synchronize()
        e0.record()
        res = call()
        e1.record()
        torch.cuda.synchronize()
        t += e0.elapsed_time(e1) * 1e-3
    t /= args.steps
    nnz_eff = res.nnz_vv
    alg = 12 * B.nnz_vp * 2 + 8 * B.n_p + 12 * nnz_eff
    out = {"metric": "H_eff SpGEMM (get_ham_eff)", "form": "eff", "dtype": "f64", "data": "synthetic", "n_gpus": 1,
           "config": {"workload": f"28 orbitals 7a+7b, V = first {nv} config-1 dets, static heat-bath eps1=1e-6, "
                                  "sigma regularisation eps 1e-12"},
           "V": nv, "nnz_vv": B.nnz_vv, "nnz_vp": B.nnz_vp, "n_p": B.n_p, "nnz_eff": nnz_eff, "ms": t * 1e3,
           "eff_elements_per_s": nnz_eff / t,
           "roofline": {"bound": "hbm", "achieved": alg / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": alg / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes": alg,
                        "note": "whole call (CSR/CSC builds, triple emission, stable sorts, folds)"}}
    if not args.no_cpu_baseline and ref_available():
        R = Lib("ref")
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "w.FCIDUMP")
            synth.write_fcidump(p, ints)
            rctx = R.ctx(p, 28)
        rctx.hb_prepare(1e-15)
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        m = 64
        while True:
            blk = rctx.ham_conn(dets[:m], 28, True, 1e-6)
            h = rctx.ham_diag(blk["det_P"])
            t0 = time.perf_counter()
            eff = R.ham_eff(blk["H_VV"], blk["H_VP"], h, float(h.min()) - 1.0, "sigma", 1e-12)
            dt = time.perf_counter() - t0
            if dt >= args.cpu_seconds * 0.4 or m >= nv:
                break
            m = min(nv, int(m * max(2.0, min(4.0, (args.cpu_seconds / max(dt, 1e-3)) ** 0.5))))
        out["cpu_baseline"] = {"value": eff["val"].size / dt, "unit": "H_eff elements/s", "cores": cores,
                               "kind": "reference",
                               "sample": f"V = first {m} rows ({eff['val'].size} H_eff elements, {dt:.1f}s), "
                                         f"OpenMP {cores}"}
    print(json.dumps(out), flush=True)


def _pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else None


def _serve_clients(args):
    """16 serving clients (42 orbitals, 12a+12b, 120 Cholesky vectors, one
    integral seed each) and a shared pool of 262,144 determinants (rank ~ 1/r^2,
    drawn on the GPU) that requests slice their V from."""
    from detnqs_amd import core, synth
    t = time.time()
    clients = []
    for ints in synth.config4_clients(16):
        ctx = core.IntCtx.from_integrals(42, 24, ints.h1e, ints.h2e, ints.e_nuc)
        ctx.hb_prepare(1e-15)
        clients.append(ctx)
    pool = synth.config1_dets_device(max(args.serve_max, 4096), 4100, 42, 12, 12).cpu().numpy().view(np.uint64)
    log(f"serve: 16 clients and a {pool.shape[0]}-det pool in {time.time() - t:.1f}s")
    return clients, pool


def run_serve(args, quick=False):
    """Serving form on the config-4 workload (heat-bath eps1 = 1e-8, requests of
    4k-256k dets log-uniform, host inputs).

    vmc: the DetNQS inner loop (driver.py:303-323): each of the 16 clients is
        a closed loop (submit, wait, submit) over `serve_vs` V sets, `serve_inner`
        steps per V with new amplitudes each step (the V kept under a v_key).
        Served = the continuous-batching server (its V sets keep their
        lane-ordered H_VV); also without that layout.  Plain = the same
        requests one at a time through compute_variational_energy, the drop-in
        that rebuilds everything per call, as the reference's does
        (local_conn.cpp:148-229).
    fresh: every request a new V, arriving Poisson at `serve_util` of the plain
        path's measured capacity (open loop): p50/p99 latency both ways.
    walk: the reference's algorithm (candidate walk + lookup) per request on a
        sample of the vmc requests, for scale."""
    import threading
    import torch
    from detnqs_amd import core
    from detnqs_amd.batch import Tuning
    from detnqs_amd.serve import Server
    n_orb, eps1 = 42, 1e-8
    clients, pool = _serve_clients(args)
    rng = np.random.Generator(np.random.PCG64(4005))
    n_vs = 1 if quick else args.serve_vs
    inner = args.serve_inner
    tune = Tuning(args.serve_wpb, True, 0)
    budget = int(args.serve_budget_gb * (1 << 30))

    def draw():
        s_ = int(min(round(np.exp(rng.uniform(np.log(args.serve_min), np.log(args.serve_max)))), pool.shape[0]))
        lo = int(rng.integers(0, pool.shape[0] - s_ + 1))
        return lo, s_

    count_memo = {}

    def elements(w, lo, s_):
        if (w, lo, s_) not in count_memo:
            d = core.dets_to_device(pool[lo:lo + s_])
            count_memo[(w, lo, s_)] = core.local_conn_count(d, clients[w], n_orb, True, eps1)[1]
        return count_memo[(w, lo, s_)]

    # ---- vmc: closed-loop clients
    work = []  # per client: list of (lo, size, key, [coefficients per step])
    key = 0
    for w in range(16):
        vs = []
        for _ in range(n_vs):
            key += 1
            lo, s_ = draw()
            vs.append((lo, s_, key, [rng.standard_normal(s_) for _ in range(inner)]))
        work.append(vs)
    el_total = sum(elements(w, lo, s_) * inner for w in range(16) for lo, s_, _, _ in work[w])
    n_req = 16 * n_vs * inner
    rows_total = sum(s_ * inner for w in range(16) for _, s_, _, _ in work[w])

    def served(layout):
        S = Server(budget, args.serve_rows_per_launch, 2, tune, vv_layout=layout)
        res = [[None] * (n_vs * inner) for _ in range(16)]
        lat = [[0.0] * (n_vs * inner) for _ in range(16)]

        def client(w):
            j = 0
            for lo, s_, k, cs in work[w]:
                for c in cs:
                    r = S.wait(S.submit(clients[w], pool[lo:lo + s_], c, n_orb, True, eps1, v_key=k))
                    res[w][j], lat[w][j] = r.energy, r.latency
                    j += 1

        th = [threading.Thread(target=client, args=(w,)) for w in range(16)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        span = time.perf_counter() - t0
        st = S.stats()
        S.close()
        return res, [x for row in lat for x in row], span, st

    def plain():
        """One worker serving the same closed-loop clients first come, first
        served: a client's next request is issued when its previous one
        returns; latency counts the wait behind other clients' requests."""
        res = [[None] * (n_vs * inner) for _ in range(16)]
        lat = []
        nxt = [0] * 16
        issued = [0.0] * 16  # issue time of each client's pending request
        t0 = time.perf_counter()
        queue = list(range(16))  # clients with a pending request, in issue order
        while queue:
            w = queue.pop(0)
            j = nxt[w]
            lo, s_, _, cs = work[w][j // inner]
            res[w][j] = core.compute_variational_energy(pool[lo:lo + s_], cs[j % inner], clients[w], n_orb, True, eps1)
            done = time.perf_counter() - t0
            lat.append(done - issued[w])
            nxt[w] += 1
            if nxt[w] < n_vs * inner:
                issued[w] = done
                queue.append(w)
        return res, lat, time.perf_counter() - t0

    def same(a, b):
        return sum(np.float64(x).tobytes() == np.float64(y).tobytes() for ra, rb in zip(a, b) for x, y in zip(ra, rb))

    # warm-up: module loads, code objects, the block cache
    w0 = work[0][0]
    core.compute_variational_energy(pool[w0[0]:w0[0] + 4096], w0[3][0][:4096], clients[0], n_orb, True, eps1)
    S = Server(budget, args.serve_rows_per_launch, 2, tune)
    S.wait(S.submit(clients[0], pool[:4096], w0[3][0][:4096], n_orb, True, eps1, v_key=10 ** 9))
    S.close()
    e_lay, lat_lay, span_lay, st_lay = served(True)
    e_nol, lat_nol, span_nol, _ = served(False)
    e_pl, lat_pl, span_pl = plain()
    vmc = {"clients": 16, "v_sets_per_client": n_vs, "inner_steps_per_v": inner, "requests": n_req,
           "rows": rows_total, "elements": int(el_total),
           "served_elems_per_s": el_total / span_lay, "plain_elems_per_s": el_total / span_pl,
           "speedup": span_pl / span_lay,
           "served_without_layout_elems_per_s": el_total / span_nol, "speedup_without_layout": span_pl / span_nol,
           "served_p50_s": _pct(lat_lay, 50), "served_p99_s": _pct(lat_lay, 99),
           "plain_p50_s": _pct(lat_pl, 50), "plain_p99_s": _pct(lat_pl, 99),
           "bit_identical": f"{same(e_lay, e_pl)}/{n_req}",
           "bit_identical_without_layout": f"{same(e_nol, e_pl)}/{n_req}",
           "layouts_built": st_lay.vv_builds, "layout_rows": st_lay.vv_rows, "launches": st_lay.launches,
           "arena_peak_gb": st_lay.arena_peak / 2 ** 30}
    log(f"serve vmc: served {vmc['served_elems_per_s']:.3e} elem/s, plain {vmc['plain_elems_per_s']:.3e}, "
        f"x{vmc['speedup']:.2f} (no layout x{vmc['speedup_without_layout']:.2f}), {vmc['bit_identical']}")

    # ---- fresh: open-loop Poisson at serve_util of plain capacity
    n_fresh = args.requests if not quick else min(args.requests, 120)
    reqs = []
    for _ in range(n_fresh):
        w = int(rng.integers(0, 16))
        lo, s_ = draw()
        reqs.append((w, lo, s_, rng.standard_normal(s_)))
    el_fresh = sum(elements(w, lo, s_) for w, lo, s_, _ in reqs)
    t0 = time.perf_counter()
    for w, lo, s_, c in reqs[:32]:
        core.compute_variational_energy(pool[lo:lo + s_], c, clients[w], n_orb, True, eps1)
    cap = 32 / (time.perf_counter() - t0)  # requests/s one at a time
    lam = args.serve_util * cap
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_fresh))
    arrivals -= arrivals[0]

    def fresh_served():
        S = Server(budget, args.serve_rows_per_launch, 2, tune)
        tickets = [None] * n_fresh
        ready = [threading.Event() for _ in reqs]
        t0 = time.perf_counter()

        def submitter():
            for i, (w, lo, s_, c) in enumerate(reqs):
                d = t0 + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = S.submit(clients[w], pool[lo:lo + s_], c, n_orb, True, eps1)
                ready[i].set()

        th = threading.Thread(target=submitter)
        th.start()
        res = []
        for i in range(n_fresh):
            ready[i].wait()
            res.append(S.wait(tickets[i]))
        th.join()
        S.close()
        return [r.energy for r in res], [r.latency for r in res], max(r.t_done for r in res) - min(
            r.t_submit for r in res)

    def fresh_plain():
        t0 = time.perf_counter()
        en, lat = [], []
        for i, (w, lo, s_, c) in enumerate(reqs):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            en.append(core.compute_variational_energy(pool[lo:lo + s_], c, clients[w], n_orb, True, eps1))
            lat.append(time.perf_counter() - (t0 + arrivals[i]))
        return en, lat, time.perf_counter() - t0

    ef_s, lat_fs, span_fs = fresh_served()
    ef_p, lat_fp, span_fp = fresh_plain()
    fresh = {"requests": n_fresh, "rate_per_s": lam, "util_of_plain_capacity": args.serve_util,
             "served_p50_s": _pct(lat_fs, 50), "served_p99_s": _pct(lat_fs, 99),
             "plain_p50_s": _pct(lat_fp, 50), "plain_p99_s": _pct(lat_fp, 99),
             "served_elems_per_s": el_fresh / span_fs, "plain_elems_per_s": el_fresh / span_fp,
             "bit_identical": f"{sum(np.float64(a).tobytes() == np.float64(b).tobytes() for a, b in zip(ef_s, ef_p))}"
                              f"/{n_fresh}"}
    log(f"serve fresh: p50/p99 served {fresh['served_p50_s']:.4f}/{fresh['served_p99_s']:.4f}, "
        f"plain {fresh['plain_p50_s']:.4f}/{fresh['plain_p99_s']:.4f}")
    out = {"vmc": vmc, "fresh": fresh}
    if args.serve_walk:
        smp = [(w, v) for w in range(16) for v in work[w]][: 8 if quick else 16]
        el_w = sum(elements(w, lo, s_) for w, (lo, s_, _, _) in smp)
        t0 = time.perf_counter()
        for w, (lo, s_, _, cs) in smp:
            core.compute_variational_energy(pool[lo:lo + s_], cs[0], clients[w], n_orb, True, eps1, method="walk")
        out["walk"] = {"requests": len(smp), "plain_walk_elems_per_s": el_w / (time.perf_counter() - t0)}
        out["walk"]["served_vmc_vs_plain_walk"] = vmc["served_elems_per_s"] / out["walk"]["plain_walk_elems_per_s"]
    return out


def bench_serve(args):
    out = run_serve(args)
    print(json.dumps({"metric": METRIC, "form": "serve", "unit": "connected H-elements/s", "dtype": "f64",
                      "data": "synthetic",
                      "config": {"workload": "serving: 16 clients, 42 orbitals, 12a+12b, 120 Cholesky vectors, "
                                             f"requests {args.serve_min}-{args.serve_max} dets log-uniform, "
                                             "heat-bath eps1=1e-8, host inputs",
                                 "elements": "connected H-elements of the requests' rows (count pass)"},
                      **out}), flush=True)


def run_batched(args, quick=False):
    """Batched form on the config-2 workload: 64 heterogeneous instances
    (orbitals U{16..60}, electrons U{8..24}, dets log-uniform 4k-512k drawn on
    the GPU, heat-bath eps1 = 1e-7) and one FP64 slogdet per determinant at the
    instance's own size n = 4-24.  A step = every instance's energy + its
    slogdets.
      plain: 64 sequential per-instance calls (compute_variational_energy, then
             the instance's slogdet batch), synchronised, as a per-instance caller;
      batched: one HIP graph: the persistent multi-instance energy launch +
             the slogdet launches (launch shape from the autotuner);
      batched_vmc: `serve_inner` steps on the same V with new amplitudes: the
             instances' lane-ordered H_VV built once (timed, amortised over
             the steps), each step one graph over the layouts + slogdets."""
    import torch
    from detnqs_amd import core, synth, tune
    from detnqs_amd.batch import ElocPlan, Instance, Tuning
    from detnqs_amd.slogdet import slogdet_fwd
    specs = synth.config2_specs(args.instances, max_dets=args.batch_max)
    eps1 = args.batch_eps
    t = time.time()
    insts, mats, elems, cost = [], [], [], []
    for k, sp in enumerate(specs):
        ints, dets = synth.config2_instance(sp)
        ctx = core.IntCtx.from_integrals(sp.n_orb, sp.n_elec, ints.h1e, ints.h2e, ints.e_nuc)
        ctx.hb_prepare(1e-15)
        g = torch.Generator(device="cuda").manual_seed(sp.seed)
        c = torch.randn(dets.shape[0], dtype=torch.float64, device="cuda", generator=g)
        c /= torch.linalg.norm(c)
        x = Instance.create(ctx, dets, c, sp.n_orb, True, eps1)
        _, tot = core.local_conn_count(x.dets, ctx, sp.n_orb, True, eps1)
        mats.append(torch.randn((x.n_rows, sp.slog_n, sp.slog_n), dtype=torch.float64, device="cuda", generator=g))
        insts.append(x)
        elems.append(tot)
        cost.append(tot / max(x.n_rows, 1))
    total_el = int(sum(elems))
    total_mats = int(sum(m.shape[0] for m in mats))
    log(f"batched: {len(insts)} instances, {sum(x.n_rows for x in insts)} dets, {total_el:.3e} elements, "
        f"{total_mats} matrices, built in {time.time() - t:.0f}s")
    signs = [torch.empty(m.shape[0], dtype=torch.float64, device="cuda") for m in mats]
    logs = [torch.empty_like(s) for s in signs]
    reps = 1 if quick else max(args.steps, 1)

    def plain_step():
        en = []
        for x, m, sg, lg in zip(insts, mats, signs, logs):
            en.append(core.compute_variational_energy_device(x.dets, x.coeffs, x.ctx, x.n_orb, True, eps1))
            slogdet_fwd(m, out=(sg, lg))
            torch.cuda.synchronize()
        return en

    tuned = tune.autotune(insts, cost, sample_frac=args.tune_frac, reps=1 if quick else 2, log=log)
    tn = Tuning(**tuned["tuning"])
    order = tuned["order"]
    energy = torch.empty(len(insts), dtype=torch.float64, device="cuda")

    def graph_of(plan):
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            plan.launch(energy, stream=s)  # warm-up outside capture
            for m, sg, lg in zip(mats, signs, logs):
                slogdet_fwd(m, out=(sg, lg))
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=s):
            plan.launch(energy, stream=s)
            for m, sg, lg in zip(mats, signs, logs):
                slogdet_fwd(m, out=(sg, lg))
        return g

    def timed(fn, n):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n

    t_plain = timed(plain_step, reps)
    e_plain = plain_step()
    plan = ElocPlan([(insts[k], 0, insts[k].n_rows, k) for k in order], tn)
    g = graph_of(plan)
    t_batch = timed(g.replay, reps)
    plan.check()
    e_b = energy.cpu().numpy()
    ident = sum(np.float64(a).tobytes() == np.float64(b).tobytes() for a, b in zip(e_b, e_plain))
    del g
    plan.free()
    # VMC steps on the same V: layouts once, then products
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for x in insts:
        x.cache_vv()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    plan = ElocPlan([(insts[k], 0, insts[k].n_rows, k) for k in order], tn)
    g = graph_of(plan)
    t_lay = timed(g.replay, max(reps, 3))
    plan.check()
    e_l = energy.cpu().numpy()
    ident_l = sum(np.float64(a).tobytes() == np.float64(b).tobytes() for a, b in zip(e_l, e_plain))
    layout_gb = sum(x.vv.nbytes for x in insts) / 2 ** 30
    del g
    plan.free()
    for x in insts:
        x.vv = None
    inner = args.serve_inner
    t_vmc = (t_build + inner * t_lay) / inner
    t_slog = timed(lambda: [slogdet_fwd(m, out=(sg, lg)) for m, sg, lg in zip(mats, signs, logs)], reps)
    out = {"instances": len(insts), "dets": int(sum(x.n_rows for x in insts)), "elements": total_el,
           "matrices": total_mats, "plain_step_s": t_plain, "batched_step_s": t_batch,
           "speedup": t_plain / t_batch, "plain_elems_per_s": total_el / t_plain,
           "batched_elems_per_s": total_el / t_batch, "energies_bit_identical": f"{ident}/{len(insts)}",
           "vmc": {"inner_steps_per_v": inner, "layout_build_s": t_build, "layout_step_s": t_lay,
                   "amortised_step_s": t_vmc, "speedup": t_plain / t_vmc,
                   "elems_per_s": total_el / t_vmc, "energies_bit_identical": f"{ident_l}/{len(insts)}",
                   "layout_gb": layout_gb},
           "slogdet_s": t_slog, "slogdets_per_s": total_mats / t_slog,
           "autotune": {"choice": tuned["choice"], "tuning": tuned["tuning"]}}
    log(f"batched: x{out['speedup']:.2f} one step, vmc x{out['vmc']['speedup']:.2f} "
        f"(build {t_build:.3f}s, step {t_lay:.4f}s), {ident}/{len(insts)} and {ident_l}/{len(insts)} identical")
    return out


def bench_batched(args):
    out = run_batched(args)
    print(json.dumps({"metric": METRIC, "form": "batched", "unit": "connected H-elements/s", "dtype": "f64",
                      "data": "synthetic",
                      "config": {"workload": f"{args.instances} instances, orbitals 16-60, electrons 8-24, dets "
                                             f"4k-{args.batch_max} log-uniform, slogdet n 4-24, heat-bath "
                                             f"eps1={args.batch_eps:g}"},
                      **out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--chunk-elems", type=float, default=1.5e9)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--project-rows", type=int, default=20000, help="project: rows of V")
    ap.add_argument("--eff-rows", type=int, default=5000, help="eff: rows of V")
    ap.add_argument("--form", choices=("plain", "diff", "slogdet", "serve", "batched", "spaces", "project", "eff"),
                    default="plain",
                    help="plain: the drop-in step (default); slogdet: FP64 forward slogdet at n = 4, 7, 8; "
                         "diff: slogdet forward+backward, "
                         "262,144 matrices at n = 12 and 16 (LU reused by the backward); serve: the "
                         "continuous-batching server against one-at-a-time calls; batched: 64 "
                         "heterogeneous instances in one graph against 64 sequential calls")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests in the open-loop trace")
    ap.add_argument("--serve-min", type=int, default=4096)
    ap.add_argument("--serve-max", type=int, default=262144)
    ap.add_argument("--serve-util", type=float, default=0.7)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0,
                    help="serve: HBM arena (V sets, their layouts, request buffers)")
    ap.add_argument("--serve-rows-per-launch", type=int, default=65536)
    ap.add_argument("--serve-wpb", type=int, default=4)
    ap.add_argument("--serve-inner", type=int, default=8, help="serve/batched: inner steps per V (vmc traces)")
    ap.add_argument("--serve-vs", type=int, default=2, help="serve: V sets per client in the vmc trace")
    ap.add_argument("--forms", type=int, default=1,
                    help="plain form: also run the serving and batched forms (quick sizes) into the line")
    ap.add_argument("--serve-walk", type=int, default=1,
                    help="serve: also time the saturation set through the candidate-walk method")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--batch-max", type=int, default=524288)
    ap.add_argument("--batch-eps", type=float, default=1e-7)
    ap.add_argument("--tune-frac", type=float, default=0.02)
    ap.add_argument("--mode", choices=("step", "copy", "write"), default="step",
                    help="copy: a plain 1 GiB device copy (calibrates HBM byte counters); "
                         "write: a plain 8 GiB device fill (practical write bandwidth)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        return cpu_child(args.cpu_child, args.cpu_seconds)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs are not part of it "
              "(rerun with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from detnqs_amd import core

    if args.mode == "write":
        n = (8 << 30) // 8
        a = torch.empty(n, dtype=torch.float64, device=core.device())
        for _ in range(args.warmup):
            a.fill_(1.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a.fill_(1.0)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "write", "bytes_written_per_launch": n * 8, "ms": dt * 1e3,
                          "GBps": n * 8 / dt / 1e9}), flush=True)
        return
    if args.mode == "copy":
        n = (1 << 30) // 8
        a = torch.ones(n, dtype=torch.float64, device=core.device())
        b = torch.empty_like(a)
        for _ in range(args.warmup):
            b.copy_(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        print(json.dumps({"mode": "copy", "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8,
                          "ms": dt * 1e3, "GBps": 2 * n * 8 / dt / 1e9}), flush=True)
        return

    if args.form == "diff":
        return bench_diff(args)
    if args.form == "slogdet":
        return bench_slogdet(args)
    if args.form == "serve":
        return bench_serve(args)
    if args.form == "batched":
        return bench_batched(args)
    if args.form == "spaces":
        return bench_spaces(args)
    if args.form == "project":
        return bench_project(args)
    if args.form == "eff":
        return bench_eff(args)

    ints, dets, A_host = build_workload(args.rows)
    dev = core.device()
    ctx = core.IntCtx.from_integrals(ints.n_orb, ints.n_elec, ints.h1e, ints.h2e, ints.e_nuc)
    t = time.time()
    ctx.hb_prepare(1e-15)
    log(f"heat-bath table built in {time.time() - t:.2f}s")
    n_orb, eps1 = 28, 1e-7
    d_dets = core.dets_to_device(dets)
    d_A = torch.from_numpy(A_host).to(dev)
    n_rows = d_dets.shape[0]
    n_mat = d_A.shape[0]
    del A_host
    offsets = torch.empty(n_rows + 1, dtype=torch.int64, device=dev)
    cap = int(args.chunk_elems)
    buf_d = torch.empty((cap, 2), dtype=torch.int64, device=dev)
    buf_v = torch.empty(cap, dtype=torch.float64, device=dev)
    sign = torch.empty(n_mat, dtype=torch.float64, device=dev)
    logabs = torch.empty(n_mat, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()

    def plan_chunks(offs):
        chunks, start = [], 0
        while start < n_rows:
            lim = offs[start] + cap
            stop = int(np.searchsorted(offs, lim, side="right")) - 1
            stop = max(stop, start + 1)
            stop = min(stop, n_rows)
            if offs[stop] - offs[start] > cap:
                raise RuntimeError("a single row exceeds the chunk buffer")
            chunks.append((start, stop))
            start = stop
        return chunks

    def step(timing):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        _, total = core.local_conn_count(d_dets, ctx, n_orb, True, eps1, offsets=offsets)
        e1.record(stream)
        offs = offsets.cpu().numpy()
        fills = []
        for (a, b) in plan_chunks(offs):
            fa = torch.cuda.Event(enable_timing=True)
            fb = torch.cuda.Event(enable_timing=True)
            fa.record(stream)
            core.local_conn_fill(d_dets[a:b], ctx, n_orb, True, eps1, offsets[a:b + 1], buf_d, buf_v)
            fb.record(stream)
            fills.append((fa, fb, b - a, int(offs[b] - offs[a])))
            if timing is not None and len(timing["fill"]) == 0:
                log(f"fill launch rows={b - a} nnz={int(offs[b] - offs[a])} "
                    f"algorithmic_bytes={24 * int(offs[b] - offs[a]) + 24 * (b - a)}")
        s0 = torch.cuda.Event(enable_timing=True)
        s1 = torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        from detnqs_amd import _lib
        _lib.call("dq_slogdet_fwd_f64", core._ptr(d_A), n_mat, 7, core._ptr(sign), core._ptr(logabs),
                  core.stream_handle())
        s1.record(stream)
        if timing is not None:
            torch.cuda.synchronize()
            timing["count_ms"].append(e0.elapsed_time(e1))
            timing["slogdet_ms"].append(s0.elapsed_time(s1))
            for fa, fb, rows, nnz in fills:
                timing["fill"].append((fa.elapsed_time(fb), rows, nnz))
        return total

    log(f"warmup x{args.warmup}")
    total = 0
    for _ in range(args.warmup):
        total = step(None)
    torch.cuda.synchronize()
    log(f"timed x{args.steps} (nnz per step = {total})")
    timing = {"count_ms": [], "slogdet_ms": [], "fill": []}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        total = step(timing)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    fill_ms = sum(f[0] for f in timing["fill"])
    fill_bytes = sum(24 * f[2] + 24 * f[1] for f in timing["fill"])
    n_launch = len(timing["fill"])
    avg_launch_ms = fill_ms / n_launch
    avg_launch_bytes = fill_bytes / n_launch
    achieved = avg_launch_bytes / (avg_launch_ms * 1e-3) / 1e9
    slog_ms = float(np.mean(timing["slogdet_ms"]))
    count_ms = float(np.mean(timing["count_ms"]))
    slog_gbs = n_mat * (8 * 49 + 16) / (slog_ms * 1e-3) / 1e9
    res = {
        "metric": METRIC,
        "value": total / (ms_per_step * 1e-3),
        "unit": "connected H-elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "N2/cc-pVDZ-sized plain GPU drop-in: 28 orbitals, 7a+7b electrons, 80 Cholesky vectors, "
                        f"{n_rows} unique dets (rank ~ 1/r^2), heat-bath eps1=1e-7, get_local_connections layout; "
                        f"{n_mat} FP64 7x7 slogdets",
            "rows": n_rows,
            "nnz_per_step": int(total),
            "slogdet_batch": n_mat,
            "chunk_elems": cap,
        },
        "assembly_elems_per_s": total / ((count_ms + fill_ms / args.steps) * 1e-3),
        "slogdets_per_s": n_mat / (slog_ms * 1e-3),
        "phase_ms": {"count_scan": count_ms, "fill": fill_ms / args.steps, "slogdet": slog_ms},
        "roofline": {
            "bound": "hbm",
            "kernel": "local_conn_kernel<FILL>",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": avg_launch_bytes * FILL_TRAFFIC_RATIO,
            "l2_hit_rate": FILL_L2_HIT_RATE,
            "counters": "profiles/local_conn_fill_config1.txt",
            "algorithmic_bytes_per_launch": avg_launch_bytes,
            "avg_launch_ms": avg_launch_ms,
        },
        "slogdet_roofline": {"bound": "hbm", "achieved": slog_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": slog_gbs / HBM_PEAK_GBS,
                             "fp64_tflops": n_mat * 203 / (slog_ms * 1e-3) / 1e12,
                             "fp64_frac": n_mat * 203 / (slog_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS},
    }
    del buf_d, buf_v, d_A
    torch.cuda.empty_cache()
    if args.forms:
        # the metric's other parts: served vs plain with p50/p99 (16 clients,
        # 42 orbitals) and the batched form (64 instances), at quick sizes
        log("serving form")
        res["serve"] = run_serve(args, quick=True)
        log("batched form")
        res["batched"] = run_batched(args, quick=True)
    if not args.no_cpu_baseline:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline(ints, dets, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
