"""Autotuners: the multi-instance energy launch (batched form) and the tile of
the batched slogdet kernels.

The choices it searches are the ones that change speed but never results
(every row's arithmetic is fixed, so any choice gives bit-identical sigma and
energies):

* ``waves_per_block`` (1, 2, 4, 8): the workgroup tile; more waves share a CU's
  LDS row lists and scheduling slots;
* partition: one wave per row (``persistent=False``), or a persistent grid of
  ``grid_blocks`` workgroups pulling rows from a work counter (16, 32 or 64
  waves per CU);
* row order: instances in the caller's order, or by estimated cost per row
  (connected elements) descending, so the longest rows start first and the
  tail of the launch is short.

Each candidate runs a sampled launch (a prefix of every instance's rows,
``sample_frac`` of the work) ``reps`` times after one warm-up; the fastest
median wins.  Results are cached per workload signature in memory and, if
``cache_path`` is given, in a JSON file.

``tune_slogdet(n)`` picks the tile of the 8 < n <= 64 slogdet kernels (lanes
per matrix: a group of lanes holds one matrix, each lane every lanes-th row;
at n equal to a bucket's column count also the kernel with the bounds fixed
at compile time or read at run time), per kernel kind (forward, forward keeping the LU, backward from A, backward from
the LU): each choice the library is built for is timed on a batch of random
matrices (the mean of back-to-back launches) and the fastest is set for every
later launch of that kind at that n.  Every choice runs the same operation sequence per element, so results
are bit-identical whichever wins.
"""

from __future__ import annotations

import json
import os
import time
from dataclasses import asdict, dataclass
from typing import Dict, List, Optional, Sequence

import torch

from .batch import ElocPlan, Instance, Tuning

_MEMO: Dict[str, dict] = {}


@dataclass
class Choice:
    waves_per_block: int
    persistent: bool
    waves_per_cu: int  # persistent grid size, in waves per CU (0 when not persistent)
    order: str  # "given" or "cost"

    def tuning(self, n_cu: int) -> Tuning:
        if not self.persistent:
            return Tuning(self.waves_per_block, False, 0)
        return Tuning(self.waves_per_block, True, max(1, n_cu * self.waves_per_cu // self.waves_per_block))


def candidates() -> List[Choice]:
    out = []
    for wpb in (1, 2, 4, 8):
        out.append(Choice(wpb, False, 0, "given"))
        for wpc in (16, 32, 64):
            for order in ("given", "cost"):
                out.append(Choice(wpb, True, wpc, order))
    return out


def order_instances(instances: Sequence[Instance], cost_per_row: Optional[Sequence[float]], order: str) -> List[int]:
    idx = list(range(len(instances)))
    if order == "cost" and cost_per_row is not None:
        idx.sort(key=lambda k: -cost_per_row[k])
    return idx


def signature(instances: Sequence[Instance]) -> str:
    parts = [f"{x.n_orb}:{x.n_rows}:{int(x.use_hb)}:{x.eps1:g}" for x in instances]
    dev = torch.cuda.get_device_name() if torch.cuda.is_available() else "none"
    return dev + "|" + ",".join(parts)


def _time_plan(plan: ElocPlan, energy: torch.Tensor, reps: int) -> float:
    plan.launch(energy)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.launch(energy)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def autotune(instances: Sequence[Instance], cost_per_row: Optional[Sequence[float]] = None,
             sample_frac: float = 0.05, reps: int = 3, cache_path: Optional[str] = None,
             cands: Optional[List[Choice]] = None, log=None) -> dict:
    """Pick the fastest launch choice for this set of instances.  Returns
    ``{"choice": Choice fields, "tuning": Tuning fields, "order": [...], "trials": [...]}``."""
    sig = signature(instances)
    if sig in _MEMO:
        return _MEMO[sig]
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            disk = json.load(f)
        if sig in disk:
            _MEMO[sig] = disk[sig]
            return disk[sig]
    n_cu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    energy = torch.empty(len(instances), dtype=torch.float64, device="cuda")
    trials = []
    best = None
    for ch in (cands or candidates()):
        idx = order_instances(instances, cost_per_row, ch.order)
        ranges = []
        for k in idx:
            x = instances[k]
            m = max(1, int(round(x.n_rows * sample_frac))) if x.n_rows else 0
            ranges.append((x, 0, min(m, x.n_rows), -1))
        plan = ElocPlan(ranges, ch.tuning(n_cu))
        t = _time_plan(plan, energy, reps)
        plan.free()
        trials.append({**asdict(ch), "seconds": t})
        if log:
            log(f"autotune {asdict(ch)}: {t * 1e3:.2f} ms")
        if best is None or t < best[1]:
            best = (ch, t)
    ch = best[0]
    res = {"choice": asdict(ch), "tuning": asdict(ch.tuning(n_cu)),
           "order": order_instances(instances, cost_per_row, ch.order), "trials": trials,
           "sample_frac": sample_frac}
    _MEMO[sig] = res
    if cache_path:
        disk = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                disk = json.load(f)
        disk[sig] = res
        with open(cache_path, "w") as f:
            json.dump(disk, f)
    return res


def tune_slogdet(n: int, batch: int = 262_144, kinds=("forward_lu", "backward_lu"), reps: int = 5,
                 log=None) -> dict:
    """Time every tile choice of each kernel kind at n (8 < n <= 64) and set
    the fastest per kind.  Kinds: "forward", "forward_lu", "backward" (from A)
    and "backward_lu".  Returns {"n", "lanes": {kind: lanes}, "trials": {kind:
    [{lanes, ms}]}}."""
    from . import slogdet as S
    key = f"slogdet|{torch.cuda.get_device_name()}|{n}|{batch}|{','.join(kinds)}"
    if key in _MEMO:
        for kind, lanes in _MEMO[key]["lanes"].items():
            S.set_tile(kind, n, lanes)
        return _MEMO[key]
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.randn(batch, n, n, dtype=torch.float64, device="cuda", generator=g)
    A += n * torch.eye(n, dtype=torch.float64, device="cuda")
    cot = torch.randn(batch, dtype=torch.float64, device="cuda", generator=g)
    _, _, lu, perm = S.slogdet_fwd_lu(A)
    ops = {"forward": lambda: S.slogdet_fwd(A), "forward_lu": lambda: S.slogdet_fwd_lu(A),
           "backward": lambda: S.slogdet_bwd_from_a(A, None, cot),
           "backward_lu": lambda: S.slogdet_bwd_from_lu(lu, perm, cot)}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    trials, chosen = {}, {}
    for kind in kinds:
        run = ops[kind]
        trials[kind] = []
        for lanes in S.tile_choices(n):
            S.set_tile(kind, n, lanes)
            run()
            torch.cuda.synchronize()
            # launches back to back between two events: host gaps between
            # single calls would be counted otherwise
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            trials[kind].append({"lanes": lanes, "ms": ms})
            if log:
                log(f"tune_slogdet n={n} {kind} lanes={lanes}: {ms:.4f} ms")
        chosen[kind] = min(trials[kind], key=lambda t: t["ms"])["lanes"]
        S.set_tile(kind, n, chosen[kind])
    res = {"n": n, "batch": batch, "lanes": chosen, "trials": trials}
    _MEMO[key] = res
    return res


__all__ = ["Choice", "candidates", "autotune", "order_instances", "signature", "tune_slogdet"]
