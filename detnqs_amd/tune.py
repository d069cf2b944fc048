"""Autotuner for the multi-instance energy launch (batched form).

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


__all__ = ["Choice", "candidates", "autotune", "order_instances", "signature"]
