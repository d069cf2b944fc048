"""get_local_connections at the benchmark's full size: 28 orbitals 7a+7b,
1,000,000 config-1 determinants, heat-bath eps1 = 1e-7 (3.07e10 elements),
streamed through a 4e8-element buffer as a caller would.

Every element is checked for size-independent properties on the device
(orbital range, 7 alpha and 7 beta electrons, |H| > 1e-15 and finite, each
row led by its own determinant), and 64 rows spread over the space (the first
8, the last 8 and 48 between) are compared bit for bit with the oracle: the
same rows computed inside the full-size call, chunk boundaries included."""

import time

import numpy as np
import pytest
import torch

from helpers import assert_conn_equal

pytestmark = pytest.mark.gpu

N_ROWS = 1_000_000
EPS1 = 1e-7
CAP = 400_000_000


def _popcount(x: torch.Tensor) -> torch.Tensor:
    """Bit count of non-negative int64 values (SWAR)."""
    x = x - ((x >> 1) & 0x5555555555555555)
    x = (x & 0x3333333333333333) + ((x >> 2) & 0x3333333333333333)
    x = (x + (x >> 4)) & 0x0F0F0F0F0F0F0F0F
    return (x * 0x0101010101010101) >> 56


def _chunks(offs: np.ndarray, cap: int):
    out, start, n = [], 0, offs.size - 1
    while start < n:
        stop = int(np.searchsorted(offs, offs[start] + cap, side="right")) - 1
        stop = min(max(stop, start + 1), n)
        assert offs[stop] - offs[start] <= cap
        out.append((start, stop))
        start = stop
    return out


def test_config1_full_size():
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    dets = synth.config1_dets(N_ROWS, seed=1)
    n = dets.shape[0]
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    d = core.dets_to_device(dets)
    print(f"{n} determinants on the device", flush=True)
    off, total = core.local_conn_count(d, ctx, 28, True, EPS1)
    offs = off.cpu().numpy()
    print(f"count pass: {total} elements", flush=True)
    assert offs[0] == 0 and offs[-1] == total
    assert np.all(np.diff(offs) >= 1)  # every row keeps at least its diagonal
    assert total > 3e10

    sample = sorted(set(np.linspace(0, n - 1, 48).astype(np.int64).tolist()) | set(range(8)) |
                    set(range(n - 8, n)))
    got = {}
    buf_d = torch.empty((CAP, 2), dtype=torch.int64, device=d.device)
    buf_v = torch.empty(CAP, dtype=torch.float64, device=d.device)
    chunks = _chunks(offs, CAP)
    t0 = time.perf_counter()
    print(f"{len(chunks)} chunks", flush=True)
    for c, (a, b) in enumerate(chunks):
        core.local_conn_fill(d[a:b], ctx, 28, True, EPS1, off[a:b + 1], buf_d, buf_v)
        m = int(offs[b] - offs[a])
        od, ov = buf_d[:m], buf_v[:m]
        assert bool(((od >= 0) & (od < (1 << 28))).all())
        assert bool((_popcount(od) == 7).all())
        assert bool(torch.isfinite(ov).all()) and bool((ov.abs() > 1e-15).all())
        starts = off[a:b] - off[a]
        assert bool((od[starts] == d[a:b]).all())  # every row starts with its own determinant
        print(f"chunk {c + 1}/{len(chunks)}: rows {a}-{b}, {m} elements checked, {time.perf_counter() - t0:.1f} s",
              flush=True)
        for r in sample:
            if a <= r < b:
                s0, s1 = int(offs[r] - offs[a]), int(offs[r + 1] - offs[a])
                got[r] = (od[s0:s1].cpu().numpy().view(np.uint64), ov[s0:s1].cpu().numpy())
    assert sorted(got) == sample

    exp = Lib("oracle").ctx(ints=ints)
    exp.hb_prepare(1e-15)
    want = exp.local_connections(dets[sample], 28, True, EPS1)
    sizes = np.array([got[r][1].size for r in sample])
    have = {"offsets": np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64),
            "dets": np.concatenate([got[r][0] for r in sample]),
            "values": np.concatenate([got[r][1] for r in sample])}
    assert_conn_equal(have, want)
