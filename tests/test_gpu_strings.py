"""GPU: V-driven local energy (string index of V) against the candidate walk and
the oracle.

Both methods must find the same (y, H_xy) pairs with the same screening, so
sigma and the energy agree to rounding: 1e-12 relative (per row, relative to
max(|sigma|, 1)).  Cases: exhaustive and heat-bath doubles, a heat-bath table
built with a coarse threshold (the table, not eps1, decides), duplicates in V
(first occurrence counts), an instance using fewer orbitals than its context,
and a larger 28-orbital space where most rows have many partners in V.
"""

import numpy as np
import pytest

from helpers import write_tmp_fcidump

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0(tmp_path_factory):
    from detnqs_amd import core, synth
    from oracle.oracle import Lib
    ints, dets = synth.config0()
    path = write_tmp_fcidump(tmp_path_factory.mktemp("st"), ints)
    ctx = core.IntCtx(path, 13)
    ctx.hb_prepare(1e-15)
    octx = Lib("oracle").ctx(path, 13)
    octx.hb_prepare(1e-15)
    return path, dets, ctx, octx


def _both(V, c, ctx, n_orb, hb, eps):
    from detnqs_amd import core
    es, ss = core.compute_variational_energy_device(V, c, ctx, n_orb, hb, eps, want_sigma=True, method="strings")
    ew, sw = core.compute_variational_energy_device(V, c, ctx, n_orb, hb, eps, want_sigma=True, method="walk")
    return es, ss.cpu().numpy(), ew, sw.cpu().numpy()


def _close(a, b):
    return np.all(np.abs(a - b) <= 1e-12 * np.maximum(np.abs(b), 1.0))


@pytest.mark.parametrize("hb,eps", [(False, 0.0), (True, 1e-4), (True, 0.0), (True, 1e-2)])
def test_strings_vs_walk_and_oracle(c0, hb, eps):
    _, dets, ctx, octx = c0
    V = dets[:4000]
    c = np.random.default_rng(11).standard_normal(4000)
    c /= np.linalg.norm(c)
    c[5] = 1e-7
    es, ss, ew, sw = _both(V, c, ctx, 13, hb, eps)
    eo, so = octx.variational_energy(V, c, 13, hb, eps, want_sigma=True)
    assert _close(ss, sw) and _close(ss, so)
    assert abs(es - eo) <= 1e-12 * abs(eo) and abs(es - ew) <= 1e-12 * abs(ew)
    assert ss[5] == 0.0


def test_coarse_heatbath_table(c0):
    """hb_prepare(1e-2): doubles below the table threshold are absent even with eps1 = 0."""
    from detnqs_amd import core
    from oracle.oracle import Lib
    path, dets, _, _ = c0
    ctx = core.IntCtx(path, 13)
    ctx.hb_prepare(1e-2)
    octx = Lib("oracle").ctx(path, 13)
    octx.hb_prepare(1e-2)
    V = dets[1000:4000]
    c = np.random.default_rng(12).standard_normal(3000)
    for eps in (0.0, 1e-3, 5e-2):
        es, ss, ew, sw = _both(V, c, ctx, 13, True, eps)
        eo, so = octx.variational_energy(V, c, 13, True, eps, want_sigma=True)
        assert _close(ss, so) and _close(ss, sw), eps
        assert abs(es - eo) <= 1e-12 * abs(eo)


def test_duplicates_first_occurrence(c0):
    _, dets, ctx, octx = c0
    V = np.concatenate([dets[:600], dets[100:250], dets[:3]])
    c = np.random.default_rng(13).standard_normal(V.shape[0])
    es, ss, ew, sw = _both(V, c, ctx, 13, True, 1e-5)
    eo, so = octx.variational_energy(V, c, 13, True, 1e-5, want_sigma=True)
    assert _close(ss, so) and _close(ss, sw)
    assert abs(es - eo) <= 1e-12 * abs(eo)


def test_duplicates_inside_v(c0):
    """Duplicates before other rows: the reference's tolerant DetMap numbers
    first occurrences consecutively and reads coeffs[that number]
    (det_space.cpp:40-54, local_conn.cpp:192-194); both methods follow it."""
    _, dets, ctx, octx = c0
    V = np.concatenate([dets[:900], dets[100:400], dets[900:1500], dets[50:60], dets[1500:2000]])
    c = np.random.default_rng(16).standard_normal(V.shape[0])
    c[[3, 950, 1300]] = 1e-7
    for hb, eps in ((True, 1e-6), (False, 0.0)):
        es, ss, ew, sw = _both(V, c, ctx, 13, hb, eps)
        eo, so = octx.variational_energy(V, c, 13, hb, eps, want_sigma=True)
        assert _close(ss, so) and _close(sw, so), hb
        assert abs(es - eo) <= 1e-12 * abs(eo) and abs(ew - eo) <= 1e-12 * abs(eo), hb


def test_fewer_orbitals_than_context(c0):
    _, dets, ctx, octx = c0
    keep = ((dets[:, 0] | dets[:, 1]) >> np.uint64(12)) == 0
    V = dets[keep][:2000]
    c = np.random.default_rng(14).standard_normal(V.shape[0])
    for hb, eps in ((False, 0.0), (True, 1e-6)):
        es, ss, ew, sw = _both(V, c, ctx, 12, hb, eps)
        eo, so = octx.variational_energy(V, c, 12, hb, eps, want_sigma=True)
        assert _close(ss, so) and _close(ss, sw)


def test_n28_dense_partners():
    from detnqs_amd import core, synth
    ints = synth.make_integrals(28, 14, 80, seed=2001)
    ctx = core.IntCtx.from_integrals(28, 14, ints.h1e, ints.h2e, ints.e_nuc)
    ctx.hb_prepare(1e-15)
    V = synth.config1_dets(30000, seed=21)
    c = np.random.default_rng(15).standard_normal(V.shape[0])
    c /= np.linalg.norm(c)
    es, ss, ew, sw = _both(V, c, ctx, 28, True, 1e-7)
    assert _close(ss, sw)
    assert abs(es - ew) <= 1e-12 * abs(ew)


def test_unknown_method(c0):
    from detnqs_amd import core
    _, dets, ctx, _ = c0
    with pytest.raises(ValueError):
        core.compute_variational_energy(dets[:10], np.ones(10), ctx, 13, method="bogus")
