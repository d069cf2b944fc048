"""GPU: multi-instance fused local energy (batched form).

Instances of different orbital counts, electron counts, screening modes and
sizes run in one launch; every sigma_i and every energy must be bit-identical
to the single-instance path (compute_variational_energy) on that instance
alone, whatever the launch shape (persistent work queue or one wave per row,
1-8 waves per block), the split of an instance into row ranges over several
launches, or HIP-graph replay.  The single-instance path is itself checked
against the oracle to 1e-12 (test_gpu_project.py); one energy is re-checked
against the oracle here.
"""

import numpy as np
import pytest

from helpers import H2O_STO3G, fci_dets, write_tmp_fcidump

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances(tmp_path_factory):
    import torch
    from detnqs_amd import core, synth
    from detnqs_amd.batch import Instance

    out = []
    ints0, dets0 = synth.config0()
    p0 = write_tmp_fcidump(tmp_path_factory.mktemp("b0"), ints0)
    c0 = core.IntCtx(p0, 13)
    c0.hb_prepare(1e-15)
    rng = np.random.default_rng(41)

    def coeffs(n):
        c = rng.standard_normal(n)
        c /= np.linalg.norm(c)
        if n > 20:
            c[7] = 1e-7  # skipped row (c^2 < 1e-12)
        return c

    out.append(("c0_hb", Instance.create(c0, dets0[:3000], coeffs(3000), 13, True, 1e-5), p0, 13))
    out.append(("c0_full", Instance.create(c0, dets0[3000:4500], coeffs(1500), 13, False, 0.0), p0, 13))
    ints1 = synth.make_integrals(16, 12, 30, seed=77)
    p1 = write_tmp_fcidump(tmp_path_factory.mktemp("b1"), ints1)
    c1 = core.IntCtx.from_integrals(16, 12, ints1.h1e, ints1.h2e, ints1.e_nuc)
    c1.hb_prepare(1e-15)
    d1 = synth.config1_dets(2500, seed=5, n_orb=16, n_a=6, n_b=6)
    out.append(("n16_hb", Instance.create(c1, d1, coeffs(d1.shape[0]), 16, True, 1e-6), p1, 16))
    out.append(("n16_hb_loose", Instance.create(c1, d1[:700], coeffs(700), 16, True, 1e-3), p1, 16))
    c2 = core.IntCtx(H2O_STO3G, 7)
    F = fci_dets(7, 5, 5)
    out.append(("h2o_fci", Instance.create(c2, F, coeffs(F.shape[0]), 7, False, 0.0), H2O_STO3G, 7))
    out.append(("tiny", Instance.create(c2, F[:3], coeffs(3), 7, False, 0.0), H2O_STO3G, 7))
    out.append(("c0_hb_walk", Instance.create(c0, dets0[500:2500], coeffs(2000), 13, True, 1e-5, method="walk"),
                p0, 13))
    torch.cuda.synchronize()
    return out


def _alone(inst):
    from detnqs_amd import core
    e, s = core.compute_variational_energy_device(inst.dets, inst.coeffs, inst.ctx, inst.n_orb, inst.use_hb,
                                                  inst.eps1, want_sigma=True, method=inst.method)
    return e, s.clone()


@pytest.fixture(scope="module")
def alone(instances):
    return [_alone(x) for _, x, _, _ in instances]


def _check(instances, alone, energy):
    import torch
    e = energy.cpu().numpy()
    for k, ((name, inst, _, _), (ea, sa)) in enumerate(zip(instances, alone)):
        assert torch.equal(inst.sigma, sa), name
        assert e[k].tobytes() == np.float64(ea).tobytes(), (name, e[k], ea)


@pytest.mark.parametrize("wpb,persistent", [(1, True), (2, True), (4, True), (8, True), (4, False), (1, False)])
def test_one_launch_matches_alone(instances, alone, wpb, persistent):
    import torch
    from detnqs_amd.batch import Tuning, eloc_batched
    for _, x, _, _ in instances:
        x.sigma.fill_(float("nan"))
    e = eloc_batched([x for _, x, _, _ in instances], Tuning(wpb, persistent, 0))
    torch.cuda.synchronize()
    _check(instances, alone, e)


def test_small_grid_drains_queue(instances, alone):
    """A persistent grid far smaller than the row count still covers every row."""
    from detnqs_amd.batch import Tuning, eloc_batched
    for _, x, _, _ in instances:
        x.sigma.fill_(float("nan"))
    e = eloc_batched([x for _, x, _, _ in instances], Tuning(2, True, 3))
    _check(instances, alone, e)


def test_row_ranges_over_launches(instances, alone):
    """Each instance split in three ranges, interleaved over three launches; the
    energies are formed in the last one."""
    import torch
    from detnqs_amd.batch import ElocPlan, Tuning
    insts = [x for _, x, _, _ in instances]
    for x in insts:
        x.sigma.fill_(float("nan"))
    cuts = [(0, x.n_rows // 3, 2 * x.n_rows // 3, x.n_rows) for x in insts]
    energy = torch.full((len(insts),), float("nan"), dtype=torch.float64, device="cuda")
    for part in range(3):
        ranges = []
        for k, (x, c) in enumerate(zip(insts, cuts)):
            ranges.append((x, c[part], c[part + 1], k if part == 2 else -1))
        plan = ElocPlan(ranges[::-1] if part == 1 else ranges, Tuning(4, True, 0))
        plan.launch(energy)
        plan.check()
        plan.free()
    _check(instances, alone, energy)


def test_graph_replays(instances, alone):
    """Captured launch replayed three times with eager work in between; every
    replay gives the alone results."""
    import torch
    from detnqs_amd.batch import ElocPlan, Tuning
    insts = [x for _, x, _, _ in instances]
    energy = torch.empty(len(insts), dtype=torch.float64, device="cuda")
    plan = ElocPlan([(x, 0, x.n_rows, k) for k, x in enumerate(insts)], Tuning(4, True, 0))
    plan.capture(energy)
    for rep in range(3):
        for x in insts:
            x.sigma.fill_(float("nan"))
        energy.fill_(float("nan"))
        _ = torch.randn(1 << 16, device="cuda").sum()  # unrelated eager work
        plan.replay()
        torch.cuda.synchronize()
        _check(instances, alone, energy)
    plan.check()
    plan.free()


def test_energy_vs_oracle(instances):
    from oracle.oracle import Lib
    name, inst, path, n = instances[2]
    from detnqs_amd.batch import eloc_batched
    e = eloc_batched([inst]).cpu().numpy()[0]
    octx = Lib("oracle").ctx(path, n)
    octx.hb_prepare(1e-15)
    eo = octx.variational_energy(inst.dets.cpu().numpy().view(np.uint64), inst.coeffs.cpu().numpy(), n, True,
                                 inst.eps1)
    assert abs(e - eo) <= 1e-12 * abs(eo)


def test_errors(instances):
    import torch
    from detnqs_amd import core
    from detnqs_amd.batch import ElocPlan, Instance, Tuning
    _, inst, _, _ = instances[5]
    with pytest.raises(ValueError, match="waves_per_block"):
        ElocPlan([(inst, 0, 3, 0)], Tuning(3, True, 0))
    with pytest.raises(ValueError, match="row range"):
        ElocPlan([(inst, 0, 4, 0)])
    bad = np.array([[0b11111, 0b11111], [0b11111 | (1 << 9), 0b11111]], dtype=np.uint64)
    x = Instance.create(inst.ctx, bad, np.ones(2), 7, False, 0.0)
    plan = ElocPlan([(x, 0, 2, 0)])
    plan.launch(torch.empty(1, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError, match="n_orb"):
        plan.check()
    plan.free()
    c = core.IntCtx(H2O_STO3G, 7)
    with pytest.raises(ValueError, match="Heat-Bath"):
        Instance.create(c, fci_dets(7, 5, 5)[:4], np.ones(4), 7, True, 1e-6)
