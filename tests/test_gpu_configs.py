"""GPU: the batched and serving forms at their stated workloads.

Batched (64 heterogeneous instances: orbitals 16-60, electrons 8-24,
determinants log-uniform 4k-512k, heat-bath eps1 1e-7, slogdet n 4-24): one
multi-instance launch (and its HIP-graph replay, and the same with every
instance's lane-ordered H_VV) gives every sigma and energy bit-identical to
the instance alone; rows of every instance of 32k dets or fewer are checked
against the oracle (1e-12); each instance's Slater blocks against the oracle.

Serving (16 clients, 42 orbitals, 12a+12b, 120 Cholesky vectors, requests
4k-256k dets, eps1 1e-8): clients run VMC inner steps (same V under a v_key,
new amplitudes) concurrently with fresh requests; every served energy is
bit-identical to compute_variational_energy, and rows of the smallest requests
match the oracle (1e-12).
"""

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _close_rows(got, exp):
    return np.all(np.abs(got - exp) <= 1e-12 * np.maximum(np.abs(exp), 1.0))


@pytest.fixture(scope="module")
def config2():
    import torch
    from detnqs_amd import core, synth
    from detnqs_amd.batch import Instance
    specs = synth.config2_specs(64)
    out = []
    for k, sp in enumerate(specs):
        ints, dets = synth.config2_instance(sp)
        ctx = core.IntCtx.from_integrals(sp.n_orb, sp.n_elec, ints.h1e, ints.h2e, ints.e_nuc)
        ctx.hb_prepare(1e-15)
        g = torch.Generator(device="cuda").manual_seed(sp.seed)
        c = torch.randn(dets.shape[0], dtype=torch.float64, device="cuda", generator=g)
        c /= torch.linalg.norm(c)
        out.append((sp, ints, Instance.create(ctx, dets, c, sp.n_orb, True, 1e-7)))
        if k % 16 == 15:
            print(f"config2: {k + 1} instances built", flush=True)
    torch.cuda.synchronize()
    return out


def test_config2_batched_bit_identical_and_oracle(config2):
    import torch
    from detnqs_amd import core
    from detnqs_amd.batch import ElocPlan, Tuning
    from oracle.oracle import Lib
    insts = [x for _, _, x in config2]
    assert sum(x.n_rows for x in insts) > 5_000_000
    alone = []
    for x in insts:
        e, s = core.compute_variational_energy_device(x.dets, x.coeffs, x.ctx, x.n_orb, True, 1e-7, want_sigma=True)
        alone.append((e, s.clone()))
    energy = torch.empty(len(insts), dtype=torch.float64, device="cuda")
    plan = ElocPlan([(x, 0, x.n_rows, k) for k, x in enumerate(insts)], Tuning(4, True, 0))
    plan.capture(energy)
    for rep in range(2):  # graph replays, eager work in between
        for x in insts:
            x.sigma.zero_()
        plan.replay()
        plan.check()
        e = energy.cpu().numpy()
        for k, x in enumerate(insts):
            assert e[k].tobytes() == np.float64(alone[k][0]).tobytes(), (rep, k)
            assert torch.equal(x.sigma, alone[k][1]), (rep, k)
    plan.free()
    print("config2: batched graph bit-identical", flush=True)
    # the same with every instance's lane-ordered H_VV
    for x in insts:
        x.cache_vv()
    plan = ElocPlan([(x, 0, x.n_rows, k) for k, x in enumerate(insts)], Tuning(4, True, 0))
    plan.launch(energy)
    plan.check()
    e = energy.cpu().numpy()
    for k, x in enumerate(insts):
        assert e[k].tobytes() == np.float64(alone[k][0]).tobytes(), ("layout", k)
        assert torch.equal(x.sigma, alone[k][1]), ("layout", k)
    plan.free()
    for x in insts:
        x.vv = None
    print("config2: layouts bit-identical", flush=True)
    # oracle: spread rows of every instance of 32k dets or fewer
    lib = Lib("oracle")
    checked = 0
    for (sp, ints, x), (_, s) in zip(config2, alone):
        if x.n_rows > 32768:
            continue
        oc = lib.ctx(ints=ints)
        oc.hb_prepare(1e-15)
        V = x.dets.cpu().numpy().view(np.uint64)
        rows = np.unique(np.linspace(0, x.n_rows - 1, 6).astype(np.int64))
        got = s.cpu().numpy()[rows]
        exp = oc.sigma_rows(V, x.coeffs.cpu().numpy(), rows, sp.n_orb, True, 1e-7)
        assert _close_rows(got, exp), (sp, rows)
        checked += 1
    assert checked >= 10
    print(f"config2: {checked} instances checked against the oracle", flush=True)


def test_config2_slogdets(config2):
    """Slogdet sizes n = 4-24 are the instances' own draw: seeded random square
    blocks of each instance's size, checked against the oracle."""
    import torch
    from detnqs_amd.slogdet import slogdet_fwd
    from oracle.oracle import Lib
    lib = Lib("oracle")
    sizes = set()
    for sp, _, _ in config2:
        g = torch.Generator(device="cuda").manual_seed(sp.seed)
        A = torch.randn((4096, sp.slog_n, sp.slog_n), dtype=torch.float64, device="cuda", generator=g)
        s, la = slogdet_fwd(A)
        es, ela = lib.slogdet(A.cpu().numpy())
        assert np.array_equal(s.cpu().numpy(), es), sp
        assert np.all(np.abs(la.cpu().numpy() - ela) <= 1e-12 * np.maximum(1.0, np.abs(ela))), sp
        sizes.add(sp.slog_n)
    assert len(sizes) >= 15


def test_config4_serving(tmp_path):
    import torch
    from detnqs_amd import core, synth
    from detnqs_amd.serve import Server
    from oracle.oracle import Lib
    n_orb, eps1 = 42, 1e-8
    ints_all = synth.config4_clients(16)
    ctxs = []
    for ints in ints_all:
        c = core.IntCtx.from_integrals(n_orb, 24, ints.h1e, ints.h2e, ints.e_nuc)
        c.hb_prepare(1e-15)
        ctxs.append(c)
    pool = synth.config1_dets_device(262144, 4100, n_orb, 12, 12).cpu().numpy().view(np.uint64)
    rng = np.random.default_rng(4006)
    # per client: one V (v_key) for 3 inner steps, then one fresh request
    plans = []
    for k in range(16):
        size = int(round(np.exp(rng.uniform(np.log(4096), np.log(262144)))))
        lo = int(rng.integers(0, pool.shape[0] - size + 1))
        steps = [rng.standard_normal(size) for _ in range(3)]
        fsize = int(round(np.exp(rng.uniform(np.log(4096), np.log(65536)))))
        flo = int(rng.integers(0, pool.shape[0] - fsize + 1))
        plans.append((k, pool[lo:lo + size], steps, pool[flo:flo + fsize], rng.standard_normal(fsize)))
    print("config4: clients built", flush=True)
    plain = []
    for k, V, steps, F, cf in plans:
        plain.append([core.compute_variational_energy(V, c, ctxs[k], n_orb, True, eps1) for c in steps]
                     + [core.compute_variational_energy(F, cf, ctxs[k], n_orb, True, eps1)])
    got = [None] * 16
    errors = []
    with Server(64 << 30, 65536, 2) as S:
        def client(k, V, steps, F, cf):
            try:
                tf = S.submit(ctxs[k], F, cf, n_orb, True, eps1)
                out = [S.wait(S.submit(ctxs[k], V, c, n_orb, True, eps1, v_key=100 + k)).energy for c in steps]
                got[k] = out + [S.wait(tf).energy]
            except Exception as ex:
                errors.append(repr(ex))

        th = [threading.Thread(target=client, args=p) for p in plans]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = S.stats()
    assert not errors, errors
    for k in range(16):
        for j in range(4):
            assert np.float64(got[k][j]).tobytes() == np.float64(plain[k][j]).tobytes(), (k, j)
    assert st.vv_builds == 16 and st.vv_rows > 0
    print(f"config4: 64 served requests bit-identical; layouts {st.vv_builds}, layout rows {st.vv_rows}", flush=True)
    # oracle rows of the two smallest V sets
    lib = Lib("oracle")
    for k, V, steps, _, _ in sorted(plans, key=lambda p: p[1].shape[0])[:2]:
        oc = lib.ctx(ints=ints_all[k])
        oc.hb_prepare(1e-15)
        _, s = core.compute_variational_energy_device(V, steps[0], ctxs[k], n_orb, True, eps1, want_sigma=True)
        rows = np.unique(np.linspace(0, V.shape[0] - 1, 8).astype(np.int64))
        exp = oc.sigma_rows(V, steps[0], rows, n_orb, True, eps1)
        assert _close_rows(s.cpu().numpy()[rows], exp), k
