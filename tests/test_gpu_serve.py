"""GPU: serving form (continuous-batching local-energy server).

Every request's energy must be bit-identical to compute_variational_energy on
the same inputs, whether it runs alone or in a mix of concurrent requests from
several clients (own integral contexts), whatever the launch size, the number
of launches in flight, the HBM budget (admission), V caching, or where the
inputs live.
"""

import threading

import numpy as np
import pytest

from helpers import H2O_STO3G, fci_dets, write_tmp_fcidump

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def clients(tmp_path_factory):
    from detnqs_amd import core, synth
    ints0, dets0 = synth.config0()
    p0 = write_tmp_fcidump(tmp_path_factory.mktemp("s0"), ints0)
    c0 = core.IntCtx(p0, 13)
    c0.hb_prepare(1e-15)
    ints1 = synth.make_integrals(16, 12, 30, seed=91)
    c1 = core.IntCtx.from_integrals(16, 12, ints1.h1e, ints1.h2e, ints1.e_nuc)
    c1.hb_prepare(1e-15)
    d1 = synth.config1_dets(3000, seed=8, n_orb=16, n_a=6, n_b=6)
    c2 = core.IntCtx(H2O_STO3G, 7)
    F = fci_dets(7, 5, 5)
    return [(c0, dets0, 13, True, 1e-5), (c1, d1, 16, True, 1e-6), (c2, F, 7, False, 0.0)]


def _requests(clients, n_req, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        ctx, dets, n_orb, hb, eps = clients[k % len(clients)]
        n = int(rng.integers(1, dets.shape[0] + 1))
        lo = int(rng.integers(0, dets.shape[0] - n + 1))
        c = rng.standard_normal(n)
        out.append((ctx, dets[lo:lo + n], c, n_orb, hb, eps))
    return out


def _plain(req):
    from detnqs_amd import core
    ctx, d, c, n_orb, hb, eps = req
    return core.compute_variational_energy(d, c, ctx, n_orb, hb, eps)


def _bits(x):
    return np.float64(x).tobytes()


@pytest.mark.parametrize("rows_per_launch,inflight", [(32768, 2), (700, 1), (257, 3)])
def test_alone_and_mix_identical(clients, rows_per_launch, inflight):
    from detnqs_amd.serve import Server
    reqs = _requests(clients, 18, seed=rows_per_launch)
    plain = [_plain(r) for r in reqs]
    with Server(1 << 30, rows_per_launch, inflight) as S:
        alone = []
        for ctx, d, c, n_orb, hb, eps in reqs:
            alone.append(S.wait(S.submit(ctx, d, c, n_orb, hb, eps)).energy)
        # the mix: three client threads submitting concurrently, then waiting
        got = [None] * len(reqs)

        def client(ids):
            tickets = [(i, S.submit(*reqs[i][:4], reqs[i][4], reqs[i][5])) for i in ids]
            for i, t in tickets:
                got[i] = S.wait(t).energy

        th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 3)),)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = S.stats()
    for k in range(len(reqs)):
        assert _bits(alone[k]) == _bits(plain[k]), k
        assert _bits(got[k]) == _bits(plain[k]), k
    assert st.completed == 2 * len(reqs)
    assert st.arena_used == 0


def test_admission_by_budget(clients):
    """A budget that holds about two requests at a time: everything completes,
    the arena never exceeds the budget, and an oversize request is refused."""
    from detnqs_amd.serve import Server
    ctx, dets, n_orb, hb, eps = clients[0]
    n = 2000
    budget = 2 * Server.request_bytes(n) + 4096
    reqs = [(ctx, dets[k * 500:k * 500 + n], np.random.default_rng(k).standard_normal(n), n_orb, hb, eps)
            for k in range(8)]
    with Server(budget, 1000, 2) as S:
        tickets = [S.submit(*r) for r in reqs]
        res = [S.wait(t) for t in tickets]
        st = S.stats()
        with pytest.raises(IndexError):
            S.submit(ctx, dets[:7000], np.ones(7000), n_orb, hb, eps)
    assert st.arena_peak <= st.arena_bytes
    for r, x in zip(reqs, res):
        assert _bits(x.energy) == _bits(_plain(r))
        assert x.t_submit <= x.t_admit <= x.t_first_launch <= x.t_done
    # later requests were admitted only once earlier ones had finished
    first_done = min(x.t_done for x in res)
    assert sum(x.t_admit >= first_done for x in res) >= 4


def test_v_cache_and_device_inputs(clients):
    import torch
    from detnqs_amd import core
    from detnqs_amd.serve import Server
    ctx, dets, n_orb, hb, eps = clients[1]
    V = dets[:2500]
    rng = np.random.default_rng(3)
    cs = [rng.standard_normal(2500) for _ in range(3)]
    with Server(1 << 28, 4096, 2) as S:
        e = [S.wait(S.submit(ctx, V, c, n_orb, hb, eps, v_key=42)).energy for c in cs]
        Vd = core.dets_to_device(V)
        cd = torch.from_numpy(cs[0]).cuda()
        e_dev = S.wait(S.submit(ctx, Vd, cd, n_orb, hb, eps, v_key=42)).energy
        st = S.stats()
    for c, x in zip(cs, e):
        assert _bits(x) == _bits(core.compute_variational_energy(V, c, ctx, n_orb, hb, eps))
    assert _bits(e_dev) == _bits(e[0])
    assert st.v_misses == 1 and st.v_hits == 3 and st.v_cached == 1


def test_errors_and_empty(clients):
    from detnqs_amd.serve import Server
    ctx, dets, n_orb, hb, eps = clients[2]
    bad = np.array([[0b11111, 0b11111], [0b11111 | (1 << 9), 0b11111]], dtype=np.uint64)
    with Server(1 << 26, 1024, 2) as S:
        t_bad = S.submit(ctx, bad, np.ones(2), n_orb, False, 0.0)
        t_ok = S.submit(ctx, dets[:50], np.ones(50), n_orb, False, 0.0)
        t_empty = S.submit(ctx, dets[:0], np.ones(0), n_orb, False, 0.0)
        with pytest.raises(ValueError, match="n_orb"):
            S.wait(t_bad)
        assert _bits(S.wait(t_ok).energy) == _bits(_plain((ctx, dets[:50], np.ones(50), n_orb, False, 0.0)))
        assert S.wait(t_empty).energy == 0.0
        with pytest.raises(ValueError, match="Heat-Bath"):
            S.submit(ctx, dets[:5], np.ones(5), n_orb, True, 1e-6)
        with pytest.raises(ValueError, match="ticket"):
            S.wait(t_ok)


def test_walk_method_requests(clients):
    from detnqs_amd import core
    from detnqs_amd.serve import Server
    ctx, dets, n_orb, hb, eps = clients[0]
    c = np.random.default_rng(5).standard_normal(1200)
    with Server(1 << 27, 500, 2) as S:
        e_w = S.wait(S.submit(ctx, dets[:1200], c, n_orb, hb, eps, method="walk")).energy
        e_s = S.wait(S.submit(ctx, dets[:1200], c, n_orb, hb, eps)).energy
    assert _bits(e_w) == _bits(core.compute_variational_energy(dets[:1200], c, ctx, n_orb, hb, eps, method="walk"))
    assert _bits(e_s) == _bits(core.compute_variational_energy(dets[:1200], c, ctx, n_orb, hb, eps))
    assert abs(e_w - e_s) <= 1e-12 * abs(e_w)


def test_shared_context_threads_and_server(clients, tmp_path):
    """Several threads on one context, each on its own stream, each asking for
    heat-bath cutoffs the context has not seen yet (so the kept-prefix tables
    are built concurrently), while a server runs requests on the same context:
    every result equals the serial calls bit for bit."""
    import torch
    from detnqs_amd import core, synth
    from detnqs_amd.serve import Server
    ints, dets = synth.config0()
    path = write_tmp_fcidump(tmp_path, ints)
    serial = core.IntCtx(path, 13)
    serial.hb_prepare(1e-15)
    eps_list = [1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6]
    sub = dets[:2500]
    exp_off = {e: core.local_conn_count(core.dets_to_device(sub), serial, 13, True, e)[0].cpu().numpy()
               for e in eps_list}
    rng = np.random.default_rng(17)
    coefs = [rng.standard_normal(1500) for _ in range(6)]
    exp_e = [core.compute_variational_energy(dets[k * 700:k * 700 + 1500], coefs[k], serial, 13, True, 2e-6)
             for k in range(6)]

    shared = core.IntCtx(path, 13)
    shared.hb_prepare(1e-15)
    got_off, errors = {}, []

    def worker(eps_sub):
        try:
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                d = core.dets_to_device(sub)
                for e in eps_sub:
                    off, _ = core.local_conn_count(d, shared, 13, True, e)
                    st.synchronize()
                    got_off[e] = off.cpu().numpy()
        except Exception as ex:  # reported below
            errors.append(repr(ex))

    with Server(1 << 28, 512, 2) as S:
        tickets = [S.submit(shared, dets[k * 700:k * 700 + 1500], coefs[k], 13, True, 2e-6) for k in range(6)]
        th = [threading.Thread(target=worker, args=(eps_list[k::3],)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        served = [S.wait(t).energy for t in tickets]
    assert not errors, errors
    for e in eps_list:
        assert np.array_equal(got_off[e], exp_off[e]), e
    for k in range(6):
        assert _bits(served[k]) == _bits(exp_e[k]), k


def test_vmc_inner_steps_use_the_layout(clients):
    """A V kept under a v_key for several requests (VMC inner steps: same V,
    new amplitudes): its lane-ordered H_VV is built once, later requests are a
    product over it, and every energy is bit-identical to
    compute_variational_energy, sequentially (closed loop) and in a mix; a
    request on the same V with other screening takes the string path."""
    from detnqs_amd.serve import Server
    ctx, dets, n_orb, hb, eps = clients[1]
    V = dets[:2500]
    rng = np.random.default_rng(31)
    cs = [rng.standard_normal(2500) for _ in range(6)]
    cs[2][::50] = 1e-7  # skipped rows
    plain = [_plain((ctx, V, c, n_orb, hb, eps)) for c in cs]
    other = _plain((ctx, V, cs[0], n_orb, hb, 1e-3))
    for layout in (True, False):
        with Server(1 << 28, 700, 2, vv_layout=layout) as S:
            seq = [S.wait(S.submit(ctx, V, c, n_orb, hb, eps, v_key=9)).energy for c in cs]
            mix_t = [S.submit(ctx, V, c, n_orb, hb, eps, v_key=9) for c in cs]
            t_other = S.submit(ctx, V, cs[0], n_orb, hb, 1e-3, v_key=9)
            mix = [S.wait(t).energy for t in mix_t]
            e_other = S.wait(t_other).energy
            st = S.stats()
        for k in range(len(cs)):
            assert _bits(seq[k]) == _bits(plain[k]), (layout, k)
            assert _bits(mix[k]) == _bits(plain[k]), (layout, k)
        assert _bits(e_other) == _bits(other)
        if layout:
            assert st.vv_builds == 1
            assert st.vv_rows == 2 * len(cs) * 2500
        else:
            assert st.vv_builds == 0 and st.vv_rows == 0
        assert st.arena_used > 0  # the keyed V (and its layout) stay resident
