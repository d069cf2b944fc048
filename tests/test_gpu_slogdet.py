"""GPU parity: batched FP64 slogdet forward/backward against the oracle
(LAPACK-getf2 restatement) and numpy.linalg.  Tolerances: log|det| within
1e-12 relative (|logabs| >= 1) or 1e-12 absolute; sign exact; the forward LU
is the oracle's operation sequence, so sign and pivots match exactly;
gradients within 1e-10 relative to max|grad| per matrix for well-conditioned
inputs."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(B, n, seed, dominant=False):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((B, n, n))
    if dominant:
        A += n * np.eye(n)
    return A


def _to(A):
    import torch
    return torch.from_numpy(np.ascontiguousarray(A)).cuda()


def _close_logabs(got, exp):
    scale = np.maximum(np.abs(exp), 1.0)
    return np.all(np.abs(got - exp) <= 1e-12 * scale)


@pytest.mark.parametrize("n", [2, 3, 5, 7, 8, 9, 12, 16, 17, 24, 31, 32, 33, 48, 64])
def test_fwd_matches_oracle_and_numpy(n):
    from detnqs_amd.slogdet import slogdet_fwd
    from oracle.oracle import Lib
    A = _rand(257, n, seed=n)
    s, la = slogdet_fwd(_to(A))
    s, la = s.cpu().numpy(), la.cpu().numpy()
    os_, ola = Lib("oracle").slogdet(A)
    assert np.array_equal(s, os_)
    assert _close_logabs(la, ola)
    ns, nla = np.linalg.slogdet(A)
    assert np.array_equal(s, ns)
    assert np.allclose(la, nla, rtol=1e-11, atol=1e-11)


def test_fwd_config1_shape_and_edge_cases():
    from detnqs_amd.slogdet import slogdet_fwd
    from oracle.oracle import Lib
    n = 7
    A = _rand(5000, n, seed=1)
    A[0] = 0.0                                     # singular: (0, -inf)
    A[1, 3] = A[1, 5]                               # exactly repeated row
    A[2] = np.eye(n)[np.random.default_rng(0).permutation(n)]  # permutation
    A[3] = np.triu(A[3])                            # triangular: product of diag
    A[4, 2, :] *= 1e-8                              # near-singular, finite
    A[5, 0, 0] = np.nan
    A[6, 1, 1] = np.inf
    s, la = (t.cpu().numpy() for t in slogdet_fwd(_to(A)))
    os_, ola = Lib("oracle").slogdet(A)
    assert np.array_equal(s, os_)
    fin = np.isfinite(ola)
    assert np.array_equal(np.isfinite(la), fin)
    assert _close_logabs(la[fin], ola[fin])
    assert s[0] == 0 and la[0] == -np.inf
    assert s[1] == 0 and la[1] == -np.inf
    assert abs(la[2]) < 1e-15 and s[2] == np.linalg.det(A[2])
    d = np.diag(A[3])
    assert np.isclose(la[3], np.sum(np.log(np.abs(d))), rtol=1e-13) and s[3] == np.prod(np.sign(d))
    assert s[5] == 0 and la[5] == -np.inf and s[6] == 0 and la[6] == -np.inf


def test_fwd_full_benchmark_batch():
    """The benchmark's whole batch: 2,000,000 FP64 7x7 Slater blocks (the
    alpha and beta occupied rows of a random 28x7 orbital matrix plus 0.05
    N(0,1) per sample, for the 1M config-1 determinants), every matrix against
    the oracle: sign exact, log|det| within 1e-12."""
    from detnqs_amd import synth
    from detnqs_amd.slogdet import slogdet_fwd
    from oracle.oracle import Lib
    dets = synth.config1_dets(1_000_000, seed=1)
    A = synth.slater_matrices(dets, 28, 7, 7, seed=3001, noise=0.05)
    assert A.shape == (2_000_000, 7, 7)
    s, la = (t.cpu().numpy() for t in slogdet_fwd(_to(A)))
    os_, ola = Lib("oracle").slogdet(A)
    assert np.array_equal(s, os_)
    assert np.all(np.isfinite(la)) and _close_logabs(la, ola)


def test_padding_invariance():
    """det(blockdiag(A, I)) = det(A): n=7 vs the same matrices padded to 12."""
    from detnqs_amd.slogdet import slogdet_fwd
    A = _rand(300, 7, seed=3)
    P = np.zeros((300, 12, 12))
    P[:, :7, :7] = A
    P[:, 7:, 7:] = np.eye(5)
    s1, l1 = (t.cpu().numpy() for t in slogdet_fwd(_to(A)))
    s2, l2 = (t.cpu().numpy() for t in slogdet_fwd(_to(P)))
    assert np.array_equal(s1, s2)
    assert np.allclose(l1, l2, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [2, 5, 7, 8, 12, 16, 24, 32])
def test_bwd_matches_oracle(n):
    from detnqs_amd.slogdet import slogdet_bwd_from_a, slogdet_bwd_from_lu, slogdet_fwd_lu
    from oracle.oracle import Lib
    A = _rand(129, n, seed=100 + n, dominant=True)
    cot = np.random.default_rng(n).standard_normal(129)
    G = slogdet_bwd_from_a(_to(A), None, _to(cot)).cpu().numpy()
    EG = Lib("oracle").slogdet_bwd(A, cot)
    scale = np.abs(EG).reshape(129, -1).max(axis=1)[:, None, None]
    assert np.all(np.abs(G - EG) <= 1e-10 * scale)
    # LU-reuse path gives the same numbers as refactorising
    _, _, lu, perm = slogdet_fwd_lu(_to(A))
    G2 = slogdet_bwd_from_lu(lu, perm, _to(cot)).cpu().numpy()
    assert np.all(np.abs(G2 - EG) <= 1e-10 * scale)
    # closed form: cot * inv(A)^T
    ref = cot[:, None, None] * np.transpose(np.linalg.inv(A), (0, 2, 1))
    assert np.all(np.abs(G - ref) <= 1e-9 * scale)


def test_bwd_singular_is_zero():
    from detnqs_amd.slogdet import slogdet_bwd_from_a
    A = _rand(4, 6, seed=9)
    A[1] = 0.0
    A[2, 4] = A[2, 1]
    G = slogdet_bwd_from_a(_to(A), None, _to(np.ones(4))).cpu().numpy()
    assert np.all(G[1] == 0) and np.all(G[2] == 0) and np.any(G[0] != 0)


def test_autograd_central_fd():
    import torch
    from detnqs_amd.slogdet import slogdet
    A = _to(_rand(64, 7, seed=5, dominant=True)).requires_grad_(True)
    cot = torch.from_numpy(np.random.default_rng(1).standard_normal(64)).cuda()
    s, la = slogdet(A)
    (la * cot).sum().backward()
    g = A.grad.cpu().numpy()
    A0 = A.detach().cpu().numpy()
    h = 1e-6
    rng = np.random.default_rng(2)
    for _ in range(40):
        b, i, j = rng.integers(64), rng.integers(7), rng.integers(7)
        Ap, Am = A0[b].copy(), A0[b].copy()
        Ap[i, j] += h
        Am[i, j] -= h
        fd = (np.linalg.slogdet(Ap)[1] - np.linalg.slogdet(Am)[1]) / (2 * h) * cot.cpu().numpy()[b]
        assert abs(fd - g[b, i, j]) <= 1e-6 * max(1.0, abs(fd))


def test_dispatch_and_errors():
    import torch
    from detnqs_amd.slogdet import slogdet, slogdet_fwd
    A = _to(_rand(3, 70, seed=1, dominant=True))
    s, la = slogdet(A)  # n > 64 -> library routine, as the reference
    assert torch.allclose(la, torch.linalg.slogdet(A)[1])
    with pytest.raises(ValueError):
        slogdet_fwd(_to(_rand(2, 3, seed=0))[:, :, :2])
    with pytest.raises(ValueError):
        slogdet_fwd(_to(np.zeros((2, 1, 1))))


@pytest.mark.parametrize("n", [2, 5, 7, 8, 9, 16, 24, 33, 64])
def test_f32_targets(n):
    """FP32 entry points: the FP32 input is computed in FP64 and the results are
    rounded once, so they equal the oracle's FP64 results (on the same values)
    rounded to FP32, within one FP32 ulp (the device log may differ from
    glibc's by an FP64 ulp, which can move a value across an FP32 rounding
    boundary)."""
    import torch
    from detnqs_amd.slogdet import slogdet_bwd_from_a, slogdet_fwd
    from oracle.oracle import Lib
    O = Lib("oracle")
    A32 = _rand(301, n, seed=100 + n, dominant=(n > 8)).astype(np.float32)
    A64 = A32.astype(np.float64)
    A32[7] = 0.0  # singular: (0, -inf) and a zero gradient
    A64[7] = 0.0
    s, la = slogdet_fwd(_to(A32))
    assert s.dtype == torch.float32 and la.dtype == torch.float32
    s, la = s.cpu().numpy(), la.cpu().numpy()
    es, ela = O.slogdet(A64)
    assert np.array_equal(s, es.astype(np.float32))
    fin = np.isfinite(ela)
    assert not fin[7] and la[7] == -np.inf
    assert np.all(np.abs(la[fin] - ela[fin].astype(np.float32)) <= np.spacing(np.abs(ela[fin]).astype(np.float32)))
    cot = np.random.default_rng(n).standard_normal(301).astype(np.float32)
    g = slogdet_bwd_from_a(_to(A32), None, torch.from_numpy(cot).cuda()).cpu().numpy()
    assert g.dtype == np.float32
    eg = O.slogdet_bwd(A64, cot.astype(np.float64)).astype(np.float32)
    scale = np.abs(eg).reshape(301, -1).max(axis=1)[:, None, None] + 1e-30
    assert np.all(np.abs(g - eg) <= 2e-7 * scale)
    assert np.all(g[7] == 0.0)


@pytest.mark.parametrize("n", [9, 11, 12, 13, 15, 16, 17, 20, 24, 25, 31, 32, 40, 48, 49, 64])
def test_every_tile_matches_oracle(n):
    """Every tile the n > 8 kernels are built for (lanes per matrix, and at
    n = 12, 16, 24, 32 the kernels with fixed and with run-time bounds): forward,
    forward + LU, both backwards and the FP32 forward and backward, on random matrices plus
    the edge cases (zero matrix, a zero row, a column scaled to subnormal
    pivots so the multipliers divide instead of using the reciprocal, NaN,
    near-singular).  Sign exact and log|det| within 1e-12 of the oracle; the
    LU, permutation and gradients bit-identical across tiles (each tile runs
    the same operation sequence per element)."""
    import torch
    from detnqs_amd.slogdet import (get_tile, set_tile, slogdet_bwd_from_a, slogdet_bwd_from_lu,
                                    slogdet_fwd, slogdet_fwd_lu, tile_choices)
    from oracle.oracle import Lib
    O = Lib("oracle")
    B = 203
    A = _rand(B, n, seed=7000 + n, dominant=True)
    A[3] = 0.0
    A[4, n - 2] = 0.0
    A[5, :, 0] *= 1e-310
    A[6, 2, 2] = np.nan
    A[7, 1, :] *= 1e-8
    cot = np.random.default_rng(n).standard_normal(B)
    es, ela = O.slogdet(A)
    fin = np.isfinite(ela)
    EG = O.slogdet_bwd(A, cot)
    ok = np.isfinite(EG).reshape(B, -1).all(axis=1)
    scale = np.abs(EG).reshape(B, -1).max(axis=1)[:, None, None]
    default = [get_tile(kind, n) for kind in range(4)]
    first = None
    try:
        for lanes in tile_choices(n):
            for kind in range(4):
                set_tile(kind, n, lanes)
            s, la = (t.cpu().numpy() for t in slogdet_fwd(_to(A)))
            assert np.array_equal(s[fin], es[fin]), lanes
            assert np.array_equal(np.isfinite(la), fin), lanes
            assert _close_logabs(la[fin], ela[fin]), lanes
            s2, la2, lu, perm = slogdet_fwd_lu(_to(A))
            assert np.array_equal(s2.cpu().numpy(), s) and la2.cpu().numpy().tobytes() == la.tobytes(), lanes
            G = slogdet_bwd_from_a(_to(A), None, _to(cot)).cpu().numpy()
            G2 = slogdet_bwd_from_lu(lu, perm, _to(cot)).cpu().numpy()
            assert G.tobytes() == G2.tobytes(), lanes
            assert np.all(np.abs(G[ok] - EG[ok]) <= 1e-10 * scale[ok]), lanes
            assert np.all(G[3] == 0.0) and np.all(G[4] == 0.0), lanes
            A32 = A.astype(np.float32)
            s32, la32 = (t.cpu().numpy() for t in slogdet_fwd(_to(A32)))
            assert s32.dtype == np.float32, lanes
            # FP32 in, FP64 inside: the FP64 results on the same values, rounded once
            es32, ela32 = O.slogdet(A32.astype(np.float64))
            f32 = np.isfinite(ela32)
            assert np.array_equal(s32[f32], es32[f32].astype(np.float32)), lanes
            assert np.all(np.abs(la32[f32] - ela32[f32].astype(np.float32))
                          <= np.spacing(np.abs(ela32[f32]).astype(np.float32))), lanes
            g32 = slogdet_bwd_from_a(_to(A32), None, torch.from_numpy(cot.astype(np.float32)).cuda()).cpu().numpy()
            eg32 = O.slogdet_bwd(A32.astype(np.float64), cot.astype(np.float32).astype(np.float64))
            ok32 = np.isfinite(eg32).reshape(B, -1).all(axis=1) & f32
            sc32 = np.abs(eg32).reshape(B, -1).max(axis=1)[:, None, None] + 1e-30
            assert np.all(np.abs(g32[ok32] - eg32[ok32].astype(np.float32)) <= 2e-7 * sc32[ok32]), lanes
            got = (la.tobytes(), lu.cpu().numpy().tobytes(), perm.cpu().numpy().tobytes(), G.tobytes())
            if first is None:
                first = got
            assert got == first, lanes
    finally:
        for kind in range(4):
            set_tile(kind, n, 0)
    assert [get_tile(kind, n) for kind in range(4)] == default
    with pytest.raises(ValueError):
        set_tile(0, n, 5)


@pytest.mark.parametrize("n", [10, 12, 16, 24])
def test_unaligned_buffers_take_scalar_path(n):
    """Matrices and outputs that start 8 bytes past a 16-byte boundary use the
    kernels' scalar loads and stores (the vector path needs 16-byte rows):
    results bit-identical to the same batch in aligned buffers, for the
    forward, the forward keeping the LU, and both backwards."""
    import torch
    from detnqs_amd import _lib
    from detnqs_amd.core import _ptr, stream_handle
    from detnqs_amd.slogdet import slogdet_bwd_from_a, slogdet_bwd_from_lu, slogdet_fwd, slogdet_fwd_lu
    B = 129
    A = _rand(B, n, seed=900 + n, dominant=True)
    cot = np.random.default_rng(n).standard_normal(B)
    dA = _to(A)
    s0, la0 = slogdet_fwd(dA)
    _, _, lu0, perm0 = slogdet_fwd_lu(dA)
    g0 = slogdet_bwd_from_a(dA, None, _to(cot))
    h0 = slogdet_bwd_from_lu(lu0, perm0, _to(cot))
    nn = n * n
    raw = torch.empty(B * nn + 1, dtype=torch.float64, device="cuda")
    ua = raw[1:].view(B, n, n)
    ua.copy_(dA)
    assert ua.data_ptr() % 16 == 8
    sign = torch.empty(B + 1, dtype=torch.float64, device="cuda")[1:]
    logabs = torch.empty(B + 1, dtype=torch.float64, device="cuda")[1:]
    _lib.call("dq_slogdet_fwd_f64", _ptr(ua), B, n, _ptr(sign), _ptr(logabs), stream_handle())
    assert sign.cpu().numpy().tobytes() == s0.cpu().numpy().tobytes()
    assert logabs.cpu().numpy().tobytes() == la0.cpu().numpy().tobytes()
    lu = torch.empty(B * nn + 1, dtype=torch.float64, device="cuda")[1:].view(B, n, n)
    perm = torch.empty((B, n), dtype=torch.int32, device="cuda")
    _lib.call("dq_slogdet_fwd_lu_f64", _ptr(ua), B, n, _ptr(sign), _ptr(logabs), _ptr(lu), _ptr(perm),
              stream_handle())
    assert lu.cpu().numpy().tobytes() == lu0.cpu().numpy().tobytes()
    assert perm.cpu().numpy().tobytes() == perm0.cpu().numpy().tobytes()
    g = torch.empty(B * nn + 1, dtype=torch.float64, device="cuda")[1:].view(B, n, n)
    c = _to(cot)
    _lib.call("dq_slogdet_bwd_f64", _ptr(ua), _ptr(None), _ptr(c), B, n, _ptr(g), stream_handle())
    assert g.cpu().numpy().tobytes() == g0.cpu().numpy().tobytes()
    _lib.call("dq_slogdet_bwd_lu_f64", _ptr(lu), _ptr(perm), _ptr(c), B, n, _ptr(g), stream_handle())
    assert g.cpu().numpy().tobytes() == h0.cpu().numpy().tobytes()
