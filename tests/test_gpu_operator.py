"""GPU: device Hamiltonian operator and energy steps (the consumers of the
projected blocks, operator/kernel.py:48-145 and operator/functional.py:31-268).

* Products against SciPy's CSR/CSC products on the same blocks
  (hamiltonian.py:186-199 builds them as coo.tocsr() and H_VP^T): bit for bit,
  since each entry is the same sequential column-order sum.
* Energy steps against closed forms on the full matrices: the variational
  step at the FCI ground state of H2O/STO-3G gives E = E_0 and a zero gradient
  (zero-variance); proxy and asymmetric steps give the Rayleigh quotient
  psi^T H_T psi / psi^T psi on the assembled T-space matrix (and its V-only
  variant) and
  cot = 2 psi (H psi - E psi) / den, within 1e-12 relative.
"""

import numpy as np
import pytest

from helpers import H2O_STO3G, fci_dets, same_bits, write_tmp_fcidump

pytestmark = pytest.mark.gpu

@pytest.fixture(scope="module")
def proxy_case(tmp_path_factory):
    import scipy.sparse as sp
    from detnqs_amd import core, synth
    ints, dets = synth.config0()
    path = write_tmp_fcidump(tmp_path_factory.mktemp("op"), ints)
    ctx = core.IntCtx(path, 13)
    ctx.hb_prepare(1e-15)
    V = dets[:1500]
    res = core.get_ham_conn(V, ctx, 13, True, 1e-3)
    nv, npp = V.shape[0], int(res["size_P"])
    h_pp = core.get_ham_diag(res["det_P"], ctx) + 0.25  # with a diag_shift
    vv, vp = res["H_VV"], res["H_VP"]
    H_vv = sp.coo_matrix((vv["val"], (vv["row"], vv["col"])), shape=(nv, nv)).tocsr()
    H_vp = sp.coo_matrix((vp["val"], (vp["row"], vp["col"])), shape=(nv, npp)).tocsr()
    H_pv = H_vp.transpose().conjugate()
    return res, nv, npp, h_pp, H_vv, H_vp, H_pv

def test_products_match_scipy(proxy_case):
    from detnqs_amd.operator import build_proxy_operator, build_vv_operator
    res, nv, npp, h_pp, H_vv, H_vp, H_pv = proxy_case
    assert npp > 100
    rng = np.random.default_rng(3)
    xv, xp = rng.standard_normal(nv), rng.standard_normal(npp)
    op = build_proxy_operator(res["H_VV"], res["H_VP"], h_pp)
    out = op(xv, xp)
    yv_ref = H_vv.dot(xv) + H_vp.dot(xp)
    yp_ref = H_pv.dot(xv) + h_pp * xp
    assert same_bits(out.V.cpu().numpy(), yv_ref)
    assert same_bits(out.P.cpu().numpy(), yp_ref)
    opv = build_vv_operator(res["H_VV"])
    assert same_bits(opv(xv).V.cpu().numpy(), H_vv.dot(xv))

@pytest.mark.parametrize("mode", ["proxy", "asymmetric"])
def test_proxy_and_asymmetric_closed_form(proxy_case, mode):
    from detnqs_amd.operator import DeviceOperator
    import scipy.sparse as sp
    res, nv, npp, h_pp, H_vv, H_vp, H_pv = proxy_case
    HT = sp.bmat([[H_vv, H_vp], [H_pv, sp.diags(h_pp)]]).tocsr()  # H_T on T = V + P
    rng = np.random.default_rng(4)
    n = nv + npp
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    logabs = rng.normal(0.0, 2.0, n)
    op = DeviceOperator(nv, res["H_VV"], res["H_VP"], npp, h_pp)
    e, cot, psi, hpsi = op.energy_step(sign, logabs, mode=mode, eps=1e-12)
    p = sign * np.exp(logabs - logabs.max())
    hp = HT @ p
    if mode == "proxy":
        den = p @ p
        E = (p @ hp) / den
        g = (hp - E * p) / den
        cref = 2 * p * g
    else:
        den = p[:nv] @ p[:nv]
        E = (p[:nv] @ hp[:nv]) / den
        cref = np.concatenate([2 * p[:nv] * (hp[:nv] - E * p[:nv]) / den, np.zeros(npp)])
    assert abs(e - E) <= 1e-12 * abs(E)
    assert np.allclose(psi.cpu().numpy(), p, rtol=1e-14, atol=0)
    c = cot.cpu().numpy()
    assert np.all(np.abs(c - cref) <= 1e-10 * np.abs(cref).max())
    if mode == "asymmetric":
        assert np.all(c[nv:] == 0.0)

def test_variational_zero_variance():
    """psi = FCI ground state of H2O/STO-3G: E = E_0 and cot = 0 (to rounding)."""
    from detnqs_amd import core
    from detnqs_amd.operator import DeviceOperator
    ctx = core.IntCtx(H2O_STO3G, 7)
    F = fci_dets(7, 5, 5)
    vv = core.get_ham_vv(F, ctx, 7)
    n = F.shape[0]
    H = np.zeros((n, n))
    H[vv["row"].astype(np.int64), vv["col"].astype(np.int64)] = vv["val"]
    w, U = np.linalg.eigh(H)
    psi0 = U[:, 0]
    sign = np.sign(psi0)
    sign[sign == 0] = 1.0
    logabs = np.log(np.abs(psi0) + 1e-300)
    op = DeviceOperator(n, vv)
    e, cot, psi, hpsi = op.energy_step(sign, logabs, mode="variational", eps=1e-12)
    assert abs(e - w[0]) <= 1e-10 * abs(w[0])
    assert abs(e + ctx.get_e_nuc() - (-75.01240139)) < 1e-6  # E_FCI, tests/cpp/test_hamiltonian.cpp:111
    assert np.abs(cot.cpu().numpy()).max() < 1e-9

def test_errors(proxy_case):
    from detnqs_amd.operator import DeviceOperator
    res, nv, npp, h_pp, *_ = proxy_case
    bad = dict(res["H_VV"])
    bad["col"] = np.asarray(bad["col"]).copy()
    bad["col"][0] = nv + 5
    with pytest.raises(IndexError):
        DeviceOperator(nv, bad)
    op = DeviceOperator(nv, res["H_VV"], res["H_VP"], npp, h_pp)
    with pytest.raises(ValueError):
        op.energy_step(np.ones(nv), np.zeros(nv), mode="proxy")
    with pytest.raises(ValueError):
        op.energy_step(np.ones(nv), np.zeros(nv), mode="bogus")
