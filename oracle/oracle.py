"""ctypes bindings for the ORACLE (liboracle.so) and the reference build
(oracle/_ref/libdetnqs_ref.so).  TEST INFRASTRUCTURE ONLY: imported by tests/,
by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
product package.

Both libraries expose the same operations; ``Lib("oracle")`` wraps the
restatement in oracle.cpp, ``Lib("ref")`` wraps the reference core compiled
from /root/reference/csrc by oracle/build_ref.sh (absent on hosts without the
reference; ``ref_available()`` says so).

What pins each oracle function (tests/test_oracle.py):

* FCIDUMP parse, packed integrals (integral_mo.*): known H2/6-31G values and
  8-fold symmetry from tests/cpp/test_integral.cpp:65-117; bit-identical
  round trip of the synthetic writer; equality with the reference build.
* analyze_excitation / phase (det_ops.cpp:45-196): the phase KATs of
  tests/cpp/test_determinant.cpp:78-149; brute-force second-quantised sign
  on random tiny dets; equality with the reference build.
* singles/doubles generators (det_ops.hpp:123-268): closed-form counts and
  uniqueness (test_determinant.cpp:155-189); brute-force classification of
  all FCI pairs by popcount degree.
* Slater–Condon (ham_eval.cpp): H2O/STO-3G FCI energy -75.01240139 within
  1e-7 (test_hamiltonian.cpp:111,282-295); hermiticity; dense second-quantised
  Hamiltonian built from first principles on tiny systems; bit-equality with
  the reference build.
* HB table (hb_table.cpp): bit-equality of (a, b, w, row order) with the
  reference build; descending order and brute-force counts.
* get_local_connections / get_ham_vv / block / conn / conn_amp / ham_eff /
  variational energy: bit-equality with the reference build on config 0 and
  H2O/STO-3G; golden digests in tests/golden/ (made from the reference build);
  dense closed forms (FCI H restricted to V; H_VV + H_VP diag(dinv) H_VP^T).
* slogdet fwd/bwd: numpy.linalg.slogdet / inv (LAPACK), closed forms
  (triangular, permutation), padding invariance, central finite differences.
* prepare_det_batch (det_batch.cpp): equality with the reference build on
  random batches and every flag set, the reference's error messages; the
  phase against a brute-force second-quantised sign; occ against bit lists.
* gen_fci (det_enum.cpp, not buildable here: <format> is missing from g++
  11.4): itertools combinations in ascending order, binomial counts.
* gen_connected / gen_perturbative (det_space.cpp, ham_utils.cpp:546-701):
  equality with the reference build (none / static / dynamic); brute force
  over the FCI space by excitation degree for the unscreened sets.
* compute_pt2 (local_conn.cpp:247-418): equality with the reference build run
  with one thread (bit-exact internal term and n_ext, external term to
  rounding); dense FCI-matrix closed forms (sum over P of (H_PV c)^2 /
  (E - H_PP), and the residual r = H_VV c - E c); zero internal term at an
  eigenvector.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdetnqs_ref.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def oracle_available() -> bool:
    return os.path.exists(ORACLE_SO)


class Lib:
    """Thin wrapper over either C-ABI (prefix 'orc_' or 'ref_')."""

    def __init__(self, which: str = "oracle"):
        path = ORACLE_SO if which == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{which} library not built: {path}")
        self.which = which
        self.pfx = "orc_" if which == "oracle" else "ref_"
        self.L = ctypes.CDLL(path)
        f = self._f
        f("last_error").restype = ctypes.c_char_p
        f("ctx_destroy").argtypes = [_vp]
        f("ctx_e_nuc").restype = ctypes.c_double
        f("ctx_e_nuc").argtypes = [_vp]
        f("ctx_nelec").argtypes = [_vp]
        f("get_h1e").restype = ctypes.c_double
        f("get_h1e").argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        f("get_h2e").restype = ctypes.c_double
        f("get_h2e").argtypes = [_vp] + [ctypes.c_int] * 4
        f("hb_prepare").argtypes = [_vp, ctypes.c_double]
        f("hb_export").restype = _vp
        f("hb_export").argtypes = [_vp]
        f("analyze").argtypes = [_u64p, _u64p, _i32p, _f64p]
        f("compute_elem").argtypes = [_vp, _u64p, _u64p, _f64p]
        f("ham_diag").restype = _vp
        f("ham_diag").argtypes = [_vp, _u64p, ctypes.c_size_t]
        f("local_connections").restype = _vp
        f("local_connections").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        f("ham_block").restype = _vp
        f("ham_block").argtypes = [_vp, _u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int]
        f("ham_conn").restype = _vp
        f("ham_conn").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        f("ham_conn_amp").restype = _vp
        f("ham_conn_amp").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double]
        f("ham_eff").restype = _vp
        f("ham_eff").argtypes = [_u32p, _u32p, _f64p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                 _u32p, _u32p, _f64p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                 _f64p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        f("bundle_size").restype = ctypes.c_int64
        f("bundle_size").argtypes = [_vp, ctypes.c_char_p]
        f("bundle_copy").argtypes = [_vp, ctypes.c_char_p, _vp]
        f("bundle_free").argtypes = [_vp]
        prep = [_u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
        f("prepare_det_batch").argtypes = prep
        f("gen_connected").restype = _vp
        f("gen_connected").argtypes = [_u64p, ctypes.c_size_t, ctypes.c_int]
        f("gen_perturbative").restype = _vp
        f("gen_perturbative").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int, _f64p, ctypes.c_int,
                                          ctypes.c_double]
        if which == "oracle":
            f("gen_fci").restype = _vp
            f("gen_fci").argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f("compute_pt2").restype = _vp
            f("compute_pt2").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double]
            f("ctx_from_fcidump").restype = _vp
            f("ctx_from_fcidump").argtypes = [ctypes.c_char_p, ctypes.c_int]
            f("ctx_from_arrays").restype = _vp
            f("ctx_from_arrays").argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_double, ctypes.c_int]
            f("variational_energy").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, _f64p, _f64p]
            f("local_conn_stats").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, _i64p, _f64p]
            f("sigma_rows").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, _i64p, ctypes.c_size_t, _f64p]
            f("slogdet").argtypes = [_f64p, ctypes.c_int64, ctypes.c_int, _f64p, _f64p]
            f("slogdet_bwd").argtypes = [_f64p, _f64p, ctypes.c_int64, ctypes.c_int, _f64p]
        else:
            f("ctx_create").restype = _vp
            f("ctx_create").argtypes = [ctypes.c_char_p, ctypes.c_int]
            f("ham_vv").restype = _vp
            f("ham_vv").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int]
            f("variational_energy").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, _f64p]
            f("local_conn_stream").argtypes = [_vp, _u64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_double, _u64p, _f64p]
            f("compute_pt2").argtypes = [_vp, _u64p, _f64p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p]

    def _f(self, name):
        return getattr(self.L, self.pfx + name)

    def err(self) -> str:
        return self._f("last_error")().decode()

    def _check(self, h):
        if not h:
            raise RuntimeError(f"{self.which}: {self.err()}")
        return h

    # ------------------------------------------------------------ bundles
    def _take(self, b, key, dtype, shape=None):
        n = self._f("bundle_size")(b, key.encode())
        if n < 0:
            raise KeyError(key)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype=dtype)
        self._f("bundle_copy")(b, key.encode(), out.ctypes.data_as(_vp))
        return out.reshape(shape) if shape is not None else out

    def _coo(self, b, p):
        return {
            "row": self._take(b, p + ".row", np.uint32),
            "col": self._take(b, p + ".col", np.uint32),
            "val": self._take(b, p + ".val", np.float64),
            "shape": tuple(int(x) for x in self._take(b, p + ".shape", np.uint32)),
        }

    # ------------------------------------------------------------ API
    def ctx(self, fcidump: Optional[str] = None, n_orb: int = 0, ints=None):
        if self.which == "ref":
            if fcidump is None:
                raise ValueError("reference build takes an FCIDUMP path")
            return Ctx(self, self._check(self._f("ctx_create")(fcidump.encode(), n_orb)))
        if fcidump is not None:
            return Ctx(self, self._check(self._f("ctx_from_fcidump")(fcidump.encode(), n_orb)))
        h1 = np.ascontiguousarray(ints.h1e, dtype=np.float64)
        h2 = np.ascontiguousarray(ints.h2e, dtype=np.float64)
        return Ctx(self, self._check(self._f("ctx_from_arrays")(ints.n_orb, _ptr(h1, _f64p), _ptr(h2, _f64p),
                                                               float(ints.e_nuc), int(ints.n_elec))))

    def analyze(self, bra, ket):
        b = np.ascontiguousarray(bra, dtype=np.uint64)
        k = np.ascontiguousarray(ket, dtype=np.uint64)
        o = np.zeros(3, dtype=np.int32)
        ph = ctypes.c_double()
        self._f("analyze")(_ptr(b, _u64p), _ptr(k, _u64p), _ptr(o, _i32p), ctypes.byref(ph))
        return int(o[0]), int(o[1]), int(o[2]), ph.value

    def ham_eff(self, H_VV, H_VP, h_pp_diag, e_ref, reg_type="sigma", epsilon=1e-12):
        def arr(d, k, t):
            return np.ascontiguousarray(d[k], dtype=t)
        vr, vc, vv = arr(H_VV, "row", np.uint32), arr(H_VV, "col", np.uint32), arr(H_VV, "val", np.float64)
        pr, pc, pv = arr(H_VP, "row", np.uint32), arr(H_VP, "col", np.uint32), arr(H_VP, "val", np.float64)
        hpp = np.ascontiguousarray(h_pp_diag, dtype=np.float64)
        b = self._check(self._f("ham_eff")(
            _ptr(vr, _u32p), _ptr(vc, _u32p), _ptr(vv, _f64p), vv.size, H_VV["shape"][0], H_VV["shape"][1],
            _ptr(pr, _u32p), _ptr(pc, _u32p), _ptr(pv, _f64p), pv.size, H_VP["shape"][0], H_VP["shape"][1],
            _ptr(hpp, _f64p), hpp.size, float(e_ref), 1 if reg_type == "sigma" else 0, float(epsilon)))
        try:
            return self._coo(b, "H")
        finally:
            self._f("bundle_free")(b)

    def prepare_det_batch(self, dets, ref, n_orb, n_alpha, n_beta, kmax=0, need_k=False, need_phase=False,
                          need_hp=False, need_hp_pos=False):
        """prepare_det_batch (det_batch.cpp:189-319) with the bridge's outputs
        (bridge.cpp:347-444).  Raises ValueError with the reference's message."""
        d = np.ascontiguousarray(dets, dtype=np.uint64).reshape(-1, 2)
        r = np.asarray(ref, dtype=np.uint64).reshape(2)
        B = d.shape[0]
        ne = n_alpha + n_beta
        need_k = need_k or need_hp
        K = max(kmax, 0)
        occ = np.zeros((B, max(ne, 0)), dtype=np.int32)
        k = np.zeros(B, dtype=np.int8)
        ph = np.zeros(B, dtype=np.int8)
        holes = np.zeros((B, K), dtype=np.int32)
        parts = np.zeros((B, K), dtype=np.int32)
        mask = np.zeros((B, K), dtype=np.uint8)
        hpos = np.zeros((B, K), dtype=np.int32)
        ppos = np.zeros((B, K), dtype=np.int32)
        flags = (1 if need_k else 0) | (2 if need_phase else 0) | (4 if need_hp else 0) | (8 if need_hp_pos else 0)
        ptr = lambda a: a.ctypes.data_as(_vp)  # noqa: E731
        rc = self._f("prepare_det_batch")(_ptr(d, _u64p), B, int(r[0]), int(r[1]), n_orb, n_alpha, n_beta, kmax,
                                          flags, ptr(occ), ptr(k), ptr(ph), ptr(holes), ptr(parts), ptr(mask),
                                          ptr(hpos), ptr(ppos))
        if rc != 0:
            raise ValueError(self.err())
        out = {"occ": occ}
        if need_k:
            out["k"] = k
        if need_phase:
            out["phase"] = ph
        if need_hp:
            out.update(holes=holes, parts=parts, hp_mask=mask.astype(bool))
            if need_hp_pos:
                out.update(holes_pos=hpos, parts_pos=ppos)
        return out

    def gen_fci(self, n_orb, n_alpha, n_beta):
        assert self.which == "oracle"
        b = self._f("gen_fci")(n_orb, n_alpha, n_beta)
        if not b:
            raise ValueError(self.err())
        try:
            return self._take(b, "dets", np.uint64).reshape(-1, 2)
        finally:
            self._f("bundle_free")(b)

    def gen_connected(self, dets, n_orb):
        d = np.ascontiguousarray(dets, dtype=np.uint64).reshape(-1, 2)
        b = self._check(self._f("gen_connected")(_ptr(d, _u64p), d.shape[0], n_orb))
        try:
            return self._take(b, "dets", np.uint64).reshape(-1, 2)
        finally:
            self._f("bundle_free")(b)

    def slogdet(self, A: np.ndarray):
        assert self.which == "oracle"
        A = np.ascontiguousarray(A, dtype=np.float64)
        n = A.shape[-1]
        B = int(np.prod(A.shape[:-2])) if A.ndim > 2 else 1
        s = np.empty(B)
        la = np.empty(B)
        self._f("slogdet")(_ptr(A, _f64p), B, n, _ptr(s, _f64p), _ptr(la, _f64p))
        return s.reshape(A.shape[:-2]), la.reshape(A.shape[:-2])

    def slogdet_bwd(self, A: np.ndarray, cot_logabs: np.ndarray):
        assert self.which == "oracle"
        A = np.ascontiguousarray(A, dtype=np.float64)
        c = np.ascontiguousarray(cot_logabs, dtype=np.float64).reshape(-1)
        n = A.shape[-1]
        G = np.empty_like(A)
        self._f("slogdet_bwd")(_ptr(A, _f64p), _ptr(c, _f64p), c.size, n, _ptr(G, _f64p))
        return G


class Ctx:
    def __init__(self, lib: Lib, h):
        self.lib = lib
        self.h = h

    def __del__(self):
        try:
            if self.h:
                self.lib._f("ctx_destroy")(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def e_nuc(self) -> float:
        return self.lib._f("ctx_e_nuc")(self.h)

    @property
    def n_elec(self) -> int:
        return self.lib._f("ctx_nelec")(self.h)

    def h1e(self, p, q):
        return self.lib._f("get_h1e")(self.h, p, q)

    def h2e(self, p, q, r, s):
        return self.lib._f("get_h2e")(self.h, p, q, r, s)

    def hb_prepare(self, threshold: float = 1e-15):
        if self.lib._f("hb_prepare")(self.h, float(threshold)) != 0:
            raise RuntimeError(self.lib.err())

    def hb_export(self):
        b = self.lib._check(self.lib._f("hb_export")(self.h))
        try:
            return {
                "row_offsets": self.lib._take(b, "row_offsets", np.uint64),
                "a": self.lib._take(b, "a", np.int32),
                "b": self.lib._take(b, "b", np.int32),
                "w": self.lib._take(b, "w", np.float64),
            }
        finally:
            self.lib._f("bundle_free")(b)

    def compute_elem(self, bra, ket) -> float:
        b = np.ascontiguousarray(bra, dtype=np.uint64)
        k = np.ascontiguousarray(ket, dtype=np.uint64)
        out = ctypes.c_double()
        self.lib._f("compute_elem")(self.h, _ptr(b, _u64p), _ptr(k, _u64p), ctypes.byref(out))
        return out.value

    def ham_diag(self, dets):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        b = self.lib._check(self.lib._f("ham_diag")(self.h, _ptr(d, _u64p), d.shape[0]))
        try:
            return self.lib._take(b, "diag", np.float64)
        finally:
            self.lib._f("bundle_free")(b)

    def local_connections(self, dets, n_orb, use_heatbath=False, eps1=1e-6):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        b = self.lib._check(self.lib._f("local_connections")(self.h, _ptr(d, _u64p), d.shape[0], n_orb,
                                                             int(bool(use_heatbath)), float(eps1)))
        try:
            off_t = np.int64 if self.lib.which == "oracle" else np.int32
            return {
                "offsets": self.lib._take(b, "offsets", off_t).astype(np.int64),
                "dets": self.lib._take(b, "dets", np.uint64).reshape(-1, 2),
                "values": self.lib._take(b, "values", np.float64),
            }
        finally:
            self.lib._f("bundle_free")(b)

    def local_conn_stats(self, dets, n_orb, use_heatbath=False, eps1=1e-6):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        n = d.shape[0]
        counts = np.empty(n, dtype=np.int64)
        vs = np.empty(n)
        if self.lib._f("local_conn_stats")(self.h, _ptr(d, _u64p), n, n_orb, int(bool(use_heatbath)),
                                           float(eps1), _ptr(counts, _i64p), _ptr(vs, _f64p)) != 0:
            raise RuntimeError(self.lib.err())
        return counts, vs

    def local_conn_stream(self, dets, n_orb, use_heatbath=False, eps1=1e-6):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        nnz = ctypes.c_uint64()
        vs = ctypes.c_double()
        if self.lib._f("local_conn_stream")(self.h, _ptr(d, _u64p), d.shape[0], n_orb, int(bool(use_heatbath)),
                                            float(eps1), ctypes.byref(nnz), ctypes.byref(vs)) != 0:
            raise RuntimeError(self.lib.err())
        return int(nnz.value), float(vs.value)

    def _blocks(self, b):
        try:
            out = {"H_VV": self.lib._coo(b, "H_VV"), "H_VP": self.lib._coo(b, "H_VP"),
                   "det_P": self.lib._take(b, "det_P", np.uint64).reshape(-1, 2)}
            out["size_P"] = int(out["det_P"].shape[0])
            return out
        finally:
            self.lib._f("bundle_free")(b)

    def ham_vv(self, dets, n_orb):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        if self.lib.which == "ref":
            b = self.lib._check(self.lib._f("ham_vv")(self.h, _ptr(d, _u64p), d.shape[0], n_orb))
            try:
                return self.lib._coo(b, "H")
            finally:
                self.lib._f("bundle_free")(b)
        return self.ham_block(d, None, n_orb)["H_VV"]

    def ham_block(self, dets_V, dets_P, n_orb):
        v = np.ascontiguousarray(dets_V, dtype=np.uint64)
        if dets_P is None:
            p = np.zeros((0, 2), dtype=np.uint64)
            pp = None
        else:
            p = np.ascontiguousarray(dets_P, dtype=np.uint64)
            pp = _ptr(p, _u64p)
        b = self.lib._check(self.lib._f("ham_block")(self.h, _ptr(v, _u64p), v.shape[0], pp, p.shape[0], n_orb))
        return self._blocks(b)

    def ham_conn(self, dets, n_orb, use_heatbath=False, eps1=1e-6):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        b = self.lib._check(self.lib._f("ham_conn")(self.h, _ptr(d, _u64p), d.shape[0], n_orb,
                                                    int(bool(use_heatbath)), float(eps1)))
        return self._blocks(b)

    def ham_conn_amp(self, dets, psi, n_orb, eps1=1e-6):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        p = np.ascontiguousarray(psi, dtype=np.float64)
        b = self.lib._check(self.lib._f("ham_conn_amp")(self.h, _ptr(d, _u64p), _ptr(p, _f64p), d.shape[0], n_orb,
                                                        float(eps1)))
        return self._blocks(b)

    def gen_perturbative(self, dets, n_orb, psi=None, mode="none", eps1=1e-6):
        """gen_perturbative_dets (bridge.cpp:278-338)."""
        d = np.ascontiguousarray(dets, dtype=np.uint64).reshape(-1, 2)
        m = {"none": 0, "static": 1, "dynamic": 2}[mode]
        p = np.ascontiguousarray(psi if psi is not None else np.zeros(d.shape[0]), dtype=np.float64)
        b = self.lib._f("gen_perturbative")(self.h, _ptr(d, _u64p), d.shape[0], n_orb, _ptr(p, _f64p), m,
                                            float(eps1))
        if not b:
            raise ValueError(self.lib.err())
        try:
            return self.lib._take(b, "dets", np.uint64).reshape(-1, 2)
        finally:
            self.lib._f("bundle_free")(b)

    def compute_pt2(self, dets, coeffs, n_orb, e_ref, use_heatbath=False, eps1=1e-6, threads=1):
        """compute_pt2 (local_conn.cpp:247-418): dict with e_pt2_internal,
        e_pt2_external, n_ext; the oracle also returns ext_dets / ext_vals (v_a)."""
        d = np.ascontiguousarray(dets, dtype=np.uint64).reshape(-1, 2)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        if self.lib.which == "ref":
            out = np.zeros(3)
            rc = self.lib._f("compute_pt2")(self.h, _ptr(d, _u64p), _ptr(c, _f64p), d.shape[0], n_orb,
                                            float(e_ref), int(bool(use_heatbath)), float(eps1), int(threads),
                                            _ptr(out, _f64p))
            if rc != 0:
                raise ValueError(self.lib.err())
            return {"e_pt2_internal": float(out[0]), "e_pt2_external": float(out[1]), "n_ext": int(out[2])}
        b = self.lib._f("compute_pt2")(self.h, _ptr(d, _u64p), _ptr(c, _f64p), d.shape[0], n_orb, float(e_ref),
                                       int(bool(use_heatbath)), float(eps1))
        if not b:
            raise ValueError(self.lib.err())
        try:
            return {"e_pt2_internal": float(self.lib._take(b, "e_int", np.float64)[0]),
                    "e_pt2_external": float(self.lib._take(b, "e_ext", np.float64)[0]),
                    "n_ext": int(self.lib._take(b, "n_ext", np.int64)[0]),
                    "ext_dets": self.lib._take(b, "ext_dets", np.uint64).reshape(-1, 2),
                    "ext_vals": self.lib._take(b, "ext_vals", np.float64)}
        finally:
            self.lib._f("bundle_free")(b)

    def sigma_rows(self, dets, coeffs, rows, n_orb, use_heatbath=False, eps1=1e-6):
        """sigma_i of compute_variational_energy for the listed rows of V only."""
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        r = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.empty(r.size)
        rc = self.lib._f("sigma_rows")(self.h, _ptr(d, _u64p), _ptr(c, _f64p), d.shape[0], n_orb,
                                       int(bool(use_heatbath)), float(eps1), _ptr(r, _i64p), r.size,
                                       _ptr(out, _f64p))
        if rc != 0:
            raise RuntimeError(self.lib.err())
        return out

    def variational_energy(self, dets, coeffs, n_orb, use_heatbath=False, eps1=1e-6, want_sigma=False):
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        out = ctypes.c_double()
        if self.lib.which == "oracle":
            sig = np.empty(d.shape[0]) if want_sigma else None
            rc = self.lib._f("variational_energy")(self.h, _ptr(d, _u64p), _ptr(c, _f64p), d.shape[0], n_orb,
                                                   int(bool(use_heatbath)), float(eps1), ctypes.byref(out),
                                                   _ptr(sig, _f64p) if sig is not None else None)
            if rc != 0:
                raise RuntimeError(self.lib.err())
            return (out.value, sig) if want_sigma else out.value
        rc = self.lib._f("variational_energy")(self.h, _ptr(d, _u64p), _ptr(c, _f64p), d.shape[0], n_orb,
                                               int(bool(use_heatbath)), float(eps1), ctypes.byref(out))
        if rc != 0:
            raise RuntimeError(self.lib.err())
        return out.value
