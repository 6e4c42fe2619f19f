"""Python mirror of the lapw1 hot-path interface (HAMILT, HNS, SECLR4) over the
C-ABI of libw2klapw1.so (include/w2k_lapw1.h).

The reference reaches these routines from CALKPT (calkpt.F:143-197) with
module state; here that state is one `RunData` (per run) plus a basis per
k-point, and the three routines are methods of `Lapw1` with the same names
and argument meaning.  Errors follow the reference's fatal paths
(seclr4.F:200-205 Cholesky INFO, 230-235 GST INFO, 304-320 SYEVX INFO / NE >
NUME) as exceptions carrying the same numbers.

PyTorch is not needed here: arrays are numpy, the device memory lives in the
library.  The library must be built (hack-wein2k_amd/build.py); importing this
module on a machine without it raises at the first use.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libw2klapw1.so")

LMAX = 13
LOMAX = 3
NLOAT = 3
NSLMAX = 5

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class RunDesc(ctypes.Structure):
    """ctypes image of w2k_run_desc (include/w2k_lapw1.h)."""
    _fields_ = [
        ("is_complex", ctypes.c_int), ("nat", ctypes.c_int), ("ndf", ctypes.c_int),
        ("ltop", ctypes.c_int), ("lmaxd", ctypes.c_int), ("nnsl", ctypes.c_int),
        ("lmmxd", ctypes.c_int), ("ngaud", ctypes.c_int), ("ins", ctypes.c_int),
        ("mult", _c_int_p), ("pos", _c_dbl_p), ("rmt", _c_dbl_p), ("v", _c_dbl_p),
        ("vi", ctypes.c_double),
        ("u", _c_dbl_p), ("up", _c_dbl_p), ("du", _c_dbl_p), ("dup", _c_dbl_p),
        ("chn", _c_dbl_p), ("e", _c_dbl_p),
        ("lapw", _c_int_p),
        ("nlo", ctypes.c_int), ("ilo", _c_int_p), ("loor", _c_int_p),
        ("alo", _c_dbl_p), ("blo", _c_dbl_p), ("clo", _c_dbl_p), ("elo", _c_dbl_p),
        ("pi12lo", _c_dbl_p), ("pe12lo", _c_dbl_p), ("dplo", _c_dbl_p),
        ("rotloc", _c_dbl_p), ("rotij", _c_dbl_p),
        ("warp1", ctypes.c_int), ("warp2", ctypes.c_int), ("warp3", ctypes.c_int),
        ("warp", _c_dbl_p),
        ("lqind", _c_int_p), ("lqns", _c_int_p), ("gfac", _c_dbl_p),
        ("vns1", _c_dbl_p), ("vns2", _c_dbl_p), ("vns3", _c_dbl_p),
        ("vns1lo", _c_dbl_p), ("vns2lo", _c_dbl_p), ("vns3lo", _c_dbl_p),
    ]


def _f(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class RunData:
    """Per-run state of HAMILT/HNS in the reference's Fortran layouts.

    Keys (as dumped from the reference modules): NAT, NDF, LTOP, LNSMAX, INS,
    NLO, MULT, POS, RMT, V, VI, P, PE, DP, DPE, PEI, E, LAPW, ILO, LOOR, ALO,
    BLO, CLO, ELO, PI12LO, PE12LO, DPLO, ROTLOC, ROTIJ, WARP1..3, WARP, LQIND,
    LQNS, GFAC, VNS1..3, VNS1LO..3LO; `is_complex` selects lapw1c.
    """

    def __init__(self, d: Dict[str, np.ndarray], is_complex: bool):
        self.is_complex = bool(is_complex)
        g = lambda k: d[k]
        self.nat = int(g("NAT"))
        self.ndf = int(g("NDF"))
        self.ltop = int(g("LTOP"))
        self.nnsl = int(g("LNSMAX")) + 1
        self.ins = int(g("INS"))
        self.nlo = int(g("NLO"))
        self.vi = float(g("VI"))
        self.arr = {}
        A = self.arr
        A["mult"] = _f(np.atleast_1d(g("MULT")), np.int32)
        A["pos"] = _f(g("POS"), np.float64)
        A["rmt"] = _f(np.atleast_1d(g("RMT")), np.float64)
        A["v"] = _f(np.atleast_1d(g("V")), np.float64)
        for k, src in (("u", "P"), ("up", "PE"), ("du", "DP"), ("dup", "DPE"), ("chn", "PEI"), ("e", "E")):
            A[k] = _f(np.asarray(g(src)).reshape(LMAX, self.nat, order="F") if np.ndim(g(src)) == 1 else g(src), np.float64)
        A["lapw"] = _f(np.asarray(g("LAPW")).reshape(LMAX, self.nat, order="F"), np.int32)
        A["ilo"] = _f(np.asarray(g("ILO")).reshape(LOMAX + 1, self.nat, order="F"), np.int32)
        A["loor"] = _f(np.asarray(g("LOOR")).reshape(LOMAX + 1, self.nat, order="F"), np.int32)
        for k in ("ALO", "BLO", "CLO", "ELO"):
            A[k.lower()] = _f(np.asarray(g(k)).reshape(LOMAX + 1, NLOAT, self.nat, order="F"), np.float64)
        for k in ("PI12LO", "PE12LO", "DPLO"):
            A[k.lower()] = _f(np.asarray(g(k)).reshape(LOMAX + 1, self.nat, order="F"), np.float64)
        A["rotloc"] = _f(np.asarray(g("ROTLOC")).reshape(3, 3, self.nat, order="F"), np.float64)
        A["rotij"] = _f(np.asarray(g("ROTIJ")).reshape(3, 3, self.ndf, order="F"), np.float64)
        if self.ins == 0:
            self.warp = (int(g("WARP1")), int(g("WARP2")), int(g("WARP3")))
            A["warp"] = _f(g("WARP"), np.complex128)
        else:
            self.warp = (0, 0, 0)
            A["warp"] = np.zeros((1, 1, 1), np.complex128, order="F")
        A["lqind"] = _f(np.atleast_1d(g("LQIND")), np.int32)
        lqns = np.asarray(g("LQNS"))
        if lqns.ndim == 2:
            lqns = lqns.reshape(6, -1, 1, order="F")
        A["lqns"] = _f(lqns, np.int32)
        gfac = np.asarray(g("GFAC"))
        if gfac.ndim == 1:
            gfac = gfac.reshape(-1, 1, order="F")
        A["gfac"] = _f(gfac, np.complex128)
        self.ngaud = int(A["lqns"].shape[1])
        for k in ("VNS1", "VNS2", "VNS3", "VNS1LO", "VNS2LO", "VNS3LO"):
            v = np.asarray(g(k))
            if v.ndim == 3:
                v = v.reshape(v.shape + (1,), order="F")
            A[k.lower()] = _f(v, np.float64)
        self.lmmxd = int(A["vns1"].shape[1])
        self._desc = None

    def desc(self) -> RunDesc:
        A = self.arr
        ip = lambda k: A[k].ctypes.data_as(_c_int_p)
        dp = lambda k: A[k].ctypes.data_as(_c_dbl_p)
        r = RunDesc()
        r.is_complex = int(self.is_complex)
        r.nat, r.ndf, r.ltop, r.lmaxd = self.nat, self.ndf, self.ltop, LMAX
        r.nnsl, r.lmmxd, r.ngaud, r.ins = self.nnsl, self.lmmxd, self.ngaud, self.ins
        r.mult, r.pos, r.rmt, r.v, r.vi = ip("mult"), dp("pos"), dp("rmt"), dp("v"), self.vi
        r.u, r.up, r.du, r.dup, r.chn, r.e = (dp(k) for k in ("u", "up", "du", "dup", "chn", "e"))
        r.lapw = ip("lapw")
        r.nlo, r.ilo, r.loor = self.nlo, ip("ilo"), ip("loor")
        r.alo, r.blo, r.clo, r.elo = (dp(k) for k in ("alo", "blo", "clo", "elo"))
        r.pi12lo, r.pe12lo, r.dplo = dp("pi12lo"), dp("pe12lo"), dp("dplo")
        r.rotloc, r.rotij = dp("rotloc"), dp("rotij")
        r.warp1, r.warp2, r.warp3 = self.warp
        r.warp = dp("warp")
        r.lqind, r.lqns, r.gfac = ip("lqind"), ip("lqns"), dp("gfac")
        r.vns1, r.vns2, r.vns3 = dp("vns1"), dp("vns2"), dp("vns3")
        r.vns1lo, r.vns2lo, r.vns3lo = dp("vns1lo"), dp("vns2lo"), dp("vns3lo")
        self._desc = r  # keep alive with the arrays
        return r


class Basis:
    """Per-k basis as COORS/LOPW leave it (coors.f, lopw.f): NV plane waves
    then NLO local-orbital vectors."""

    def __init__(self, nv: int, nlo: int, kzz, xk, yk, zk, rk):
        self.nv, self.nlo = int(nv), int(nlo)
        n = self.nv + self.nlo
        self.kzz = np.asfortranarray(np.asarray(kzz, np.int32).reshape(3, -1, order="F")[:, :n])
        self.xk, self.yk, self.zk, self.rk = (np.ascontiguousarray(np.asarray(a, np.float64)[:n]) for a in (xk, yk, zk, rk))

    @property
    def n(self) -> int:
        return self.nv + self.nlo


class _BasisDesc(ctypes.Structure):
    """w2k_basis of include/w2k_lapw1.h."""
    _fields_ = [("nv", ctypes.c_int), ("nlo", ctypes.c_int), ("kzz", ctypes.POINTER(ctypes.c_int)),
                ("xk", ctypes.POINTER(ctypes.c_double)), ("yk", ctypes.POINTER(ctypes.c_double)),
                ("zk", ctypes.POINTER(ctypes.c_double)), ("rk", ctypes.POINTER(ctypes.c_double))]


class _StarsDesc(ctypes.Structure):
    _fields_ = [("nkk", ctypes.c_int), ("ja", _c_int_p), ("jb", _c_int_p), ("jc", _c_int_p),
                ("potk", _c_dbl_p), ("iord", ctypes.c_int), ("imat", _c_int_p), ("tau", _c_dbl_p)]


class _RadialDesc(ctypes.Structure):
    _fields_ = [("rel", ctypes.c_int), ("jrid", ctypes.c_int), ("jrj", _c_int_p), ("ro", _c_dbl_p), ("dh", _c_dbl_p),
                ("zz", _c_dbl_p), ("vr", _c_dbl_p), ("lmmxd", ctypes.c_int), ("vlm", _c_dbl_p)]


class Lattice(ctypes.Structure):
    """w2k_lattice of include/w2k_lapw1.h: what COORS reads besides k
    (coors.f:1-18)."""
    _fields_ = [("br2", ctypes.c_double * 9), ("pia", ctypes.c_double * 3), ("ortho", ctypes.c_int),
                ("cxz", ctypes.c_int), ("rkm", ctypes.c_double), ("nmatmax", ctypes.c_int)]

    @classmethod
    def make(cls, br2, pia, ortho: bool, cxz: bool, rkm: float, nmatmax: int = 13000) -> "Lattice":
        """BR2(3,3) as an array (Fortran order is taken from its [i, j] indices)."""
        b = np.asarray(br2, np.float64).reshape(3, 3)
        lat = cls()
        lat.br2[:] = [float(b[i, j]) for j in range(3) for i in range(3)]
        lat.pia[:] = [float(x) for x in np.asarray(pia, np.float64).ravel()[:3]]
        lat.ortho, lat.cxz = int(bool(ortho)), int(bool(cxz))
        lat.rkm, lat.nmatmax = float(rkm), int(nmatmax)
        return lat


class Lapw1Error(RuntimeError):
    def __init__(self, where: str, code: int, detail: str = ""):
        msg = f"{where} failed with status {code}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)
        self.code = code
        self.detail = detail


_LIB = None


def load_library(path: Optional[str] = None):
    """Load libw2klapw1.so (raises OSError when it is missing: there is no CPU fallback)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    lib = ctypes.CDLL(path or LIB_PATH)
    V = ctypes.c_void_p
    lib.w2k_create.argtypes = [ctypes.POINTER(V), ctypes.c_int, ctypes.POINTER(RunDesc)]
    lib.w2k_set_run.argtypes = [V, ctypes.POINTER(RunDesc)]
    lib.w2k_destroy.argtypes = [V]
    lib.w2k_destroy.restype = None
    lib.w2k_set_basis.argtypes = [V, ctypes.c_int, ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    lib.w2k_coors.argtypes = [V, ctypes.POINTER(Lattice), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              _c_int_p, _c_int_p]
    lib.w2k_get_basis.argtypes = [V, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    for fn in ("w2k_hamilt", "w2k_swap_diag", "w2k_hns", "w2k_sync"):
        getattr(lib, fn).argtypes = [V]
    lib.w2k_seclr4.argtypes = [V, ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_int_p, _c_int_p]
    lib.w2k_calkpt.argtypes = [V, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _c_int_p, _c_int_p]
    lib.w2k_get_hs.argtypes = [V, _c_dbl_p, ctypes.c_int, _c_dbl_p]
    lib.w2k_put_hs.argtypes = [V, _c_dbl_p, ctypes.c_int, _c_dbl_p]
    lib.w2k_get_diag.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.w2k_put_diag.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.w2k_get_albl.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.w2k_get_eig.argtypes = [V, _c_dbl_p, _c_dbl_p, ctypes.c_int]
    lib.w2k_device_buffers.argtypes = [V, ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V), ctypes.POINTER(V)]
    lib.w2k_stream.argtypes = [V]
    lib.w2k_stream.restype = V
    lib.w2k_last_timings.argtypes = [V, _c_dbl_p, ctypes.c_int]
    lib.w2k_version.restype = ctypes.c_char_p
    lib.w2k_last_error.restype = ctypes.c_char_p
    lib.w2k_set_kernel_timing.argtypes = [V, ctypes.c_int]
    lib.w2k_set_split_precision.argtypes = [V, ctypes.c_int]
    lib.w2k_set_graphs.argtypes = [V, ctypes.c_int]
    lib.w2k_adjoint_warp.argtypes = [V, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    lib.w2k_adjoint_vns.argtypes = [V, _c_dbl_p] + [_c_dbl_p] * 6
    lib.w2k_set_stars.argtypes = [V, ctypes.POINTER(_StarsDesc)]
    lib.w2k_get_warp.argtypes = [V, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p]
    lib.w2k_adjoint_stars.argtypes = [V, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    lib.w2k_atpar.argtypes = [V, ctypes.POINTER(_RadialDesc)] + [_c_dbl_p] * 16
    lib.w2k_write_vector.argtypes = [V, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p]
    lib.w2k_density_interstitial.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.w2k_adjoint_vlm.argtypes = [V, _c_dbl_p, _c_dbl_p]
    lib.w2k_calkpt_batched.argtypes = [V, ctypes.c_int, ctypes.POINTER(_BasisDesc), ctypes.c_int, _c_dbl_p, _c_dbl_p,
                                       ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p,
                                       ctypes.c_int, ctypes.POINTER(_c_dbl_p), _c_int_p]
    if path is None:
        _LIB = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_c_dbl_p)


class Lapw1:
    """One device context: HAMILT / HNS / SECLR4 on one GPU."""

    def __init__(self, run: RunData, device: int = 0):
        self.lib = load_library()
        self.run = run
        self.ctx = ctypes.c_void_p()
        rc = self.lib.w2k_create(ctypes.byref(self.ctx), int(device), ctypes.byref(run.desc()))
        if rc != 0:
            raise Lapw1Error("w2k_create", rc)
        self.basis: Optional[Basis] = None
        self.ne = 0
        self._nkk = 0

    def set_run(self, run: RunData):
        """Replace the per-run data (w2k_set_run)."""
        self.run = run
        self._chk("w2k_set_run", self.lib.w2k_set_run(self.ctx, ctypes.byref(run.desc())))

    def close(self):
        if self.ctx:
            self.lib.w2k_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, where, rc):
        if rc != 0:
            detail = self.lib.w2k_last_error().decode() if rc < 0 else ""
            raise Lapw1Error(where, rc, detail)

    @property
    def cdt(self):
        return np.complex128 if self.run.is_complex else np.float64

    def set_basis(self, b: Basis):
        self.basis = b
        self._chk("w2k_set_basis", self.lib.w2k_set_basis(
            self.ctx, b.nv, b.nlo, b.kzz.ctypes.data_as(_c_int_p), _dp(b.xk), _dp(b.yk), _dp(b.zk), _dp(b.rk)))

    def coors(self, lat: Lattice, sx: float, sy: float, sz: float) -> Basis:
        """COORS + LOPW (coors.f, lopw.f) on the device: the k-point's basis
        becomes the context's (as set_basis would) and is returned."""
        nv, nlo = ctypes.c_int(0), ctypes.c_int(0)
        self._chk("COORS", self.lib.w2k_coors(self.ctx, ctypes.byref(lat), float(sx), float(sy), float(sz),
                                              ctypes.byref(nv), ctypes.byref(nlo)))
        self.basis = self.get_basis(nv.value, nlo.value)
        return self.basis

    def get_basis(self, nv: Optional[int] = None, nlo: Optional[int] = None) -> Basis:
        """The context's basis (w2k_get_basis)."""
        if nv is None:
            nv, nlo = self.basis.nv, self.basis.nlo
        n = nv + nlo
        kzz = np.zeros((3, n), np.int32, order="F")
        xk, yk, zk, rk = (np.zeros(n) for _ in range(4))
        self._chk("w2k_get_basis", self.lib.w2k_get_basis(self.ctx, kzz.ctypes.data_as(_c_int_p), _dp(xk), _dp(yk),
                                                          _dp(zk), _dp(rk)))
        return Basis(nv, nlo, kzz, xk, yk, zk, rk)

    def hamilt(self):
        """HAMILT (hamilt.F): H/S spherical + interstitial + LO into HS/HSDIAG."""
        self._chk("HAMILT", self.lib.w2k_hamilt(self.ctx))

    def swap_diag(self):
        """DSWAP(HSROWS,HS,HSROWS+1,HSDIAG,1) of calkpt.F:160-166."""
        self._chk("DSWAP", self.lib.w2k_swap_diag(self.ctx))

    def hns(self):
        """HNS (hns.F): non-spherical terms into the lower triangle of HS."""
        self._chk("HNS", self.lib.w2k_hns(self.ctx))

    def seclr4(self, el: float, eu: float, nume: int) -> Tuple[int, int]:
        """SECLR4 (seclr4.F): returns (NE, NBELW)."""
        ne, nb = ctypes.c_int(0), ctypes.c_int(0)
        self._chk("SECLR4", self.lib.w2k_seclr4(self.ctx, float(el), float(eu), int(nume), ctypes.byref(ne), ctypes.byref(nb)))
        self.ne = ne.value
        return ne.value, nb.value

    def calkpt(self, el: float, eu: float, nume: int, nohns: bool = False) -> Tuple[int, int]:
        ne, nb = ctypes.c_int(0), ctypes.c_int(0)
        self._chk("CALKPT", self.lib.w2k_calkpt(self.ctx, int(nohns), float(el), float(eu), int(nume), ctypes.byref(ne), ctypes.byref(nb)))
        self.ne = ne.value
        return ne.value, nb.value

    def calkpt_batched(self, bases, el, eu, nume: int, nworkers: int = 0, nohns: bool = False,
                       vectors: bool = True):
        """Batched k-points (w2k_calkpt_batched): CALKPT for every basis in
        `bases` (ragged N) on concurrent device streams.  el/eu: one window
        per k-point (or scalars).  Returns one (ne, nbelw, eigval, z or None)
        per k-point; raises Lapw1Error if any k-point failed."""
        nk = len(bases)
        el = np.broadcast_to(np.asarray(el, np.float64), (nk,)).copy()
        eu = np.broadcast_to(np.asarray(eu, np.float64), (nk,)).copy()
        descs = (_BasisDesc * nk)()
        for q, b in enumerate(bases):
            descs[q] = _BasisDesc(b.nv, b.nlo, b.kzz.ctypes.data_as(_c_int_p), _dp(b.xk), _dp(b.yk), _dp(b.zk),
                                  _dp(b.rk))
        ne = np.zeros(nk, np.int32)
        nb = np.zeros(nk, np.int32)
        st = np.zeros(nk, np.int32)
        ev = np.zeros((nk, int(nume)), np.float64)
        zs = [np.zeros((b.n, int(nume)), self.cdt, order="F") for b in bases] if vectors else None
        zp = (_c_dbl_p * nk)(*[_dp(a) for a in zs]) if vectors else None
        ldz = np.array([b.n for b in bases], np.int32)
        rc = self.lib.w2k_calkpt_batched(self.ctx, nk, descs, int(nohns), _dp(el), _dp(eu), int(nume), int(nworkers),
                                         ne.ctypes.data_as(_c_int_p), nb.ctypes.data_as(_c_int_p),
                                         st.ctypes.data_as(_c_int_p), _dp(ev), int(nume), zp,
                                         ldz.ctypes.data_as(_c_int_p))
        self._chk("w2k_calkpt_batched", rc)
        return [(int(ne[q]), int(nb[q]), ev[q, :ne[q]].copy(), zs[q][:, :ne[q]] if vectors else None)
                for q in range(nk)]

    def adjoint_warp(self, f) -> Tuple[float, np.ndarray]:
        """Band energy sum_n f_n eps_n (weights held fixed) of the last solved
        k-point and its gradient with respect to the warped interstitial box:
        complex array shaped like the run's WARP, dE/dRe + i dE/dIm."""
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        if f.shape != (self.ne,):
            raise ValueError(f"need {self.ne} weights")
        e = ctypes.c_double(0.0)
        w1, w2, w3 = self.warp_dims()
        g = np.zeros((2 * w1 + 1, 2 * w2 + 1, 2 * w3 + 1), np.complex128, order="F")
        self._chk("w2k_adjoint_warp", self.lib.w2k_adjoint_warp(self.ctx, _dp(f), ctypes.byref(e), _dp(g)))
        return e.value, g

    def set_stars(self, ja, jb, jc, potk, imat, tau):
        """STERN + SETWAR on the device: the run's warped box from the
        interstitial stars (JA/JB/JC, POTK) and the space group (IMAT(3,3,IORD),
        TAU(3,IORD)), replacing the box of the run data."""
        ja, jb, jc = (np.ascontiguousarray(np.asarray(x, np.int32).ravel()) for x in (ja, jb, jc))
        potk = np.ascontiguousarray(np.asarray(potk, np.complex128).ravel())
        imat = np.asfortranarray(np.asarray(imat, np.int32))
        tau = np.asfortranarray(np.asarray(tau, np.float64))
        d = _StarsDesc(len(ja), ja.ctypes.data_as(_c_int_p), jb.ctypes.data_as(_c_int_p), jc.ctypes.data_as(_c_int_p),
                       _dp(potk), int(imat.shape[2]), imat.ctypes.data_as(_c_int_p), _dp(tau))
        self._chk("w2k_set_stars", self.lib.w2k_set_stars(self.ctx, ctypes.byref(d)))
        self._nkk = len(ja)

    def warp_dims(self) -> Tuple[int, int, int]:
        w = [ctypes.c_int(0) for _ in range(3)]
        self._chk("w2k_get_warp", self.lib.w2k_get_warp(self.ctx, *[ctypes.byref(x) for x in w],
                                                        ctypes.cast(None, _c_dbl_p)))
        return tuple(x.value for x in w)

    def get_warp(self) -> np.ndarray:
        """The warped box HAMILT reads: complex (2 w1+1, 2 w2+1, 2 w3+1)."""
        w1, w2, w3 = self.warp_dims()
        g = np.zeros((2 * w1 + 1, 2 * w2 + 1, 2 * w3 + 1), np.complex128, order="F")
        w = [ctypes.c_int(0) for _ in range(3)]
        self._chk("w2k_get_warp", self.lib.w2k_get_warp(self.ctx, *[ctypes.byref(x) for x in w], _dp(g)))
        return g

    def adjoint_stars(self, f) -> Tuple[float, np.ndarray]:
        """Band energy of the last solved k-point and its gradient with respect
        to the star coefficients POTK (complex: dE/dRe + i dE/dIm)."""
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        if f.shape != (self.ne,):
            raise ValueError(f"need {self.ne} weights")
        e = ctypes.c_double(0.0)
        g = np.zeros(self._nkk, np.complex128)
        self._chk("w2k_adjoint_stars", self.lib.w2k_adjoint_stars(self.ctx, _dp(f), ctypes.byref(e), _dp(g)))
        return e.value, g

    def adjoint_vns(self, f) -> dict:
        """Gradient of the band energy sum_n f_n eps_n (weights held fixed) of
        the last solved k-point with respect to the VNS radial integrals:
        arrays shaped like the run's VNS1..3 and VNS1LO..3LO."""
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        if f.shape != (self.ne,):
            raise ValueError(f"need {self.ne} weights")
        out = {k: np.zeros(self.run.arr[k].shape, np.float64, order="F")
               for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo")}
        self._chk("w2k_adjoint_vns", self.lib.w2k_adjoint_vns(
            self.ctx, _dp(f), *[_dp(out[k]) for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo")]))
        return out

    ATPAR_NAMES = ("p", "dp", "pe", "dpe", "pei", "e", "plo", "dplo", "pi12lo", "pe12lo",
                   "vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo")

    def atpar(self, rel: bool, jrj, ro, dh, zz, vr, vlm=None) -> dict:
        """ATPAR's radial part on the device from the radial mesh (JRJ, RO, DH),
        the nuclear charges, r V00(r) in Hartree (jrid, NAT) and V_LM(r) in Ry
        (jrid, lmmx, NAT).  Returns the radial parameters and the VNS
        integrals (entries the coupling list reads) in the reference layouts."""
        jrj = np.ascontiguousarray(np.asarray(jrj, np.int32).ravel())
        ro, dh, zz = (np.ascontiguousarray(np.asarray(x, np.float64).ravel()) for x in (ro, dh, zz))
        vr = np.asfortranarray(np.asarray(vr, np.float64))
        nat = self.run.nat
        lmaxp, lomaxp = 13, 4
        shapes = {k: (lmaxp, nat) for k in ("p", "dp", "pe", "dpe", "pei", "e")}
        shapes.update({k: (lomaxp, nat) for k in ("plo", "dplo", "pi12lo", "pe12lo")})
        for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo"):
            shapes[k] = self.run.arr[k].shape
        out = {k: np.zeros(sh, np.float64, order="F") for k, sh in shapes.items()}
        if vlm is not None:
            vlm = np.asfortranarray(np.asarray(vlm, np.float64))
        d = _RadialDesc(int(bool(rel)), int(vr.shape[0]), jrj.ctypes.data_as(_c_int_p), _dp(ro), _dp(dh), _dp(zz),
                        _dp(vr), int(vlm.shape[1]) if vlm is not None else 0,
                        _dp(vlm) if vlm is not None else ctypes.cast(None, _c_dbl_p))
        self._chk("w2k_atpar", self.lib.w2k_atpar(self.ctx, ctypes.byref(d),
                                                  *[_dp(out[k]) for k in self.ATPAR_NAMES]))
        self._vlm_shape = (int(vr.shape[0]), int(vlm.shape[1]) if vlm is not None else 0, nat)
        return out

    def adjoint_vlm(self, f) -> np.ndarray:
        """Gradient of the band energy of the last solved k-point (with HNS)
        with respect to V_LM(r_k): (jrid, lmmx, NAT) of the last atpar call."""
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        if f.shape != (self.ne,):
            raise ValueError(f"need {self.ne} weights")
        g = np.zeros(self._vlm_shape, np.float64, order="F")
        self._chk("w2k_adjoint_vlm", self.lib.w2k_adjoint_vlm(self.ctx, _dp(f), _dp(g)))
        return g

    def write_vector(self, path: str, kname: str, sx: float, sy: float, sz: float, weight: float, ipgr: str,
                     append: bool = True):
        """Append (or write) the vector-file record of the last solved k-point
        in TAPEWF's layout."""
        self._chk("w2k_write_vector", self.lib.w2k_write_vector(
            self.ctx, path.encode(), int(bool(append)), kname.encode()[:10], float(sx), float(sy), float(sz),
            float(weight), ipgr.encode()[:3]))

    def density_interstitial(self, f) -> np.ndarray:
        """rho(Q) = sum_n f_n sum_{KZZ_J - KZZ_I = Q} conj z_I z_J on the run's box."""
        f = np.ascontiguousarray(np.asarray(f, np.float64))
        if f.shape != (self.ne,):
            raise ValueError(f"need {self.ne} weights")
        w1, w2, w3 = self.warp_dims()
        rho = np.zeros((2 * w1 + 1, 2 * w2 + 1, 2 * w3 + 1), np.complex128, order="F")
        self._chk("w2k_density_interstitial", self.lib.w2k_density_interstitial(self.ctx, _dp(f), _dp(rho)))
        return rho

    def get_hs(self) -> Tuple[np.ndarray, np.ndarray]:
        n = self.basis.n
        hs = np.zeros((n, n), self.cdt, order="F")
        hd = np.zeros(n, self.cdt)
        self._chk("w2k_get_hs", self.lib.w2k_get_hs(self.ctx, _dp(hs), n, _dp(hd)))
        return hs, hd

    def put_hs(self, hs: np.ndarray, hsdiag: np.ndarray):
        hs = np.asfortranarray(hs, dtype=self.cdt)
        hd = np.ascontiguousarray(hsdiag, dtype=self.cdt)
        self._chk("w2k_put_hs", self.lib.w2k_put_hs(self.ctx, _dp(hs), hs.shape[0], _dp(hd)))

    def get_albl(self) -> Tuple[np.ndarray, np.ndarray]:
        n, lt, nat = self.basis.n, self.run.ltop, self.run.nat
        al = np.zeros((n, lt, nat), np.float64, order="F")
        bl = np.zeros((n, lt, nat), np.float64, order="F")
        self._chk("w2k_get_albl", self.lib.w2k_get_albl(self.ctx, _dp(al), _dp(bl)))
        return al, bl

    def get_eig(self) -> Tuple[np.ndarray, np.ndarray]:
        n, ne = self.basis.n, self.ne
        ev = np.zeros(max(ne, 1), np.float64)
        z = np.zeros((n, max(ne, 1)), self.cdt, order="F")
        self._chk("w2k_get_eig", self.lib.w2k_get_eig(self.ctx, _dp(ev), _dp(z), n))
        return ev[:ne], z[:, :ne]

    TIMING_NAMES = ("hamilt", "hns", "seclr4", "cholesky", "reduction", "tridiag", "bisect_invit",
                    "backtransform", "hamilt_pw_kernel", "hns_gemm", "tri_symv_total", "tri_symv_launches",
                    "tri_launch_mode")

    def set_kernel_timing(self, on: bool):
        """Per-launch events around the tridiagonalisation SYMV kernels."""
        self._chk("w2k_set_kernel_timing", self.lib.w2k_set_kernel_timing(self.ctx, int(bool(on))))

    def set_graphs(self, on: bool):
        """Tridiagonalisation launches captured into a HIP graph and replayed
        where that measures faster (on), or always launched directly (off, the
        default)."""
        self._chk("w2k_set_graphs", self.lib.w2k_set_graphs(self.ctx, int(bool(on))))

    def set_split_precision(self, slices: int):
        """HNS rank-K update in split precision: 0 = FP64 (default), 3..7 = INT8
        digit slices per element with FP64 recombination."""
        self._chk("w2k_set_split_precision", self.lib.w2k_set_split_precision(self.ctx, int(slices)))

    def timings(self) -> np.ndarray:
        """Device milliseconds of the last calls (order of TIMING_NAMES)."""
        t = np.zeros(len(self.TIMING_NAMES), np.float64)
        self.lib.w2k_last_timings(self.ctx, _dp(t), len(t))
        return t

    def sync(self):
        self._chk("w2k_sync", self.lib.w2k_sync(self.ctx))
