"""ATPAR's radial part on the device (w2k_atpar: OUTWIN outwin.f, the
normalisation and energy derivative of atpar.f:212-330, the non-spherical
integrals of atpar.f:470-561 with RINT13 rint13.f) and the band-energy
gradient with respect to the non-spherical potential V_LM(r)
(w2k_adjoint_vlm, an extended form with no counterpart in the reference).

Pinned by:
- the radial parameters P, DP, PE, DPE, PEI, PLO, DPLO, PI12LO, PE12LO and
  every VNS1..3 / VNS1LO..3LO integral that the HNS coupling list reads,
  against the reference's own values (its ATPAR on the same deck);
- the k-point solved with the device ATPAR's parameters against the
  reference eigenvalues;
- central finite differences of E_band (weights fixed, Fermi-Dirac at 1 mRy)
  through the library, with V_LM(r_k) at single mesh points moved by +-h and
  the integrals recomputed by w2k_atpar, to 1e-6 relative."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from conftest import Fixture, PKG, REPO

pytestmark = pytest.mark.gpu
L = PKG.lapw1
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import w2k_inputs as W  # noqa: E402

CASES = ["c1_fcc", "lo_mult2", "lo_cplx", "rs8", "ins_lo"]


def _radial(fx):
    r = fx.run_arrays
    if "JRJ" not in r:
        pytest.skip("fixture without radial data")
    case = W.CASES[fx.name]()
    zz = np.array([a.z for a in case.atoms], np.float64)
    vlm = r.get("VLM")
    return dict(rel=bool(int(r["REL"])), jrj=r["JRJ"], ro=r["RO"], dh=r["DH"], zz=zz, vr=r["VR"], vlm=vlm)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_atpar_matches_reference(name):
    fx = Fixture(name)
    rad = _radial(fx)
    r = fx.run_arrays
    ctx = L.Lapw1(fx.run())
    out = ctx.atpar(**rad)
    nt = int(r["LTOP"])
    for k, ref in (("p", "P"), ("dp", "DP"), ("pe", "PE"), ("dpe", "DPE"), ("pei", "PEI"), ("e", "E")):
        assert _rel(out[k][:nt], r[ref][:nt]) < 1e-11, k
    loor = np.asarray(r["LOOR"]).astype(bool)
    for k, ref in (("plo", "PLO"), ("dplo", "DPLO"), ("pi12lo", "PI12LO"), ("pe12lo", "PE12LO")):
        if loor.any():
            assert _rel(out[k][loor], r[ref][loor]) < 1e-11, k
    for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo"):
        ref = r[k.upper()]
        got = out[k]
        m = got != 0
        if m.any():
            assert _rel(got[m], ref[m]) < 1e-11, k
        # every integral the coupling list reads was written
        if int(r["INS"]) == 0 and k in ("vns1", "vns2", "vns3"):
            assert m.sum() > 0
    # the k-point with the device ATPAR's parameters
    arrays = dict(fx.run_arrays)
    for k, ref in (("p", "P"), ("dp", "DP"), ("pe", "PE"), ("dpe", "DPE"), ("pei", "PEI")):
        a = np.array(r[ref], order="F")
        a[:nt] = out[k][:nt]
        arrays[ref] = a
    for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo"):
        a = np.array(r[k.upper()], order="F")
        m = out[k] != 0
        a[m] = out[k][m]
        arrays[k.upper()] = a
    ctx.close()
    kd = fx.k(0)
    c2 = L.Lapw1(L.RunData(arrays, fx.is_complex))
    c2.set_basis(fx.basis(0))
    ne, nb = c2.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    assert ne == int(kd["NE"]) and nb == int(kd["NBELW"])
    ev, _ = c2.get_eig()
    assert np.max(np.abs(ev - kd["EIGVAL"])) < 1e-10
    c2.close()


def _fermi(ev, kt=0.001):
    mu = 0.5 * (ev[len(ev) // 2 - 1] + ev[len(ev) // 2])
    x = np.clip((ev - mu) / kt, -700, 700)
    return 2.0 / (1.0 + np.exp(x))


class _Vlm:
    """A k-point whose VNS integrals come from w2k_atpar on a given V_LM(r)."""

    def __init__(self, name):
        self.fx = Fixture(name)
        self.rad = _radial(self.fx)
        self.kd = self.fx.k(0)

    def arrays(self, ctx, vlm):
        """The run arrays with the VNS integrals w2k_atpar forms from vlm."""
        rad = dict(self.rad)
        rad["vlm"] = vlm
        out = ctx.atpar(**rad)
        arrays = dict(self.fx.run_arrays)
        for k in ("vns1", "vns2", "vns3", "vns1lo", "vns2lo", "vns3lo"):
            a = np.array(arrays[k.upper()], order="F")
            m = out[k] != 0
            a[m] = out[k][m]
            arrays[k.upper()] = a
        return arrays

    def run(self, vlm):
        ctx = L.Lapw1(self.fx.run())
        arrays = self.arrays(ctx, vlm)
        ctx.close()
        c2 = L.Lapw1(L.RunData(arrays, self.fx.is_complex))
        c2.set_basis(self.fx.basis(0))
        kd = self.kd
        ne, _ = c2.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
        return c2, ne


@pytest.mark.parametrize("name", ["c1_fcc", "lo_mult2", "lo_cplx"])
def test_vlm_gradient_matches_finite_differences(name):
    p = _Vlm(name)
    vlm = np.array(p.rad["vlm"], order="F")
    # the solve whose gradient is taken: VNS from the device ATPAR, radial
    # functions of that call kept on the context that runs the k-point
    ctx = L.Lapw1(p.fx.run())
    ctx.set_run(L.RunData(p.arrays(ctx, vlm), p.fx.is_complex))
    ctx.set_basis(p.fx.basis(0))
    ne, _ = ctx.calkpt(float(p.kd["EL"]), float(p.kd["EU"]), int(p.kd["NUME"]))
    ev, _ = ctx.get_eig()
    f = _fermi(ev)
    g = ctx.adjoint_vlm(f)
    ctx.close()
    assert g.shape == vlm.shape
    flat = np.argsort(-np.abs(g.ravel(order="F")))[:4]
    h = 1e-4
    for fl in flat:
        k, lm, a = np.unravel_index(fl, g.shape, order="F")

        def energy(delta):
            v = vlm.copy()
            v[k, lm, a] += delta
            c2, n2 = p.run(v)
            e2, _ = c2.get_eig()
            c2.close()
            assert n2 == ne
            return float(np.dot(f, e2))

        fd = (energy(h) - energy(-h)) / (2 * h)
        an = g[k, lm, a]
        assert abs(fd - an) <= 1e-6 * abs(an) + 1e-9, ((k, lm, a), fd, an)
