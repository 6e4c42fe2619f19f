"""The adjoint cell: 16-atom cubic cell (CsCl-type 2x2x2 supercell, real,
RKmax 7, N about 3030 with 64 local orbitals), 8 k-points, Fermi-Dirac
occupations at 1 mRy from one Fermi level over all k-points (k weights 1/8).
The band energy E = sum_k sum_n f_nk eps_nk (weights held fixed) and its
gradients with respect to
- the interstitial star coefficients POTK of the warped potential
  (w2k_set_stars / w2k_adjoint_stars), and
- the non-spherical potential V_LM(r_k) on the radial mesh (w2k_atpar /
  w2k_adjoint_vlm),
summed over the k-points, against central finite differences through the
library (h = 1e-6 Ry for POTK, 1e-4 for V_LM at single mesh points) to 1e-6
relative; the k-points on the device-built box against the reference
eigenvalues; and the warped-box gradient of one k-point against the band
density formed through FFTs (lapw2's fourir.F route)."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, Fixture, PKG, REPO

pytestmark = [pytest.mark.gpu, pytest.mark.big,
              pytest.mark.skipif(not os.path.exists(os.path.join(GOLDEN, "c5_cubic16.npz")),
                                 reason="c5_cubic16 fixture not generated")]
L = PKG.lapw1
sys.path.insert(0, os.path.join(REPO, "tests", "golden"))
import w2k_inputs as W  # noqa: E402

KT = 0.001  # Ry


class Cell:
    def __init__(self):
        self.fx = Fixture("c5_cubic16")
        r = self.fx.run_arrays
        self.stars = dict(ja=r["JA"], jb=r["JB"], jc=r["JC"], potk=np.asarray(r["POTK"], np.complex128),
                          imat=r["IMAT"], tau=r["TAU"])
        case = W.CASES["c5_cubic16"]()
        self.rad = dict(rel=bool(int(r["REL"])), jrj=r["JRJ"], ro=r["RO"], dh=r["DH"],
                        zz=np.array([a.z for a in case.atoms]), vr=r["VR"], vlm=r["VLM"])
        self.nk = self.fx.nk

    def arrays_with_vns(self, ctx, vlm):
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

    def solve_all(self, potk=None, vlm=None, keep=False):
        """Every k-point on one context; returns eigenvalues per k (and the
        context, solved at the last k-point, when keep)."""
        ctx = L.Lapw1(self.fx.run())
        if vlm is not None:
            ctx.set_run(L.RunData(self.arrays_with_vns(ctx, vlm), False))
        st = dict(self.stars)
        if potk is not None:
            st["potk"] = potk
        ctx.set_stars(**st)
        evs = []
        for q in range(self.nk):
            kd = self.fx.k(q)
            ctx.set_basis(self.fx.basis(q))
            ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
            evs.append(ctx.get_eig()[0])
        if keep:
            return ctx, evs
        ctx.close()
        return evs

    def weights(self, evs):
        allev = np.sort(np.concatenate(evs))
        nocc = len(allev) // 2
        mu = 0.5 * (allev[nocc - 1] + allev[nocc])
        return [(2.0 / self.nk) / (1.0 + np.exp(np.clip((e - mu) / KT, -700, 700))) for e in evs]


@pytest.fixture(scope="module")
def cell():
    return Cell()


def test_c5_kpoints_on_device_box_match_reference(cell):
    evs = cell.solve_all()
    for q in range(cell.nk):
        kd = cell.fx.k(q)
        assert len(evs[q]) == int(kd["NE"])
        assert np.max(np.abs(evs[q] - kd["EIGVAL"])) < 1e-10, q


def test_c5_star_gradient_matches_finite_differences(cell):
    evs = cell.solve_all()
    f = cell.weights(evs)
    ctx = L.Lapw1(cell.fx.run())
    ctx.set_stars(**cell.stars)
    g = np.zeros(len(cell.stars["ja"]), np.complex128)
    e0 = 0.0
    for q in range(cell.nk):
        kd = cell.fx.k(q)
        ctx.set_basis(cell.fx.basis(q))
        ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
        e, gq = ctx.adjoint_stars(f[q])
        e0 += e
        g += gq
    ctx.close()
    assert np.all(g.imag == 0.0)  # real build: the box is read through its real part

    def energy(potk):
        ev = cell.solve_all(potk=potk)
        return sum(float(np.dot(f[q], ev[q])) for q in range(cell.nk))

    h = 1e-6
    pick = list(dict.fromkeys([0] + list(np.argsort(-np.abs(g))[:4])))
    for s in pick:
        p = cell.stars["potk"].copy()
        p[s] += h
        ep = energy(p)
        p[s] -= 2 * h
        em = energy(p)
        fd = (ep - em) / (2 * h)
        assert abs(fd - g[s].real) <= 1e-6 * abs(g[s].real) + 2e-8, (s, fd, g[s])


def test_c5_vlm_gradient_matches_finite_differences(cell):
    vlm = np.array(cell.rad["vlm"], order="F")
    evs = cell.solve_all(vlm=vlm)
    f = cell.weights(evs)
    ctx = L.Lapw1(cell.fx.run())
    ctx.set_run(L.RunData(cell.arrays_with_vns(ctx, vlm), False))
    ctx.set_stars(**cell.stars)
    g = None
    for q in range(cell.nk):
        kd = cell.fx.k(q)
        ctx.set_basis(cell.fx.basis(q))
        ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
        gq = ctx.adjoint_vlm(f[q])
        g = gq if g is None else g + gq
    ctx.close()

    def energy(v):
        ev = cell.solve_all(vlm=v)
        return sum(float(np.dot(f[q], ev[q])) for q in range(cell.nk))

    h = 1e-4
    for fl in np.argsort(-np.abs(g.ravel(order="F")))[:3]:
        k, lm, a = np.unravel_index(fl, g.shape, order="F")
        v = vlm.copy()
        v[k, lm, a] += h
        ep = energy(v)
        v[k, lm, a] -= 2 * h
        em = energy(v)
        fd = (ep - em) / (2 * h)
        assert abs(fd - g[k, lm, a]) <= 1e-6 * abs(g[k, lm, a]) + 1e-9, ((k, lm, a), fd, g[k, lm, a])


def test_c5_warp_gradient_is_the_fft_band_density(cell):
    from test_gpu_stars import _fft_density
    ctx, evs = cell.solve_all(keep=True)
    q = cell.nk - 1
    f = cell.weights(evs)[q]
    _, g = ctx.adjoint_warp(f)
    _, z = ctx.get_eig()
    ctx.close()
    kd = cell.fx.k(q)
    w = [(s - 1) // 2 for s in g.shape]
    rho, dims = _fft_density(np.asarray(kd["KZZ"]), z.astype(np.complex128), f, int(kd["NV"]), w)
    worst = 0.0
    for qx in range(-w[0], w[0] + 1):
        for qy in range(-w[1], w[1] + 1):
            for qz in range(-w[2], w[2] + 1):
                gq = g[qx + w[0], qy + w[1], qz + w[2]].real
                gm = g[-qx + w[0], -qy + w[1], -qz + w[2]].real
                r = rho[qx % dims[0], qy % dims[1], qz % dims[2]].real
                d = abs(gq - r) if (qx, qy, qz) == (0, 0, 0) else abs(gq + gm - 2 * r)
                worst = max(worst, d)
    assert worst < 1e-12 * max(1.0, abs(rho[0, 0, 0]))
