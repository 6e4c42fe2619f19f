"""Interstitial stars on the device (w2k_set_stars: STERN, stern.F:1-97, and
SETWAR's box expansion, setwar.f:40-74) and the band-energy gradient with
respect to the star coefficients POTK (w2k_adjoint_stars, an extended form
with no counterpart in the reference).

Pinned by:
- the warped box built from the reference's own stars and space group
  against the reference's WARP box (every entry, 1e-14 absolute);
- the whole k-point on that box against the reference eigenvalues (1e-10 Ry);
- central finite differences of E_band = sum f_n eps_n (weights fixed,
  Fermi-Dirac at 1 mRy) through the library, with POTK_s moved by +-h
  (h = 1e-6 Ry), real and imaginary parts, to 1e-6 relative;
- the interstitial band density of the returned eigenvectors formed as lapw2
  does (fourir.F: psi_n(r) on an FFT grid, rho(r) = sum f_n |psi_n|^2, back
  to G), an algorithm independent of the library's pair sums: the warped-box
  gradient g satisfies g(Q) + conj g(-Q) = 2 rho(Q) for Q /= 0, g(0) = rho(0)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import Fixture, PKG

pytestmark = pytest.mark.gpu
L = PKG.lapw1

CASES = ["c1_fcc", "lo_mult2", "lo_cplx", "rs8"]


def _has_stars(fx):
    return "NKK" in fx.run_arrays


def _stars(fx):
    r = fx.run_arrays
    return dict(ja=r["JA"], jb=r["JB"], jc=r["JC"], potk=np.asarray(r["POTK"], np.complex128),
                imat=r["IMAT"], tau=r["TAU"])


def _fermi(ev, kt=0.001):
    mu = 0.5 * (ev[len(ev) // 2 - 1] + ev[len(ev) // 2])
    x = np.clip((ev - mu) / kt, -700, 700)
    return 2.0 / (1.0 + np.exp(x))


def _solve(fx, q, st, potk=None):
    ctx = L.Lapw1(fx.run())
    s = dict(st)
    if potk is not None:
        s["potk"] = potk
    ctx.set_stars(**s)
    kd = fx.k(q)
    ctx.set_basis(fx.basis(q))
    ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    return ctx, ne, nb


@pytest.mark.parametrize("name", CASES)
def test_device_box_matches_reference_warp(name):
    fx = Fixture(name)
    if not _has_stars(fx):
        pytest.skip("fixture without star data")
    ctx = L.Lapw1(fx.run())
    ctx.set_stars(**_stars(fx))
    box = ctx.get_warp()
    ref = fx.run_arrays["WARP"]
    w = [(s - 1) // 2 for s in box.shape]
    v = [(s - 1) // 2 for s in ref.shape]
    # the reference box is cropped to its nonzero entries: compare on the union
    m = [max(a, b) for a, b in zip(w, v)]
    A = np.zeros([2 * x + 1 for x in m], np.complex128)
    B = np.zeros_like(A)
    A[m[0] - w[0]:m[0] + w[0] + 1, m[1] - w[1]:m[1] + w[1] + 1, m[2] - w[2]:m[2] + w[2] + 1] = box
    B[m[0] - v[0]:m[0] + v[0] + 1, m[1] - v[1]:m[1] + v[1] + 1, m[2] - v[2]:m[2] + v[2] + 1] = ref
    assert np.max(np.abs(A - B)) < 1e-14
    # the k-point on the device box against the reference
    kd = fx.k(0)
    ctx.set_basis(fx.basis(0))
    ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    assert ne == int(kd["NE"]) and nb == int(kd["NBELW"])
    ev, _ = ctx.get_eig()
    assert np.max(np.abs(ev - kd["EIGVAL"])) < 1e-10
    ctx.close()


@pytest.mark.parametrize("name", CASES)
def test_star_gradient_matches_finite_differences(name):
    fx = Fixture(name)
    if not _has_stars(fx):
        pytest.skip("fixture without star data")
    st = _stars(fx)
    ctx, ne, _ = _solve(fx, 0, st)
    ev, _ = ctx.get_eig()
    f = _fermi(ev)
    e0, g = ctx.adjoint_stars(f)
    ctx.close()
    assert abs(e0 - float(np.dot(f, ev))) < 1e-12 * max(1.0, abs(e0))
    # the stars with the largest gradients, and the first star (G = 0)
    pick = list(dict.fromkeys([0] + list(np.argsort(-np.abs(g))[:3])))
    h = 1e-6
    parts = [(1.0, "re")] + ([(1j, "im")] if fx.is_complex else [])

    def energy(potk):
        c, n2, _ = _solve(fx, 0, st, potk)
        e2, _ = c.get_eig()
        c.close()
        assert n2 == ne
        return float(np.dot(f, e2))

    for s in pick:
        for unit, part in parts:
            p = st["potk"].copy()
            p[s] += h * unit
            ep = energy(p)
            p[s] -= 2 * h * unit
            em = energy(p)
            fd = (ep - em) / (2 * h)
            an = g[s].real if part == "re" else g[s].imag
            assert abs(fd - an) <= 1e-6 * abs(an) + 2e-8, (s, part, fd, an)


def _fft_density(kzz, z, f, nv, box):
    """rho(Q) = sum_n f_n sum over (I, J) with KZZ_J - KZZ_I = Q of
    conj z_I z_J, through psi_n(r) on an FFT grid (lapw2's fourir.F route).
    The grid holds every difference of the basis and every Q of the box
    (half-widths `box`) without wrap-around."""
    k = kzz[:, :nv].T
    lo = k.min(axis=0)
    span = k.max(axis=0) - lo + 1
    dims = tuple(int(max(2 * s, s + w)) for s, w in zip(span, box))
    rho = np.zeros(dims)
    rho = rho.astype(np.complex128)
    idx = tuple((k - lo).T)
    for n in range(z.shape[1]):
        c = np.zeros(dims, np.complex128)
        c[idx] = z[:nv, n]
        psi = np.fft.ifftn(c) * np.prod(dims)
        rho += f[n] * np.fft.fftn(np.abs(psi) ** 2) / np.prod(dims)
    return rho, dims


@pytest.mark.parametrize("name", ["c1_fcc", "lo_cplx"])
def test_warp_gradient_is_the_fft_band_density(name):
    fx = Fixture(name)
    kd = fx.k(0)
    ctx = L.Lapw1(fx.run())
    b = fx.basis(0)
    ctx.set_basis(b)
    ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    ev, z = ctx.get_eig()
    f = _fermi(ev)
    _, g = ctx.adjoint_warp(f)
    ctx.close()
    w = [(s - 1) // 2 for s in g.shape]
    rho, dims = _fft_density(np.asarray(kd["KZZ"]), z.astype(np.complex128), f, b.nv, w)
    worst = 0.0
    scale = abs(rho[0, 0, 0])
    for qx in range(-w[0], w[0] + 1):
        for qy in range(-w[1], w[1] + 1):
            for qz in range(-w[2], w[2] + 1):
                gq = g[qx + w[0], qy + w[1], qz + w[2]]
                gm = g[-qx + w[0], -qy + w[1], -qz + w[2]]
                r = rho[qx % dims[0], qy % dims[1], qz % dims[2]]
                if (qx, qy, qz) == (0, 0, 0):
                    d = abs(gq - r)
                elif fx.is_complex:
                    d = abs(gq + np.conj(gm) - 2 * r)
                else:
                    d = abs((gq + gm).real - 2 * r.real)
                worst = max(worst, d)
    assert worst < 1e-12 * max(1.0, scale)
