"""The k-point's basis on the device (w2k_coors: COORS, coors.f:47-201, and
LOPW, lopw.f:40-128).

Pinned by:
- the reference's own bases (KZZ, XK, YK, ZK, RK of every k-point of the
  fixtures, local-orbital rows included), bit for bit: orthogonal (ORTHO,
  KZZ from PIA) and general axes, real and complex builds, one to eight
  equivalent atoms per local orbital;
- the oracle (oracle_coors.c, itself pinned against those bases), which
  keeps the list as the reference does (bounded insertion, NVEC = 25 box with
  the axis check) while the device ranks a gathered sphere, bit for bit on
  seeded random k-points, on a CXZ lattice and on lists cut at NMATMAX
  (capacities from one above the sphere's count to a third of it);
- properties of the cut (coors.f:149-155): the cut list is a prefix of the
  uncut one and ends at a gap wider than 1e-5;
- the whole k-point built on the device basis: HAMILT after w2k_coors equals
  HAMILT on the reference basis, and the eigenvalues the reference's
  (1e-10 Ry);
- the CXZ rule (coors.f:135-139) is checked against the oracle only: no
  reference fixture has a CXZ lattice (parity unpinned for that branch)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import EMU, Fixture, PKG
import oracle_lib as O

pytestmark = pytest.mark.gpu
L = PKG.lapw1

CASES = ["c1_fcc", "mgo", "lo_mult2", "lo_cplx", "rs8", "ins_lo", "ins_lo_cplx", "c4_ortho8", "c5_cubic16"]


def _lattice(fx, nmatmax=13000, rkm=None):
    r = fx.run_arrays
    return L.Lattice.make(r["BR2"], r["PIA"], int(r["ORTHO"]), int(r["CXZ"]),
                          float(r["RKM"]) if rkm is None else rkm, nmatmax)


def _same(a, b):
    return (a.nv == b.nv and a.nlo == b.nlo and np.array_equal(a.kzz, b.kzz)
            and all(np.array_equal(getattr(a, f), getattr(b, f)) for f in ("xk", "yk", "zk", "rk")))


def _diff(a, b):
    if a.nv != b.nv or a.nlo != b.nlo:
        return f"NV {a.nv} vs {b.nv}, NLO {a.nlo} vs {b.nlo}"
    bad = {f: int(np.sum(getattr(a, f) != getattr(b, f))) for f in ("xk", "yk", "zk", "rk")}
    bad["kzz"] = int(np.sum(np.any(a.kzz != b.kzz, axis=0)))
    return str(bad)


@pytest.mark.parametrize("name", CASES)
def test_basis_matches_reference(name):
    fx = Fixture(name)
    if "BR2" not in fx.run_arrays:
        pytest.skip("fixture without the lattice")
    lat = _lattice(fx)
    ctx = L.Lapw1(fx.run())
    nk = fx.nk if not EMU else min(fx.nk, 4)
    for i in range(nk):
        kd = fx.k(i)
        b = ctx.coors(lat, float(kd["SX"]), float(kd["SY"]), float(kd["SZ"]))
        ref = fx.basis(i)
        assert _same(b, ref), f"{name} k{i}: {_diff(b, ref)}"


@pytest.mark.big
@pytest.mark.parametrize("name", ["c2_rs64", "c3_lowsym100"])
def test_large_cell_basis_matches_reference(name):
    """64 atoms (N = 5951) and 100 atoms with 400 local orbitals
    (N = 12447)."""
    fx = Fixture(name)
    ctx = L.Lapw1(fx.run())
    kd = fx.k(0)
    b = ctx.coors(_lattice(fx), float(kd["SX"]), float(kd["SY"]), float(kd["SZ"]))
    ref = fx.basis(0)
    assert _same(b, ref), _diff(b, ref)


@pytest.mark.parametrize("name", ["lo_mult2", "lo_cplx", "c1_fcc"])
def test_random_kpoints_match_oracle(name):
    fx = Fixture(name)
    lat = _lattice(fx)
    run = fx.run()
    ctx = L.Lapw1(run)
    nlo = int(fx.k(0)["NLO"])
    rng = np.random.default_rng(7)
    for _ in range(6 if not EMU else 2):
        s = rng.uniform(-0.5, 0.5, 3)
        b = ctx.coors(lat, *s)
        ref = O.coors(run, lat, *s, nlo)
        assert _same(b, ref), f"k = {s}: {_diff(b, ref)}"


@pytest.mark.parametrize("name", ["lo_mult2", "rs8"])
def test_cut_at_nmatmax_matches_oracle(name):
    fx = Fixture(name)
    run = fx.run()
    ctx = L.Lapw1(run)
    kd = fx.k(0)
    s = (float(kd["SX"]), float(kd["SY"]), float(kd["SZ"]))
    nlo = int(kd["NLO"])
    full = O.coors(run, _lattice(fx), *s, nlo)
    m = full.nv  # vectors inside the sphere
    # NDM1 = NMATMAX + 1 - NLO around and below the sphere's count
    ndm1s = sorted({m + 1, m, m - 1, m - 2, (2 * m) // 3, m // 2, m // 3})
    for ndm1 in ndm1s:
        lat = _lattice(fx, nmatmax=ndm1 - 1 + nlo)
        b = ctx.coors(lat, *s)
        ref = O.coors(run, lat, *s, nlo)
        assert _same(b, ref), f"NDM1 = {ndm1}: {_diff(b, ref)}"
        nv = b.nv
        assert nv <= max(m, 1)
        # prefix of the uncut list, cut at a gap wider than 1e-5
        assert np.array_equal(b.kzz[:, :nv], full.kzz[:, :nv])
        assert np.array_equal(b.rk[:nv], full.rk[:nv])
        if nv < m:
            assert full.rk[nv] - full.rk[nv - 1] > 1e-5


def test_cxz_rule_matches_oracle():
    """A monoclinic base-centred (CXZ) lattice: KZZ = (JA+JC, JB, -JA+JC)."""
    fx = Fixture("lo_mult2")
    run = fx.run()
    a, b, c, gam = 9.0, 10.0, 11.0, np.deg2rad(100.0)
    # reciprocal vectors of a C-centred (x-z) monoclinic cell, in 2 pi / a units
    br2 = 2 * np.pi * np.array([[1 / a, 0.0, 1 / a],
                                [-np.cos(gam) / (a * np.sin(gam)), 1 / (b * np.sin(gam)), -np.cos(gam) / (a * np.sin(gam))],
                                [-1 / c, 0.0, 1 / c]]).T
    lat = L.Lattice.make(br2, 2 * np.pi / np.array([a, b, c]), False, True, 3.0)
    ctx = L.Lapw1(run)
    nlo = int(fx.k(0)["NLO"])
    for s in ((0.0, 0.0, 0.0), (0.125, 0.25, -0.375)):
        bd = ctx.coors(lat, *s)
        ref = O.coors(run, lat, *s, nlo)
        assert _same(bd, ref), f"k = {s}: {_diff(bd, ref)}"


@pytest.mark.parametrize("name", ["lo_mult2", "lo_cplx"])
def test_kpoint_on_device_basis_matches_reference(name):
    fx = Fixture(name)
    kd = fx.k(0)
    s = (float(kd["SX"]), float(kd["SY"]), float(kd["SZ"]))
    ctx = L.Lapw1(fx.run())
    ctx.coors(_lattice(fx), *s)
    ctx.hamilt()
    hs, hd = ctx.get_hs()
    ref = L.Lapw1(fx.run())
    ref.set_basis(fx.basis(0))
    ref.hamilt()
    hs2, hd2 = ref.get_hs()
    assert np.array_equal(hs, hs2) and np.array_equal(hd, hd2)
    ctx.coors(_lattice(fx), *s)
    ne, _ = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    ev, _ = ctx.get_eig()
    ref_ev = np.atleast_1d(kd["EIGVAL"])
    assert ne == len(ref_ev)
    assert np.max(np.abs(ev - ref_ev)) < 1e-10


def test_lopw_out_of_plane_waves_is_an_error():
    """A sphere with fewer plane waves than a local-orbital block needs:
    LOPW's fatal path (lopw.f:139) as W2K_EINVAL, the context usable after."""
    fx = Fixture("lo_mult2")
    ctx = L.Lapw1(fx.run())
    r = fx.run_arrays
    # a radius that holds a handful of vectors around Gamma
    small = L.Lattice.make(r["BR2"], r["PIA"], int(r["ORTHO"]), int(r["CXZ"]), 1.0e-3)
    with pytest.raises(L.Lapw1Error) as ei:
        ctx.coors(small, 0.0, 0.0, 0.0)
    assert ei.value.code == -1
    kd = fx.k(0)
    b = ctx.coors(_lattice(fx), float(kd["SX"]), float(kd["SY"]), float(kd["SZ"]))
    assert _same(b, fx.basis(0))
