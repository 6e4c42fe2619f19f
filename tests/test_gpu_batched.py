"""Batched k-points (w2k_calkpt_batched): an extended form with no counterpart
in the reference, whose k-point loop is serial (lapw1.F:50-66).

Pinned per k-point against the unbatched path: every k-point of a batch must
give exactly (bit for bit) the NE, NBELW, eigenvalues and eigenvectors of a
w2k_calkpt call on its own, and the reference eigenvalues to 1e-10 Ry.  The
batches are ragged (N differs between k-points) and run on several streams."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import Fixture, PKG

pytestmark = pytest.mark.gpu
L = PKG.lapw1
TOL_E = 1e-10


def _single(fx, q, slices=0):
    """k-point q through the unbatched path on a fresh context."""
    kd = fx.k(q)
    ctx = L.Lapw1(fx.run())
    ctx.set_split_precision(slices)
    ctx.set_basis(fx.basis(q))
    ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    ev, z = ctx.get_eig()
    ctx.close()
    return ne, nb, ev, z


def _batch(fx, ks, nworkers, slices=0, ctx=None):
    own = ctx is None
    if own:
        ctx = L.Lapw1(fx.run())
    ctx.set_split_precision(slices)
    kds = [fx.k(q) for q in ks]
    nume = max(int(kd["NUME"]) for kd in kds)
    out = ctx.calkpt_batched([fx.basis(q) for q in ks], [float(kd["EL"]) for kd in kds],
                             [float(kd["EU"]) for kd in kds], nume, nworkers=nworkers)
    if own:
        ctx.close()
    return out


def _check(fx, ks, out, slices=0, ref_tol=TOL_E):
    for q, (ne, nb, ev, z) in zip(ks, out):
        kd = fx.k(q)
        ne1, nb1, ev1, z1 = _single(fx, q, slices)
        assert (ne, nb) == (ne1, nb1), f"k{q}"
        assert np.array_equal(ev, ev1), f"k{q}: eigenvalues differ from the unbatched path"
        assert np.array_equal(z, z1), f"k{q}: eigenvectors differ from the unbatched path"
        assert ne == int(kd["NE"]) and nb == int(kd["NBELW"]), f"k{q}"
        if ne:
            assert np.max(np.abs(ev - kd["EIGVAL"])) < ref_tol, f"k{q}"


@pytest.mark.parametrize("nworkers", [0, 3, 1])
def test_batched_c1_fcc_ten_kpoints(nworkers):
    """Ten k-points with N from 38 to 51 (several per stream with 1 or 3 streams)."""
    fx = Fixture("c1_fcc")
    ks = list(range(fx.nk))
    _check(fx, ks, _batch(fx, ks, nworkers))


def test_batched_mgo_one_kpoint_per_stream():
    """Twelve MgO k-points (local orbitals, N from 228 to 243), one per stream."""
    fx = Fixture("mgo")
    ks = list(range(fx.nk))
    _check(fx, ks, _batch(fx, ks, len(ks)))


def test_batched_complex_lo_cplx():
    fx = Fixture("lo_cplx")
    ks = list(range(fx.nk))
    _check(fx, ks, _batch(fx, ks, len(ks)))


def test_batched_split_precision():
    """The split-precision HNS inside a batch: identical to its unbatched run."""
    fx = Fixture("c1_fcc")
    ks = list(range(fx.nk))
    _check(fx, ks, _batch(fx, ks, 4, slices=4), slices=4, ref_tol=1e-8)


def test_batched_reuses_streams_and_subsets():
    """A second batch on the same context (streams and workspaces reused, a
    subset in another order) gives the same results."""
    fx = Fixture("c1_fcc")
    ctx = L.Lapw1(fx.run())
    ks1 = list(range(fx.nk))
    out1 = _batch(fx, ks1, 4, ctx=ctx)
    ks2 = [7, 2, 9, 0]
    out2 = _batch(fx, ks2, 2, ctx=ctx)
    ctx.close()
    for i, q in enumerate(ks2):
        assert out2[i][0] == out1[q][0]
        assert np.array_equal(out2[i][2], out1[q][2]) and np.array_equal(out2[i][3], out1[q][3])


@pytest.mark.big
def test_batched_c4_sixty_four_kpoints():
    """The batched-k cell: 64 complex k-points, N from 1936 to 1978 (no local
    orbitals), all in one batch; NE, NBELW and eigenvalues against the
    reference for every k-point, and four of them bit-identical to single calls."""
    fx = Fixture("c4_ortho8")
    ks = list(range(fx.nk))
    out = _batch(fx, ks, 0)
    worst = 0.0
    for q, (ne, nb, ev, z) in zip(ks, out):
        kd = fx.k(q)
        assert ne == int(kd["NE"]) and nb == int(kd["NBELW"]), f"k{q}"
        worst = max(worst, float(np.max(np.abs(ev - kd["EIGVAL"]))))
    print(f"c4_ortho8: 64 k-points, max |dE| = {worst:.3e} Ry")
    assert worst < TOL_E
    for q in (0, 17, 40, 63):
        ne1, nb1, ev1, z1 = _single(fx, q)
        assert np.array_equal(out[q][2], ev1) and np.array_equal(out[q][3], z1), f"k{q}"


def test_batched_rejects_bad_arguments():
    fx = Fixture("c1_fcc")
    ctx = L.Lapw1(fx.run())
    b = fx.basis(0)
    with pytest.raises(L.Lapw1Error):
        ctx.calkpt_batched([b], [-1.0], [1.0], 0)  # nume < 1
    bad = L.Basis(b.nv, 3, b.kzz, b.xk, b.yk, b.zk, b.rk)  # LO count of another run
    with pytest.raises(L.Lapw1Error):
        ctx.calkpt_batched([bad], [-1.0], [1.0], 10)
    ctx.close()
