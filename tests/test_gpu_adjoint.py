"""Band-energy adjoint with respect to the warped interstitial potential
(w2k_adjoint_warp) and to the non-spherical radial integrals VNS
(w2k_adjoint_vns): an extended form with no counterpart in the reference.

E_band = sum_n f_n eps_n with the weights held fixed; by Hellmann-Feynman its
gradient with respect to a box entry WARP(Q) is a sum of the eigenvector
products Gamma = sum_n f_n z_n z_n^H over the plane-wave pairs with
KZZ_J - KZZ_I = Q.  Pinned by:
- central finite differences of E_band through the whole k-point (the
  library itself, with the box entry moved by +-h and +-h/2, h = 1e-4,
  Richardson extrapolated) to 1e-6 relative, real and imaginary parts,
  Q = 0 included (with h = 1e-3 the curvature of the complex cases already
  shows at 1e-6);
- the same sum formed on the host from the returned eigenvectors, i.e. the
  interstitial Fourier components of the band density that lapw2 would build
  from these eigenvectors;
- a torch.autograd.Function over a few Hermitian perturbation directions,
  checked with torch.autograd.gradcheck."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import Fixture, PKG

pytestmark = pytest.mark.gpu
L = PKG.lapw1


def _fermi(ev, kt=0.001):
    mu = 0.5 * (ev[0] + ev[-1])
    x = np.clip((ev - mu) / kt, -700, 700)
    return 2.0 / (1.0 + np.exp(x))


class _Point:
    """One k-point of a fixture, recomputed for modified WARP boxes."""

    def __init__(self, name, q=0):
        self.fx = Fixture(name)
        self.kd = self.fx.k(q)
        self.basis = self.fx.basis(q)
        self.el, self.eu, self.nume = float(self.kd["EL"]), float(self.kd["EU"]), int(self.kd["NUME"])

    def solve(self, warp=None, replace=None):
        arrays = dict(self.fx.run_arrays)
        if warp is not None:
            arrays["WARP"] = warp
        if replace:
            arrays.update(replace)
        ctx = L.Lapw1(L.RunData(arrays, self.fx.is_complex))
        ctx.set_basis(self.basis)
        ne, _ = ctx.calkpt(self.el, self.eu, self.nume)
        ev, z = ctx.get_eig()
        return ctx, ne, ev, z

    def energy(self, warp, f, replace=None):
        ctx, ne, ev, _ = self.solve(warp, replace)
        ctx.close()
        assert ne == len(f), "the window lost or gained an eigenvalue"
        return float(np.dot(f, ev))


def _fd(pt, w0, idx, dirn, f, h):
    def e(t):
        w = w0.copy()
        w[idx] += t * dirn
        return pt.energy(w, f)
    d1 = (e(h) - e(-h)) / (2 * h)
    d2 = (e(h / 2) - e(-h / 2)) / h
    return (4 * d2 - d1) / 3


@pytest.mark.parametrize("name,weights", [("c1_fcc", "ones"), ("c1_fcc", "fermi"), ("lo_cplx", "ones"),
                                          ("lo_cplx", "fermi"), ("rs8", "fermi")])
def test_adjoint_warp_matches_finite_differences(name, weights):
    pt = _Point(name)
    ctx, ne, ev, _ = pt.solve()
    f = np.ones(ne) if weights == "ones" else _fermi(ev)
    e0, g = ctx.adjoint_warp(f)
    ctx.close()
    assert abs(e0 - float(np.dot(f, ev))) <= 1e-12 * max(1.0, abs(e0))
    w0 = np.asarray(pt.fx.run_arrays["WARP"], np.complex128)
    assert g.shape == w0.shape
    centre = tuple(s // 2 for s in w0.shape)
    # Q = 0 and the entries with the largest gradients
    flat = np.argsort(-np.abs(g).ravel(order="F"))
    picks = [centre] + [np.unravel_index(k, g.shape, order="F") for k in flat[:3]]
    cpx = pt.fx.is_complex
    for idx in picks:
        for dirn, comp in ((1.0, g[idx].real), (1j, g[idx].imag)):
            if dirn == 1j and (not cpx or idx == centre):
                assert comp == 0.0  # the real build and the diagonal see Re WARP only
                continue
            fd = _fd(pt, w0, idx, dirn, f, 1e-4)
            print(f"{name} {weights} Q-index {idx} {'Im' if dirn == 1j else 'Re'}: adjoint {comp:.12e} fd {fd:.12e}")
            assert abs(fd - comp) <= 1e-6 * abs(comp) + 1e-9, (idx, dirn, comp, fd)


@pytest.mark.parametrize("name", ["c1_fcc", "lo_cplx"])
def test_adjoint_warp_matches_host_density(name):
    """The gradient is the interstitial band density in plane-wave pairs:
    formed here on the host from the eigenvectors and the basis."""
    pt = _Point(name)
    ctx, ne, ev, z = pt.solve()
    f = _fermi(ev)
    _, g = ctx.adjoint_warp(f)
    ctx.close()
    nv = pt.basis.nv
    kzz = pt.basis.kzz[:, :nv].T.astype(np.int64)
    zz = z[:nv, :]
    gam = (zz * f[None, :]) @ zz.conj().T            # Gamma(I, J)
    w1, w2, w3 = (s // 2 for s in g.shape)
    ref = np.zeros_like(g)
    for i in range(nv):
        for j in range(i + 1):
            q = kzz[j] - kzz[i]
            if abs(q[0]) > w1 or abs(q[1]) > w2 or abs(q[2]) > w3:
                continue
            t = (q[0] + w1, q[1] + w2, q[2] + w3)
            ref[t] += gam[i, i].real if i == j else 2.0 * np.conj(gam[i, j])
    if not pt.fx.is_complex:
        ref = ref.real.astype(np.complex128)
    assert np.max(np.abs(g - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_adjoint_warp_torch_gradcheck():
    """A torch.autograd.Function around the library: E_band as a function of
    three Hermitian perturbations of the box (W(Q) and W(-Q) together)."""
    torch = pytest.importorskip("torch")
    pt = _Point("c1_fcc")
    ctx, ne, ev, _ = pt.solve()
    f = _fermi(ev)
    _, g0 = ctx.adjoint_warp(f)
    ctx.close()
    w0 = np.asarray(pt.fx.run_arrays["WARP"], np.complex128)
    centre = np.array([s // 2 for s in w0.shape])
    flat = np.argsort(-np.abs(g0).ravel(order="F"))
    dirs = []
    for k in flat[1:]:
        q = np.array(np.unravel_index(k, w0.shape, order="F")) - centre
        if not q.any():
            continue
        d = np.zeros_like(w0)
        d[tuple(centre + q)] += 1.0
        d[tuple(centre - q)] += 1.0
        dirs.append(d)
        if len(dirs) == 3:
            break

    class BandEnergy(torch.autograd.Function):
        @staticmethod
        def forward(fctx, p):
            w = w0 + sum(float(pi) * d for pi, d in zip(p.detach().numpy(), dirs))
            c, ne1, ev1, _ = pt.solve(w)
            assert ne1 == len(f)
            e, g = c.adjoint_warp(f)
            c.close()
            fctx.save_for_backward(torch.tensor([float(np.sum((np.conj(g) * d).real)) for d in dirs],
                                                dtype=torch.float64))
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(fctx, ge):
            (grad,) = fctx.saved_tensors
            return ge * grad

    p = torch.zeros(3, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(BandEnergy.apply, (p,), eps=1e-4, atol=1e-7, rtol=1e-5)


VNS_KEYS = ("VNS1", "VNS2", "VNS3", "VNS1LO", "VNS2LO", "VNS3LO")


@pytest.mark.parametrize("name,weights", [("c1_fcc", "fermi"), ("lo_mult2", "fermi"), ("lo_cplx", "ones"),
                                          ("lo_cplx", "fermi"), ("rs8", "ones")])
def test_adjoint_vns_matches_finite_differences(name, weights):
    """dE_band/dVNS from the library against central differences of E_band
    through the whole k-point (HNS included), h = 1e-5 with Richardson
    extrapolation (near-degenerate pairs in lo_cplx give enough curvature
    that h = 1e-4 is off by 6e-6), on the entries with the largest gradients
    of every array (the LO arrays included where the cell has local
    orbitals); the floor 2e-8 is the rounding of E_band over 2h."""
    pt = _Point(name)
    ctx, ne, ev, _ = pt.solve()
    f = np.ones(ne) if weights == "ones" else _fermi(ev)
    g = ctx.adjoint_vns(f)
    ctx.close()
    checked = 0
    for key in VNS_KEYS:
        gk = g[key.lower()]
        v0 = np.asarray(pt.fx.run_arrays[key], np.float64)
        assert gk.shape == v0.reshape(gk.shape, order="F").shape
        if not np.any(gk):
            continue
        flat = np.argsort(-np.abs(gk).ravel(order="F"))[:2]
        for k in flat:
            idx = np.unravel_index(k, gk.shape, order="F")

            def e(t):
                v = v0.reshape(gk.shape, order="F").copy(order="F")
                v[idx] += t
                return pt.energy(None, f, {key: v})
            h = 1e-5
            d1 = (e(h) - e(-h)) / (2 * h)
            d2 = (e(h / 2) - e(-h / 2)) / h
            fd = (4 * d2 - d1) / 3
            print(f"{name} {weights} {key}{tuple(int(i) for i in idx)}: adjoint {gk[idx]:.12e} fd {fd:.12e}")
            assert abs(fd - gk[idx]) <= 1e-6 * abs(gk[idx]) + 2e-8, (key, idx, gk[idx], fd)
            checked += 1
    assert checked >= 2
    if pt.fx.run_arrays.get("NLO") is not None and int(np.asarray(pt.fx.run_arrays["NLO"])) == 0:
        assert not any(np.any(g[k.lower()]) for k in VNS_KEYS[3:])


def test_adjoint_vns_needs_hns():
    pt = _Point("c1_fcc")
    ctx = L.Lapw1(pt.fx.run())
    ctx.set_basis(pt.basis)
    ne, _ = ctx.calkpt(pt.el, pt.eu, pt.nume, nohns=True)
    with pytest.raises(L.Lapw1Error):
        ctx.adjoint_vns(np.ones(ne))
    ctx.close()
