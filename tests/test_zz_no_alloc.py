"""w2k_seclr4_noalloc: SECLR4 on caller-owned device buffers with every
workspace carved from the library's static scratch (no device allocation in
the call).  Checked against w2k_seclr4 on the same HS/HSDIAG (eigenvalues
within 1e-10 Ry with the same NE/NBELW, eigenvectors within 1e-8 up to a
phase, S-orthonormality), on two calls in a row on one context, the second
with a smaller order; the context's own eigenpairs stay untouched; an order
beyond the scratch is refused before any work is queued."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import EMU, Fixture, PKG
from devmem import DeviceArray

pytestmark = pytest.mark.gpu
L = PKG.lapw1


def _to_device(a):
    """(holder, address) of a device copy of a host array (host memory on the
    CPU build of the library)."""
    d = DeviceArray(a, EMU)
    return d, d.ptr


def _to_host(holder, dtype, shape):
    return holder.get().view(dtype).reshape(shape)


def _pack(H, S):
    """HS/HSDIAG as CALKPT leaves them (complex: H conjugated in the lower triangle)."""
    if np.iscomplexobj(H):
        hs = np.triu(S, 1) + np.tril(H, -1).conj() + np.diag(np.diag(H))
        return np.asfortranarray(hs), np.diag(S).copy()
    return np.asfortranarray(np.triu(S, 1) + np.tril(H)), np.diag(S).copy()


def _unpack(hs, hd):
    if np.iscomplexobj(hs):
        Hl = np.tril(hs, -1).conj()
        H = Hl + Hl.conj().T + np.diag(np.diag(hs).real)
        Su = np.triu(hs, 1)
        return H, Su + Su.conj().T + np.diag(np.asarray(hd).real)
    return np.tril(hs) + np.tril(hs, -1).T, np.triu(hs, 1) + np.triu(hs, 1).T + np.diag(hd)


def _noalloc(ctx, hs, hd, el, eu, nume):
    lib = ctx.lib
    lib.w2k_seclr4_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    n = hs.shape[0]
    cplx = np.iscomplexobj(hs)
    dt = np.complex128 if cplx else np.float64
    # column-major HS: the transpose's C order is the Fortran order
    hs_h, hs_p = _to_device(np.asarray(hs, dt).T.copy())
    hd_h, hd_p = _to_device(np.asarray(hd, dt))
    ldz = n + 3
    m = min(nume, n) if nume > 0 else n
    ev_h, ev_p = _to_device(np.full(max(m, 1), np.nan))
    z_h, z_p = _to_device(np.zeros((max(m, 1), ldz), dt))
    ne, nb = ctypes.c_int(-1), ctypes.c_int(-1)
    rc = lib.w2k_seclr4_noalloc(ctx.ctx, n, hs_p, hd_p, el, eu, nume, ctypes.byref(ne), ctypes.byref(nb), ev_p, z_p,
                                ldz)
    if rc != 0:
        return rc, 0, 0, None, None
    ev = _to_host(ev_h, np.float64, (max(m, 1),))[:ne.value]
    z = _to_host(z_h, dt, (max(m, 1), ldz))[:ne.value, :n].T
    return rc, ne.value, nb.value, ev, z


def _regular(run, hs, hd, el, eu, nume):
    n = hs.shape[0]
    ctx = L.Lapw1(run)
    zz = np.zeros(n)
    ctx.set_basis(L.Basis(n, 0, np.zeros((3, n), np.int32), zz, zz, zz, zz))
    ctx.put_hs(hs, hd)
    ne, nb = ctx.seclr4(el, eu, nume)
    ev, z = ctx.get_eig()
    ctx.close()
    return ne, nb, ev, z


def _compare(got, ref, hs, hd):
    rc, ne, nb, ev, z = got
    ne_r, nb_r, ev_r, z_r = ref
    assert rc == 0 and ne == ne_r and nb == nb_r
    assert np.max(np.abs(ev - ev_r)) <= 1e-10
    H, S = _unpack(hs, hd)
    gaps = np.diff(ev_r)
    for j in range(ne):
        g = min(gaps[j - 1] if j > 0 else np.inf, gaps[j] if j < ne - 1 else np.inf)
        if g > 1e-6:
            ph = np.vdot(z[:, j], z_r[:, j])
            ph = ph / abs(ph)
            assert np.max(np.abs(z[:, j] * ph - z_r[:, j])) <= 1e-8 * np.max(np.abs(z_r[:, j])), j
    G = z.conj().T @ S @ z
    assert np.max(np.abs(G - np.eye(ne))) < 1e-10


def _pencil(n, cplx, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    S = X @ X.conj().T / n + np.eye(n)
    Y = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    return (Y + Y.conj().T) / 2, S


@pytest.mark.parametrize("name,small", [("mgo", 97), ("lo_cplx", 83)])
def test_seclr4_noalloc_matches_seclr4_twice(name, small):
    fx = Fixture(name)
    kd = fx.k(0)
    run = fx.run()
    # first call: the reference's post-HNS matrix of the cell
    hs1 = np.array(kd["HS_HNS"], order="F")
    hd1 = np.array(kd["HSDIAG_HNS"])
    el, eu, nume = float(kd["EL"]), float(kd["EU"]), int(kd["NUME"])
    # second call, on the same context: a smaller random pencil, full window
    H2, S2 = _pencil(small, fx.is_complex, 7)
    hs2, hd2 = _pack(H2, S2)
    ref1 = _regular(run, hs1, hd1, el, eu, nume)
    ref2 = _regular(run, hs2, hd2, 1.0, 0.0, 40)
    ctx = L.Lapw1(run)
    # the context's own eigenpairs, to check that the calls leave them alone
    ctx.set_basis(fx.basis(0))
    ctx.put_hs(hs1, hd1)
    ctx.seclr4(el, eu, nume)
    own_ev, own_z = ctx.get_eig()
    _compare(_noalloc(ctx, hs1, hd1, el, eu, nume), ref1, hs1, hd1)
    _compare(_noalloc(ctx, hs2, hd2, 1.0, 0.0, 40), ref2, hs2, hd2)
    ev, z = ctx.get_eig()
    assert np.array_equal(ev, own_ev) and np.array_equal(z, own_z)
    ctx.close()


def test_seclr4_noalloc_refuses_orders_beyond_scratch():
    fx = Fixture("c1_fcc")
    ctx = L.Lapw1(fx.run())
    n = 769
    hold, p = _to_device(np.zeros(16))
    lib = ctx.lib
    ne, nb = ctypes.c_int(0), ctypes.c_int(0)
    lib.w2k_seclr4_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    # the order is checked before any buffer is touched
    rc = lib.w2k_seclr4_noalloc(ctx.ctx, n, p, p, 1.0, 0.0, 10, ctypes.byref(ne), ctypes.byref(nb), p, p, n)
    assert rc == -1  # W2K_EINVAL
    assert b"scratch" in lib.w2k_last_error()
    ctx.close()
