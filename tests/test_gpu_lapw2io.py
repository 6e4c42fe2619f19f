"""What lapw2 reads from a solved k-point: the vector file (w2k_write_vector,
TAPEWF's records, tapewf.F:45-56) and the interstitial band density
(w2k_density_interstitial, the quantity lapw2's fourir.F forms on an FFT grid).

Pinned by:
- the reference's own vector files (c1_fcc: 10 k-points, real; lo_cplx: 2
  k-points, complex, Z conjugated): the file written by the library for the
  same k-points parses into the same records, the header and KZZ records
  byte for byte, the eigenvalues within 1e-10 Ry and the vectors of isolated
  eigenvalues within 1e-8 up to a unit phase;
- the density against the one formed through FFTs of psi_n(r) (the lapw2
  route, independent of the library's pair sums) to 1e-12."""
from __future__ import annotations

import os
import struct

import numpy as np
import pytest

from conftest import Fixture, PKG
from test_gpu_stars import _fft_density

pytestmark = pytest.mark.gpu
L = PKG.lapw1


def _records(b: bytes):
    out, o = [], 0
    while o < len(b):
        (n,) = struct.unpack_from("<i", b, o)
        out.append(b[o + 4:o + 4 + n])
        (m,) = struct.unpack_from("<i", b, o + 4 + n)
        assert m == n
        o += 8 + n
    return out


@pytest.mark.parametrize("name", ["c1_fcc", "lo_cplx"])
def test_vector_file_matches_reference(name, tmp_path):
    fx = Fixture(name)
    ref = _records(fx.raw["vector/case_vector"].tobytes())
    nat = int(fx.run_arrays["NAT"])
    ref = ref[2 * nat:]  # the per-atom E/ELO records of the main program precede the k-points
    path = str(tmp_path / "case.vector")
    open(path, "wb").close()
    ctx = L.Lapw1(fx.run())
    cpx = fx.is_complex
    heads = []
    pos = 0
    for q in range(fx.nk):
        h = ref[pos]
        sx, sy, sz = struct.unpack("<3d", h[:24])
        kname = h[24:34].decode()
        nvt, ne = struct.unpack("<2i", h[34:42])
        (weight,) = struct.unpack("<d", h[42:50])
        ipgr = h[50:53].decode()
        heads.append((pos, ne))
        pos += 2 + 2 * ne
        kd = fx.k(q)
        ctx.set_basis(fx.basis(q))
        ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
        ctx.write_vector(path, kname, sx, sy, sz, weight, ipgr, append=True)
    ctx.close()
    mine = _records(open(path, "rb").read())
    assert len(mine) == len(ref)
    dt = np.complex128 if cpx else np.float64
    for (p0, ne) in heads:
        assert mine[p0] == ref[p0], "header record"
        assert mine[p0 + 1] == ref[p0 + 1], "KZZ record"
        evm, evr, zm, zr = [], [], [], []
        for i in range(ne):
            a, b = mine[p0 + 2 + 2 * i], ref[p0 + 2 + 2 * i]
            assert a[:4] == b[:4]
            evm.append(struct.unpack("<d", a[4:12])[0])
            evr.append(struct.unpack("<d", b[4:12])[0])
            zm.append(np.frombuffer(mine[p0 + 3 + 2 * i], dt))
            zr.append(np.frombuffer(ref[p0 + 3 + 2 * i], dt))
        evm, evr = np.array(evm), np.array(evr)
        assert np.max(np.abs(evm - evr)) < 1e-10
        gaps = np.diff(evr)
        for j in range(ne):
            g = min(gaps[j - 1] if j > 0 else np.inf, gaps[j] if j < ne - 1 else np.inf)
            if g > 1e-6:
                ph = np.vdot(zm[j], zr[j])
                ph = ph / abs(ph)
                assert np.max(np.abs(zm[j] * ph - zr[j])) < 1e-8 * np.max(np.abs(zr[j])), j


def _fermi(ev, kt=0.001):
    mu = 0.5 * (ev[len(ev) // 2 - 1] + ev[len(ev) // 2])
    return 2.0 / (1.0 + np.exp(np.clip((ev - mu) / kt, -700, 700)))


@pytest.mark.parametrize("name", ["c1_fcc", "lo_cplx", "rs8"])
def test_interstitial_density_matches_fft(name):
    fx = Fixture(name)
    kd = fx.k(0)
    ctx = L.Lapw1(fx.run())
    b = fx.basis(0)
    ctx.set_basis(b)
    ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    ev, z = ctx.get_eig()
    f = _fermi(ev)
    rho = ctx.density_interstitial(f)
    ctx.close()
    w = [(s - 1) // 2 for s in rho.shape]
    ref, dims = _fft_density(np.asarray(kd["KZZ"]), z.astype(np.complex128), f, b.nv, w)
    worst = 0.0
    for qx in range(-w[0], w[0] + 1):
        for qy in range(-w[1], w[1] + 1):
            for qz in range(-w[2], w[2] + 1):
                r = ref[qx % dims[0], qy % dims[1], qz % dims[2]]
                if not fx.is_complex:
                    r = r.real
                worst = max(worst, abs(rho[qx + w[0], qy + w[1], qz + w[2]] - r))
    assert worst < 1e-12 * max(1.0, abs(ref[0, 0, 0]))
