"""HAMILT / HNS on the MI355X against the reference golden vectors (and the
FP64 MFMA GEMM against numpy).  Bar: H and S elementwise to 1e-12 relative
(|diff| / max|M| per matrix)."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from conftest import GOLDEN, Fixture, relerr, CASES, PKG

pytestmark = pytest.mark.gpu
L = PKG.lapw1
TOL = 1e-12
# cells stored as sampled entries and sums (the matrices are hundreds of MB)
LARGE = [c for c in ("c2_rs64", "c3_lowsym100") if os.path.exists(os.path.join(GOLDEN, f"{c}.npz"))]


def _dgemm(ta, tb, m, n, k, alpha, A, B, beta, C, tri=0, atri=0, btri=0):
    lib = L.load_library()
    fn = lib.w2k_test_dgemm
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
    A = np.asfortranarray(A); B = np.asfortranarray(B); C = np.asfortranarray(C.copy())
    rc = fn(ta.encode(), tb.encode(), m, n, k, alpha, A.ctypes.data_as(dp), A.shape[0], B.ctypes.data_as(dp),
            B.shape[0], beta, C.ctypes.data_as(dp), C.shape[0], tri, atri, btri)
    assert rc == 0
    return C


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (130, 70, 129), (200, 3, 47)])
def test_dgemm_matches_numpy(ta, tb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    out = _dgemm(ta, tb, m, n, k, 0.7, A, B, -1.3, C)
    ref = 0.7 * opA @ opB - 1.3 * C
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(32, 32, 5951), (32, 1184, 5951), (7, 33, 2049), (130, 20, 4097), (32, 195, 700),
                                   (7, 33, 512)])
def test_dgemm_split_k_matches_numpy(ta, tb, m, n, k):
    """Few output tiles and a long K: the split-K path (partials reduced in a fixed order)."""
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    out = _dgemm_ws(ta, tb, m, n, k, 0.5, A, B, 2.0, C, 0, 0, 0, 1 << 22)
    ref = 0.5 * (A if ta == "N" else A.T) @ (B if tb == "N" else B.T) + 2.0 * C
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


def _dgemm_ws(ta, tb, m, n, k, alpha, A, B, beta, C, tri, atri, btri, wsn):
    lib = L.load_library()
    fn = lib.w2k_test_dgemm_ws
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_long]
    A = np.asfortranarray(A); B = np.asfortranarray(B); out = np.asfortranarray(C.copy())
    rc = fn(ta.encode(), tb.encode(), m, n, k, alpha, A.ctypes.data_as(dp), A.shape[0], B.ctypes.data_as(dp),
            B.shape[0], beta, out.ctypes.data_as(dp), out.shape[0], tri, atri, btri, wsn)
    assert rc == 0
    return out


@pytest.mark.parametrize("ta,tb,m,n,k,tri,atri,btri", [
    ("T", "N", 1100, 1100, 2100, 2, 0, 0), ("T", "N", 1100, 1100, 2100, 1, 0, 0),
    ("T", "N", 1300, 1250, 1300, 0, 2, 0), ("N", "N", 1300, 1250, 1300, 0, 1, 0),
    ("N", "N", 1250, 1300, 1300, 0, 0, 1), ("T", "N", 1100, 1100, 4500, 0, 0, 0)])
def test_dgemm_structured_split_k(ta, tb, m, n, k, tri, atri, btri):
    """Products with a triangular operand or output, 256..4096 output tiles
    and K >= 1024 are split into K slices of about 1024 (partials reduced in
    a fixed order, only the output triangle written); the last case has no
    structure and stays unsplit.  Elements outside an operand's triangle
    hold values that must not be read, and those outside the output
    triangle must keep their values."""
    rng = np.random.default_rng(m + 7 * k + 3 * tri + 5 * atri + 11 * btri)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    opA = A if ta == "N" else A.T
    Am = np.triu(opA) if atri == 1 else (np.tril(opA) if atri == 2 else opA)
    Bm = np.triu(B) if btri == 1 else (np.tril(B) if btri == 2 else B)
    out = _dgemm_ws(ta, tb, m, n, k, -0.5, A, B, 1.5, C, tri, atri, btri, 3 * m * n)
    ref = -0.5 * Am @ Bm + 1.5 * C
    if tri:
        keep = np.tril(np.ones((m, n), bool), -1) if tri == 2 else np.triu(np.ones((m, n), bool), 1)
        ref[keep] = C[keep]
    assert np.max(np.abs(out - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(256, 256, 256), (300, 257, 270), (513, 384, 401), (259, 700, 131)])
def test_dgemm_large_tiles_match_numpy(ta, tb, m, n, k):
    """Outputs of at least 256 x 256 with K >= 256 take the 128 x 128 tile
    kernel; ragged edges in every dimension (and one short-K case for the
    64 x 64 tiles)."""
    test_dgemm_matches_numpy(ta, tb, m, n, k)


@pytest.mark.parametrize("tri", [1, 2])
@pytest.mark.parametrize("n,k", [(150, 40), (389, 64), (389, 300)])
def test_dgemm_triangular_output_sizes(tri, n, k):
    rng = np.random.default_rng(n + k + tri)
    A = rng.standard_normal((n, k)); B = rng.standard_normal((n, k)); C = rng.standard_normal((n, n))
    out = _dgemm("N", "T", n, n, k, 1.0, A, B, 1.0, C, tri=tri)
    full = C + A @ B.T
    mask = np.tril(np.ones((n, n), bool)) if tri == 1 else np.triu(np.ones((n, n), bool))
    assert np.max(np.abs(out[mask] - full[mask])) < 1e-11
    assert np.array_equal(out[~mask], C[~mask])


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
@pytest.mark.parametrize("atri,btri", [(1, 0), (2, 0), (0, 1), (0, 2)])
def test_dgemm_triangular_operands_large(ta, tb, atri, btri):
    """Triangular op(A) / op(B) in the 128 x 128 tile kernel (k range clipped
    per tile, elements outside the triangle masked; the storage there holds
    values that must not be read)."""
    rng = np.random.default_rng(31 + atri + 3 * btri)
    m = n = k = 301
    A = rng.standard_normal((m, k)); B = rng.standard_normal((k, n)); C = np.zeros((m, n))
    opA = A if ta == "N" else A.T  # square: op(A) has the shape of A
    Am = np.triu(opA) if atri == 1 else (np.tril(opA) if atri == 2 else opA)
    Bm = np.triu(B) if btri == 1 else (np.tril(B) if btri == 2 else B)
    out = _dgemm(ta, tb, m, n, k, 1.0, A, B, 0.0, C, atri=atri, btri=btri)
    assert np.max(np.abs(out - Am @ Bm)) < 1e-11


def _dgemm_tiles(ta, tb, m, n, k, alpha, A, B, beta, C, tri, atri, btri, variant):
    lib = L.load_library()
    fn = lib.w2k_test_dgemm_tiles
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                   dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_double, dp, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    A = np.asfortranarray(A); B = np.asfortranarray(B); C = np.asfortranarray(C.copy())
    rc = fn(ta.encode(), tb.encode(), m, n, k, alpha, A.ctypes.data_as(dp), A.shape[0], B.ctypes.data_as(dp),
            B.shape[0], beta, C.ctypes.data_as(dp), C.shape[0], tri, atri, btri, variant)
    assert rc == 0
    return C


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k,tri,atri,btri", [(300, 257, 64, 0, 0, 0), (389, 389, 300, 1, 0, 0),
                                                 (389, 389, 200, 2, 0, 0), (301, 301, 301, 0, 1, 0),
                                                 (301, 301, 301, 0, 2, 0), (301, 301, 301, 0, 0, 1),
                                                 (301, 301, 301, 0, 0, 2), (130, 700, 999, 0, 0, 0)])
def test_dgemm_tile_kernels_bit_identical(ta, tb, m, n, k, tri, atri, btri):
    """The GEMM tile autotuner switches between the 64x64 and the 128x128 tile
    kernels by measured time; both accumulate in the same order, so their
    results must agree bit for bit (triangular outputs and operands
    included)."""
    rng = np.random.default_rng(m + 7 * n + 13 * k + tri + 3 * atri + 5 * btri)
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    C = rng.standard_normal((m, n))
    c0 = _dgemm_tiles(ta, tb, m, n, k, 0.7, A, B, -1.3, C, tri, atri, btri, 0)
    c1 = _dgemm_tiles(ta, tb, m, n, k, 0.7, A, B, -1.3, C, tri, atri, btri, 1)
    assert np.array_equal(c0, c1)
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    Am = np.triu(opA) if atri == 1 else (np.tril(opA) if atri == 2 else opA)
    Bm = np.triu(opB) if btri == 1 else (np.tril(opB) if btri == 2 else opB)
    ref = 0.7 * Am @ Bm - 1.3 * C
    mask = np.ones((m, n), bool) if tri == 0 else (np.tril(np.ones((m, n), bool)) if tri == 1
                                                   else np.triu(np.ones((m, n), bool)))
    assert np.max(np.abs(c1[mask] - ref[mask])) < 1e-11 * max(1.0, np.max(np.abs(ref)))


def test_dgemm_autotuned_calls_match():
    """Repeated calls of one shape class while the autotuner measures both
    kernels and then keeps one: every call gives the same bits."""
    rng = np.random.default_rng(77)
    m, n, k = 520, 530, 300
    A = rng.standard_normal((m, k)); B = rng.standard_normal((k, n)); C = rng.standard_normal((m, n))
    outs = [_dgemm("N", "N", m, n, k, 1.0, A, B, 0.5, C) for _ in range(5)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    assert np.max(np.abs(outs[0] - (A @ B + 0.5 * C))) < 1e-10


@pytest.mark.parametrize("tri", [1, 2])
def test_dgemm_triangular_output(tri):
    rng = np.random.default_rng(5)
    n, k = 150, 40
    A = rng.standard_normal((n, k)); B = rng.standard_normal((n, k)); C = rng.standard_normal((n, n))
    out = _dgemm("N", "T", n, n, k, 1.0, A, B, 1.0, C, tri=tri)
    full = C + A @ B.T
    mask = np.tril(np.ones((n, n), bool)) if tri == 1 else np.triu(np.ones((n, n), bool))
    assert np.max(np.abs(out[mask] - full[mask])) < 1e-12
    assert np.array_equal(out[~mask], C[~mask])


@pytest.mark.parametrize("atri,btri", [(1, 0), (2, 0), (0, 1), (0, 2)])
def test_dgemm_triangular_operands(atri, btri):
    rng = np.random.default_rng(9)
    m = n = k = 131
    A = rng.standard_normal((m, k)); B = rng.standard_normal((k, n)); C = np.zeros((m, n))
    Am = np.triu(A) if atri == 1 else (np.tril(A) if atri == 2 else A)
    Bm = np.triu(B) if btri == 1 else (np.tril(B) if btri == 2 else B)
    out = _dgemm("N", "N", m, n, k, 1.0, A, B, 0.0, C, atri=atri, btri=btri)
    assert np.max(np.abs(out - Am @ Bm)) < 1e-11


def _ctx(fx):
    ctx = L.Lapw1(fx.run())
    return ctx


@pytest.mark.parametrize("name", CASES)
def test_hamilt_matches_reference(name):
    fx = Fixture(name)
    ctx = _ctx(fx)
    for ik in range(fx.nk):
        if not fx.has_full(ik):
            continue
        kd = fx.k(ik)
        ctx.set_basis(fx.basis(ik))
        ctx.hamilt()
        hs, hd = ctx.get_hs()
        ref = kd["HS_HAMILT"]
        assert relerr(np.triu(hs), np.triu(ref)) < TOL, f"S k{ik}"
        assert relerr(np.tril(hs, -1), np.tril(ref, -1)) < TOL, f"H k{ik}"
        assert relerr(hd, kd["HSDIAG_HAMILT"]) < TOL, f"HSDIAG k{ik}"
        al, bl = ctx.get_albl()
        assert relerr(al, kd["AL"]) < TOL and relerr(bl, kd["BL"]) < TOL
    ctx.close()


@pytest.mark.parametrize("name", CASES)
def test_hns_matches_reference(name):
    fx = Fixture(name)
    ctx = _ctx(fx)
    kd = fx.k(0)
    ctx.set_basis(fx.basis(0))
    ctx.hamilt()
    ctx.swap_diag()
    ctx.hns()
    hs, hd = ctx.get_hs()
    ref = kd["HS_HNS"]
    assert relerr(np.tril(hs), np.tril(ref)) < TOL, "H after HNS"
    assert relerr(np.triu(hs, 1), np.triu(ref, 1)) < TOL, "S untouched"
    assert relerr(hd, kd["HSDIAG_HNS"]) < TOL
    ctx.close()


def _check_sampled(hs, hd, kd, stage):
    """HS/HSDIAG of one stage against the sampled reference entries (per
    triangle, relative to that triangle's largest entry), the reference's
    column and row sums and its HSDIAG.  The sums hold N entries, so the
    elementwise bar gives |d sum| <= N TOL max|HS|."""
    n = hs.shape[0]
    ii, jj = kd[f"SAMPLE_I_{stage}"], kd[f"SAMPLE_J_{stage}"]
    up = ii <= jj
    smax = np.max(np.abs(np.triu(hs))) if stage == "HAMILT" else np.max(np.abs(np.triu(hs, 1)))
    hmax = np.max(np.abs(np.tril(hs, -1))) if stage == "HAMILT" else np.max(np.abs(np.tril(hs)))
    d = np.abs(hs[ii, jj] - kd[f"SAMPLE_V_{stage}"])
    # after HNS the diagonal of HS holds diag(H) (CALKPT's swap)
    if stage == "HAMILT":
        s_up, h_lo = up, ~up
    else:
        s_up, h_lo = ii < jj, ii >= jj
    assert np.max(d[s_up]) / smax < TOL, f"S samples {stage}"
    assert np.max(d[h_lo]) / hmax < TOL, f"H samples {stage}"
    mx = float(kd[f"MAXABS_{stage}"])
    assert np.max(np.abs(hs.sum(axis=0) - kd[f"COLSUM_{stage}"])) < n * TOL * mx, f"column sums {stage}"
    assert np.max(np.abs(hs.sum(axis=1) - kd[f"ROWSUM_{stage}"])) < n * TOL * mx, f"row sums {stage}"
    assert relerr(hd, kd[f"HSDIAG_{stage}"]) < TOL, f"HSDIAG {stage}"


@pytest.mark.big
@pytest.mark.parametrize("name", LARGE)
def test_assembly_at_size_matches_reference_samples(name):
    """HAMILT and HNS of the large cells (N = 5951 for the 64-atom cell)
    against the reference's sampled entries, column/row sums and HSDIAG."""
    fx = Fixture(name)
    kd = fx.k(0)
    ctx = _ctx(fx)
    ctx.set_basis(fx.basis(0))
    ctx.hamilt()
    hs, hd = ctx.get_hs()
    _check_sampled(hs, hd, kd, "HAMILT")
    del hs
    ctx.swap_diag()
    ctx.hns()
    hs, hd = ctx.get_hs()
    _check_sampled(hs, hd, kd, "HNS")
    ctx.close()


def test_hns_alone_on_reference_input():
    """HNS applied to the reference's own post-HAMILT matrix (put_hs)."""
    fx = Fixture("lo_mult2")
    kd = fx.k(0)
    ctx = _ctx(fx)
    ctx.set_basis(fx.basis(0))
    ctx.hamilt()  # AL/BL tables
    hs0 = np.array(kd["HS_HAMILT"], order="F")
    d = np.diag(hs0).copy()
    np.fill_diagonal(hs0, kd["HSDIAG_HAMILT"])
    ctx.put_hs(hs0, d)
    ctx.hns()
    hs, hd = ctx.get_hs()
    assert relerr(np.tril(hs), np.tril(kd["HS_HNS"])) < TOL
    ctx.close()
