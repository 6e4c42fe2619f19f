"""Split-precision HNS: the rank-K update on the INT8 matrix cores with 7-bit
digit slices and FP64 recombination (w2k_set_split_precision), an extended
form with no counterpart in the reference.

Pinned three ways:
- exact integer products: operands whose digit expansions are exact give
  results identical to the exact product (bit for bit), which fixes the digit
  split, the recombination weights and the output layout;
- random FP64 operands with widely varying row scales stay inside the a-priori
  bound |error(i,j)| <= K (S + 2) 2^(E_i + F_j - 7S);
- HNS and the whole k-point in split precision against the reference's golden
  vectors: H to a slice-dependent tolerance, eigenvalues to 1e-8 Ry (the
  split-precision bar) with identical NE and NBELW."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from conftest import Fixture, relerr, CASES, PKG, GOLDEN

pytestmark = pytest.mark.gpu
L = PKG.lapw1
DE_BAR = 1e-8  # Ry, split precision vs FP64 reference eigenvalues


def _split(A, B, C, tri=0, accumulate=1, slices=5):
    lib = L.load_library()
    fn = lib.w2k_test_split_gemm
    dp = ctypes.POINTER(ctypes.c_double)
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int,
                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    out = np.asfortranarray(C.copy())
    m, k = A.shape
    n = B.shape[0]
    rc = fn(m, n, k, A.ctypes.data_as(dp), m, B.ctypes.data_as(dp), n, out.ctypes.data_as(dp), m, tri, accumulate,
            slices)
    assert rc == 0
    return out


def _ints(rng, shape, bits):
    return rng.integers(-(2 ** bits) + 1, 2 ** bits, size=shape).astype(np.int64)


@pytest.mark.parametrize("slices", [3, 4, 5, 6, 7])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (64, 64, 64), (70, 33, 130), (200, 150, 1000)])
def test_split_gemm_exact_on_exact_digits(slices, m, n, k):
    """A: integers of at most min(7S, 35) bits (exact in S digits); B: integers
    in [-127, 127] (exact in one digit).  Every kept digit pair is then exact
    and every dropped pair is zero: the result must equal A B^T exactly."""
    rng = np.random.default_rng(1000 * slices + m + n + k)
    bits = min(7 * slices, 35)
    A = _ints(rng, (m, k), bits)
    B = rng.integers(-127, 128, size=(n, k)).astype(np.int64)
    A[m // 2, :] = 0  # an all-zero row
    exact = A @ B.T   # int64, |.| < 2^53
    C0 = rng.standard_normal((m, n))
    out = _split(A.astype(np.float64), B.astype(np.float64), C0, tri=0, accumulate=0, slices=slices)
    assert np.array_equal(out, exact.astype(np.float64))
    # roles swapped: A one digit, B several
    out2 = _split(B.astype(np.float64), A.astype(np.float64), C0.T.copy(), tri=0, accumulate=0, slices=slices)
    assert np.array_equal(out2, exact.T.astype(np.float64))


@pytest.mark.parametrize("slices", [3, 5])
def test_split_gemm_lower_triangle_and_accumulate(slices):
    rng = np.random.default_rng(77)
    n, k = 150, 300
    A = _ints(rng, (n, k), 14).astype(np.float64)
    B = rng.integers(-127, 128, size=(n, k)).astype(np.float64)
    C = rng.integers(-1000, 1000, size=(n, n)).astype(np.float64)
    out = _split(A, B, C, tri=1, accumulate=1, slices=slices)
    full = C + A @ B.T
    low = np.tril(np.ones((n, n), bool))
    assert np.array_equal(out[low], full[low])
    assert np.array_equal(out[~low], C[~low])


def _row_expo(X):
    mx = np.max(np.abs(X), axis=1)
    e = np.full(X.shape[0], -1100, np.int64)
    nz = mx > 0
    e[nz] = np.frexp(mx[nz])[1]  # mx < 2^e
    return e


@pytest.mark.parametrize("slices", [3, 4, 5, 6, 7])
def test_split_gemm_error_bound_random(slices):
    rng = np.random.default_rng(31 + slices)
    m, n, k = 190, 170, 777
    A = rng.standard_normal((m, k)) * (10.0 ** rng.uniform(-6, 6, size=(m, 1)))
    B = rng.standard_normal((n, k)) * (10.0 ** rng.uniform(-6, 6, size=(n, 1)))
    A[:, ::7] *= 1e-9  # small entries inside a row
    out = _split(A, B, np.zeros((m, n)), tri=0, accumulate=0, slices=slices)
    ref = A @ B.T  # FP64 rounding of the reference is far below the bound
    ea, eb = _row_expo(A), _row_expo(B)
    bound = k * (slices + 2) * np.ldexp(1.0, (ea[:, None] + eb[None, :] - 7 * slices).astype(np.int64))
    err = np.abs(out - ref)
    assert np.all(err <= bound + 1e-15 * np.abs(ref)), float(np.max(err / bound))


# H after HNS in split precision vs the reference (|diff| / max|H|): each
# digit gains a factor 2^7; six digits meet the FP64 parity bar (1e-12)
H_TOL = {3: 3e-8, 4: 3e-10, 5: 5e-12, 6: 1e-12}


@pytest.mark.parametrize("slices", [3, 4, 5, 6])
@pytest.mark.parametrize("name", CASES)
def test_split_hns_matches_reference(name, slices):
    fx = Fixture(name)
    ctx = L.Lapw1(fx.run())
    ctx.set_split_precision(slices)
    kd = fx.k(0)
    ctx.set_basis(fx.basis(0))
    ctx.hamilt()
    ctx.swap_diag()
    ctx.hns()
    hs, hd = ctx.get_hs()
    err = relerr(np.tril(hs), np.tril(kd["HS_HNS"]))
    print(f"{name} slices={slices}: H rel err {err:.3e}")
    assert err < H_TOL[slices]
    assert relerr(np.triu(hs, 1), np.triu(kd["HS_HNS"], 1)) < 1e-12, "S untouched"
    ctx.close()


@pytest.mark.parametrize("slices", [4, 5])
@pytest.mark.parametrize("name", CASES)
def test_split_calkpt_eigenvalues(name, slices):
    """Whole k-points with the split-precision HNS: identical NE, NBELW and
    eigenvalues within 1e-8 Ry of the reference."""
    fx = Fixture(name)
    ctx = L.Lapw1(fx.run())
    ctx.set_split_precision(slices)
    kd = fx.k(0)
    ctx.set_basis(fx.basis(0))
    ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    assert ne == int(kd["NE"]) and nb == int(kd["NBELW"])
    ev, _ = ctx.get_eig()
    de = float(np.max(np.abs(ev - kd["EIGVAL"]))) if ne else 0.0
    print(f"{name} slices={slices}: max |dE| = {de:.3e} Ry")
    assert de < DE_BAR
    ctx.close()


@pytest.mark.big
def test_split_c2_rocksalt_eigenvalues():
    """The 64-atom rocksalt k-point (N=5951) with the split-precision HNS
    (4 digits, the setting bench.py reports)."""
    raw = np.load(os.path.join(GOLDEN, "c2_rs64.npz"))
    run = L.RunData({k[4:]: raw[k] for k in raw.files if k.startswith("run/")}, bool(int(raw["meta/complex"])))
    kd = {k[3:]: raw[k] for k in raw.files if k.startswith("k0/")}
    ctx = L.Lapw1(run)
    ctx.set_split_precision(4)
    ctx.set_basis(L.Basis(int(kd["NV"]), int(kd["NLO"]), kd["KZZ"], kd["XK"], kd["YK"], kd["ZK"], kd["RK"]))
    ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
    assert ne == int(kd["NE"]) and nb == int(kd["NBELW"])
    ev, _ = ctx.get_eig()
    de = float(np.max(np.abs(ev - kd["EIGVAL"])))
    print(f"c2_rs64: max |dE| = {de:.3e} Ry")
    assert de < DE_BAR
    ctx.close()


def test_split_precision_rejects_bad_slices():
    fx = Fixture("c1_fcc")
    ctx = L.Lapw1(fx.run())
    for bad in (1, 2, 8, -1):
        with pytest.raises(L.Lapw1Error):
            ctx.set_split_precision(bad)
    ctx.set_split_precision(0)
    ctx.close()
