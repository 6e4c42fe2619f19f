"""Shared test setup: the `gpu` marker, the package loader and the fixtures."""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "hack-wein2k_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")

sys.path.insert(0, os.path.join(REPO, "tests"))


from oracle_lib import load_package  # noqa: E402

PKG = load_package()

# W2K_TEST_EMU=1: the `gpu` tests run on the CPU build of the library's HIP
# sources (tests/cpuemu; make -C tests/cpuemu) instead of the device library,
# so that the whole device test sequence can be replayed on the CPU with every
# device buffer against a guard page.  Tests marked `big` (cells too large for
# the emulation, or programs linked to the device library) are skipped then.
EMU = os.environ.get("W2K_TEST_EMU") == "1"
EMU_LIB = os.path.join(REPO, "tests", "cpuemu", "_build", "libw2klapw1_emu.so")
if EMU:
    PKG.lapw1._LIB = PKG.lapw1.load_library(path=EMU_LIB)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "big: too large for the CPU emulation of the device library")


def pytest_collection_modifyitems(config, items):
    if not EMU:
        return
    skip = pytest.mark.skip(reason="too large for the CPU emulation (W2K_TEST_EMU=1)")
    for it in items:
        if "big" in it.keywords:
            it.add_marker(skip)


class Fixture:
    """One golden case: run-level state, per-k data, metadata."""

    def __init__(self, name: str):
        self.name = name
        z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
        self.raw = {k: z[k] for k in z.files}
        self.is_complex = bool(int(self.raw["meta/complex"]))
        self.nk = int(self.raw["meta/nk"])
        self.run_arrays = {k[4:]: v for k, v in self.raw.items() if k.startswith("run/")}

    def run(self):
        RunData = PKG.lapw1.RunData
        return RunData(self.run_arrays, self.is_complex)

    def k(self, i: int) -> dict:
        p = f"k{i}/"
        return {k[len(p):]: v for k, v in self.raw.items() if k.startswith(p)}

    def basis(self, i: int):
        Basis = PKG.lapw1.Basis
        kd = self.k(i)
        return Basis(int(kd["NV"]), int(kd["NLO"]), kd["KZZ"], kd["XK"], kd["YK"], kd["ZK"], kd["RK"])

    def has_full(self, i: int) -> bool:
        return f"k{i}/HS_HAMILT" in self.raw


CASES = ["c1_fcc", "mgo", "lo_mult2", "lo_cplx", "rs8", "ins_lo", "ins_lo_cplx"]


@pytest.fixture(scope="session", params=CASES)
def fixture_case(request):
    return Fixture(request.param)


def load_fixture(name: str) -> Fixture:
    return Fixture(name)


def relerr(a, b) -> float:
    """max |a-b| / max |b| (the per-matrix relative measure used for H and S)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))
