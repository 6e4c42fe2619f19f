"""The library's HIP sources on the CPU (tests/cpuemu): the multi-k-point
chains with local orbitals, in which one context takes a basis that grows from
one k-point to the next, run with every device buffer against a guard page
(at its end, then at its start) and poisoned when allocated.  An access
outside a buffer, a use after free or a read of a never-written index faults
the child process (status 86) with the kernel, block and buffer named; a
deadlock or an unsupported wave operation ends it with status 87.  The child
also checks NE, NBELW and the eigenvalues against the reference, so the CPU
build is pinned to the same results as the device build.

Runs on the CPU only (no `gpu` marker); the emulated library is built with
make -C tests/cpuemu."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "cpuemu")

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, {tests!r})
from conftest import Fixture, PKG
L = PKG.lapw1
L._LIB = L.load_library(path={lib!r})
for name, nk in {cases!r}:
    fx = Fixture(name)
    ctx = L.Lapw1(fx.run())
    for ik in range(min(nk, fx.nk)):
        kd = fx.k(ik)
        ctx.set_basis(fx.basis(ik))
        ne, nb = ctx.calkpt(float(kd["EL"]), float(kd["EU"]), int(kd["NUME"]))
        ev, z = ctx.get_eig()
        de = float(np.max(np.abs(ev - kd["EIGVAL"])))
        print(name, ik, fx.basis(ik).n, ne, nb, de, flush=True)
        assert ne == int(kd["NE"]) and nb == int(kd["NBELW"]), (name, ik)
        assert de < 1e-10, (name, ik, de)
    ctx.close()
print("CHILD OK", flush=True)
"""


@pytest.fixture(scope="module")
def emu_lib():
    r = subprocess.run(["make", "-s", "-C", EMU_DIR, "-j8"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    assert r.returncode == 0, r.stdout.decode(errors="replace")[-3000:]
    lib = os.path.join(EMU_DIR, "_build", "libw2klapw1_emu.so")
    assert os.path.exists(lib)
    return lib


@pytest.mark.parametrize("align", ["end", "start"])
def test_multi_k_chains_in_bounds_on_cpu(emu_lib, align):
    cases = [("mgo", 3), ("lo_mult2", 2), ("c1_fcc", 2), ("lo_cplx", 2)]
    code = CHILD.format(tests=HERE, lib=emu_lib, cases=cases)
    env = dict(os.environ, W2K_EMU_ALIGN=align)
    r = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       timeout=900)
    out = r.stdout.decode(errors="replace")
    assert r.returncode == 0 and "CHILD OK" in out, out[-4000:]
