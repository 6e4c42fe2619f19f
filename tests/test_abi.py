"""C-ABI library: builds for gfx950, loads, exports every declared symbol
(no device calls -- CPU suite)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from conftest import REPO, PKG

LIB = os.path.join(REPO, "hack-wein2k_amd", "libw2klapw1.so")


def _declared(header):
    txt = open(os.path.join(REPO, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(w2k_[a-z0-9_]+)\s*\(", txt)))


def test_library_exists():
    assert os.path.exists(LIB), "build the library first (python hack-wein2k_amd/build.py)"


@pytest.mark.parametrize("header", ["w2k_lapw1.h", "w2k_lapw1_test.h"])
def test_library_exports_declared_symbols(header):
    lib = ctypes.CDLL(LIB)
    names = _declared(header)
    assert names
    for nm in names:
        assert hasattr(lib, nm), nm


def test_version_string():
    lib = PKG.lapw1.load_library()
    assert b"gfx950" in lib.w2k_version()


def test_run_descriptor_layout_matches_header():
    """The ctypes mirror has one field per member of w2k_run_desc, in order."""
    txt = open(os.path.join(REPO, "include", "w2k_lapw1.h")).read()
    body = txt[txt.index("typedef struct w2k_run_desc {"):txt.index("} w2k_run_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";")[:-1]:
        decl = decl.split("{")[-1].strip()
        if not decl:
            continue
        for part in decl.split(","):
            nm = part.strip().split()[-1].lstrip("*")
            names.append(nm)
    fields = [f[0] for f in PKG.lapw1.RunDesc._fields_]
    assert fields == names
