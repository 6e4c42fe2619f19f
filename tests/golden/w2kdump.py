"""Reader for the stream files written by oracle/ref_dump.F90.

Record: name (32 bytes), kind (int32: 1=int32, 2=float64, 3=complex128),
ndim (int32), dims (ndim x int32), then the data in Fortran (column-major) order.
"""
from __future__ import annotations

import numpy as np

_DT = {1: np.int32, 2: np.float64, 3: np.complex128}


def read_dump(path: str) -> dict:
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    while off < len(buf):
        name = buf[off:off + 32].decode().strip()
        off += 32
        kind, ndim = (int(x) for x in np.frombuffer(buf, np.int32, 2, off))
        off += 8
        dims = tuple(int(x) for x in np.frombuffer(buf, np.int32, ndim, off))
        off += 4 * ndim
        dt = np.dtype(_DT[kind])
        cnt = int(np.prod(dims))
        arr = np.frombuffer(buf, dt, cnt, off).reshape(dims, order="F").copy()
        off += cnt * dt.itemsize
        out[name] = arr[0] if dims == (1,) else arr
    return out
