"""Device buffers for tests that hand device pointers to the library: the HIP
runtime the library itself links (hipMalloc / hipMemcpy / hipFree through
ctypes), so that one runtime owns the device in the process.  On the CPU
build of the library (W2K_TEST_EMU=1) a "device" buffer is host memory."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        lib = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7"))
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        _HIP = lib
    return _HIP


class DeviceArray:
    """A device copy of a host array (C order bytes), or host memory on the
    CPU build."""

    def __init__(self, a: np.ndarray, emu: bool):
        a = np.ascontiguousarray(a)
        self.dtype, self.shape, self.nbytes, self.emu = a.dtype, a.shape, a.nbytes, emu
        if emu:
            self.host = a.copy()
            self.ptr = self.host.ctypes.data
            return
        hip = _hip()
        p = ctypes.c_void_p()
        rc = hip.hipMalloc(ctypes.byref(p), max(self.nbytes, 8))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value
        rc = hip.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1)  # hipMemcpyHostToDevice
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")

    def get(self) -> np.ndarray:
        if self.emu:
            return self.host.copy()
        out = np.empty(self.shape, self.dtype)
        rc = _hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2)  # hipMemcpyDeviceToHost
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    def free(self):
        if not self.emu and self.ptr:
            _hip().hipFree(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
