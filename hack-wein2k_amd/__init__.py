"""MI355X-native lapw1 hot path (HAMILT, HNS, SECLR4) behind a C-ABI.

Layout:
  csrc/            HIP kernels for gfx950 and the extern "C" entry points
                   (built into libw2klapw1.so next to this file)
  fortran/         ISO_C_BINDING drop-in for the reference's HAMILT/HNS/SECLR4
  lapw1.py         Python mirror of the interface over the C-ABI
  build.py         hipcc build of the library
"""
from . import lapw1  # noqa: F401
