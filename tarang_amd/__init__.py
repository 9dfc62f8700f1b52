"""tarang_amd: MI355X-native pseudo-spectral time step (Tarang FOUR / SCFT bases).

The product is the C-ABI library libtarang_hip.so (include/tarang_hip.h) built for gfx950.
This package only locates / builds it and gives Python a thin ctypes view of it (tests, bench).
There is no CPU fallback: if the library is missing, loading fails loudly.
"""
from .abi import (Context, Ensemble, Plan, State, field_read, field_read_reduced, field_write,  # noqa: F401
                  field_write_reduced, lib, LIB_PATH)
