"""coral_amd: CORAL's BLAS call surface on MI355X (gfx950) through libcoralhip.so.

- ``coral_amd.blas``   -- CORAL-named functions over host numpy buffers (drop-in form)
- ``coral_amd.device`` -- device buffers, streams, events for the stream-ordered C ABI
- ``coral_amd._lib``   -- the ctypes binding, generated from include/coral_hip.h
"""
from ._lib import LIB_PATH, PROTOTYPES, CoralError, call, load, status_of  # noqa: F401

__all__ = ["LIB_PATH", "PROTOTYPES", "CoralError", "call", "load", "status_of"]
