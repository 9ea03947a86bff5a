"""ctypes binding of libcoralhip.so.

Prototypes are read from ``include/coral_hip.h`` so the Python signatures cannot
drift from the C ABI.  Loading fails loudly when the library has not been built:
there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcoralhip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "coral_hip.h")


class CoralError(RuntimeError):
    """Non-zero status from a libcoralhip entry point."""

    def __init__(self, func: str, status: int, detail: str = ""):
        self.func = func
        self.status = status
        if status < 0:
            msg = f"{func}: invalid argument {-status}"
        else:
            msg = f"{func}: status {status}"
        if detail:
            msg += f" ({detail})"
        super().__init__(msg)


_CTYPE = {
    "int32_t": ctypes.c_int32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "size_t": ctypes.c_size_t,
    "float": ctypes.c_float,
    "double": ctypes.c_double,
    "int": ctypes.c_int,
}


def _param_type(decl: str):
    decl = decl.strip()
    if decl == "void":
        return None
    if "*" in decl or "[" in decl or "coral_stream_t" in decl:
        if re.match(r"^const\s+char\s*\*", decl):
            return ctypes.c_char_p
        return ctypes.c_void_p
    toks = decl.replace("const ", "").split()
    return _CTYPE[toks[0]]


def parse_header(path: str = HEADER_PATH) -> dict:
    """Return {name: (restype, [argtypes])} for every function in the header."""
    with open(path) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(int|const char \*)\s*(coralh?_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        restype = ctypes.c_char_p if "char" in ret else ctypes.c_int
        args = []
        for p in params.split(","):
            t = _param_type(" ".join(p.split()))
            if t is not None:
                args.append(t)
        protos[name] = (restype, args)
    return protos


PROTOTYPES = parse_header()
_lib = None


def load():
    """Load libcoralhip.so and attach prototypes.  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950).  There is no CPU fallback."
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (restype, argtypes) in PROTOTYPES.items():
        fn = getattr(lib, name)  # AttributeError = header/library mismatch
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def call(name: str, *args) -> None:
    """Call an entry point and raise CoralError on a non-zero status."""
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        detail = lib.coral_last_error().decode(errors="replace") if status > 0 else ""
        raise CoralError(name, status, detail)


def status_of(name: str, *args) -> int:
    """Call an entry point and return its raw status."""
    return getattr(load(), name)(*args)
