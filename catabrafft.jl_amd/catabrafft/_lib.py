"""Loader of lib/libcatabrafft.so, the C-ABI HIP library (include/catabrafft.h).

torch, when installed, is imported first: torch ships its own HIP runtime under the same
SONAME, and importing it first makes this library bind to that one runtime, so torch's streams
and device pointers are valid handles for the library.  The library is never replaced by a
CPU fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
import re

try:  # share one HIP runtime with torch (see module docstring)
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for the library itself
    torch = None

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libcatabrafft.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "catabrafft.h")

CFFT_OK = 0
CFFT_ERR_INVALID = -1
CFFT_ERR_NOMEM = -2
CFFT_ERR_HIP = -3
CFFT_ERR_UNSUPPORTED = -4
CFFT_ERR_ADMISSION = -5
CFFT_ERR_NOTREADY = -6

C32, C64, C16 = 0, 1, 2
FORWARD, BACKWARD = -1, 1

_lib = None


class CatabraFFTError(RuntimeError):
    """A non-zero status of the C ABI, with cfft_last_error()'s text."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class CatabraFFTArgumentError(CatabraFFTError, ValueError):
    """CFFT_ERR_INVALID: the arguments were rejected."""


def _set(L, name, argtypes, restype=ctypes.c_int):
    try:
        fn = getattr(L, name)
    except AttributeError:  # a build that predates this entry point
        return
    fn.argtypes = argtypes
    fn.restype = restype


def _sig(L):
    vp, i64, c_int, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    _set(L, "cfft_last_error", [], ctypes.c_char_p)
    _set(L, "cfft_version", [], ctypes.c_char_p)
    _set(L, "cfft_init", [c_int, P(vp)])
    _set(L, "cfft_finalize", [vp])
    _set(L, "cfft_plan_1d", [vp, i64, i64, i64, i64, c_int, c_int, P(vp)])
    _set(L, "cfft_plan_nd", [vp, c_int, P(i64), i64, c_int, c_int, P(vp)])
    _set(L, "cfft_plan_1d_real", [vp, i64, i64, i64, i64, c_int, c_int, P(vp)])
    _set(L, "cfft_plan_many_uneven", [vp, i64, P(i64), P(i64), c_int, c_int, P(vp)])
    _set(L, "cfft_exec", [vp, vp, vp, c_int, c_int, vp])
    _set(L, "cfft_exec_noalloc", [vp, vp, vp, c_int, c_int, vp])
    _set(L, "cfft_exec_host", [vp, vp, vp, c_int, c_int])
    _set(L, "cfft_plan_device_bytes", [vp, P(i64)])
    _set(L, "cfft_plan_describe", [vp, ctypes.c_char_p, i64])
    _set(L, "cfft_plan_destroy", [vp])
    _set(L, "cfft_empty_cache", [vp])
    _set(L, "cfft_wisdom_export", [vp, ctypes.c_char_p])
    _set(L, "cfft_wisdom_import", [vp, ctypes.c_char_p])
    _set(L, "cfft_malloc", [vp, i64, P(vp)])
    _set(L, "cfft_free", [vp, vp])
    _set(L, "cfft_memcpy_h2d", [vp, vp, vp, i64])
    _set(L, "cfft_memcpy_d2h", [vp, vp, vp, i64])
    _set(L, "cfft_synchronize", [vp])
    _set(L, "cfft_server_create", [vp, i64, c_int, P(vp)])
    _set(L, "cfft_server_submit", [vp, i64, i64, c_int, c_int, c_int, vp, vp, c_int, P(i64)])
    _set(L, "cfft_server_poll", [vp, i64])
    _set(L, "cfft_server_wait", [vp, i64])
    _set(L, "cfft_server_stats", [vp, P(dbl), P(dbl), P(dbl), P(dbl)])
    _set(L, "cfft_server_reset_stats", [vp])
    _set(L, "cfft_server_destroy", [vp])
    return L


def lib() -> ctypes.CDLL:
    """The loaded library (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `python catabrafft.jl_amd/build.py`"
                          " (or __graft_entry__.build())")
        _lib = _sig(ctypes.CDLL(LIB_PATH))
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        msg = lib().cfft_last_error().decode(errors="replace")
        cls = CatabraFFTArgumentError if rc == CFFT_ERR_INVALID else CatabraFFTError
        raise cls(rc, msg)
    return rc


def declared_functions() -> list[str]:
    """Names of the functions declared in include/catabrafft.h."""
    with open(HEADER) as f:
        text = f.read()
    return re.findall(r"CFFT_API\s+(?:const\s+char\s*\*|int)\s+(cfft_\w+)\s*\(", text)
