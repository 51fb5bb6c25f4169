"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a C restatement of CatabraFFT.jl's FFT algorithm (see catabra_oracle.c for the
file:line map).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time against the product; the product never does.

Arrays are numpy complex arrays laid out as interleaved (re, im) pairs:
complex64 <-> float, complex128 <-> double, and pairs of np.longdouble <-> long double.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liborc.so")
_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        c_long, c_int, vp = ctypes.c_long, ctypes.c_int, ctypes.c_void_p
        L.orc_is_power_of.argtypes = [c_long, c_long]
        L.orc_isprime.argtypes = [c_long]
        L.orc_closest_factors.argtypes = [c_long, ctypes.POINTER(c_long), ctypes.POINTER(c_long)]
        L.orc_std_radix_plan.argtypes = [c_long, ctypes.POINTER(c_int), c_int]
        L.orc_powmod.argtypes = [c_long, c_long, c_long]
        L.orc_powmod.restype = c_long
        L.orc_primitive_root.argtypes = [c_long]
        L.orc_primitive_root.restype = c_long
        L.orc_dft_l.argtypes = [c_long, vp, vp, c_int]
        for s in "fdl":
            getattr(L, f"orc_plan_create_{s}").argtypes = [c_long]
            getattr(L, f"orc_plan_create_{s}").restype = vp
            getattr(L, f"orc_plan_destroy_{s}").argtypes = [vp]
            getattr(L, f"orc_exec_batch_{s}").argtypes = [vp, vp, vp, c_long, c_int, c_int, c_int]
            getattr(L, f"orc_stockham_pass_{s}").argtypes = [vp, vp, c_long, c_long, c_int]
            getattr(L, f"orc_codelet_{s}").argtypes = [vp, c_int]
        assert L.orc_sizeof_long_double() == np.dtype(np.longdouble).itemsize
        _lib = L
    return _lib


# ------------------------------------------------------------------ integer planner pieces
def is_power_of(n: int, p: int) -> bool:
    return bool(lib().orc_is_power_of(n, p))


def isprime(n: int) -> bool:
    return bool(lib().orc_isprime(n))


def closest_factors(n: int) -> tuple[int, int]:
    p, m = ctypes.c_long(), ctypes.c_long()
    lib().orc_closest_factors(n, ctypes.byref(p), ctypes.byref(m))
    return p.value, m.value


def std_radix_plan(n: int) -> list[int]:
    buf = (ctypes.c_int * 64)()
    k = lib().orc_std_radix_plan(n, buf, 64)
    return [buf[i] for i in range(k)]


def primitive_root(p: int) -> int:
    return lib().orc_primitive_root(p)


def powmod(b: int, e: int, m: int) -> int:
    return lib().orc_powmod(b, e, m)


def plan_tree(n: int):
    """The strategy tree kernel.jl:72-92 builds for n, as nested tuples (restated in Python
    from the same C pieces): ('trivial',1) | ('radix', n, radices) | ('mixed', n, p, m, Fp, Fm)
    | ('rader', n, g, F(n-1))."""
    if n == 1:
        return ("trivial", 1)
    if any(is_power_of(n, b) for b in (2, 3, 5, 7)):
        return ("radix", n, std_radix_plan(n))
    if isprime(n):
        return ("rader", n, primitive_root(n), plan_tree(n - 1))
    p, m = closest_factors(n)
    return ("mixed", n, p, m, plan_tree(p), plan_tree(m))


# ------------------------------------------------------------------ buffers
_SFX = {np.dtype(np.complex64): "f", np.dtype(np.complex128): "d"}


def _as_pairs_l(x: np.ndarray) -> np.ndarray:
    """complex128 / complex64 / (..., 2) longdouble -> contiguous (..., 2) longdouble."""
    if x.dtype == np.longdouble:
        return np.ascontiguousarray(x)
    x = np.asarray(x, dtype=np.complex128)
    out = np.empty(x.shape + (2,), dtype=np.longdouble)
    out[..., 0] = x.real
    out[..., 1] = x.imag
    return out


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ------------------------------------------------------------------ transforms
class Plan:
    """An oracle plan (the cached closure of kernel.jl:72-92) for one length and precision."""

    def __init__(self, n: int, prec: str):
        assert prec in ("f", "d", "l")
        self.n, self.prec = n, prec
        self.h = getattr(lib(), f"orc_plan_create_{prec}")(n)
        if not self.h:
            raise ValueError(f"oracle cannot plan n={n}")

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            getattr(_lib, f"orc_plan_destroy_{self.prec}")(h)
            self.h = None

    def exec(self, x: np.ndarray, direction: int = -1, normalize: bool = False,
             threads: int = 1) -> np.ndarray:
        """x: (batch, n) or (n,) array.  prec f/d: complex64/complex128; prec l: (..., n, 2)
        longdouble (or complex input, promoted)."""
        if self.prec == "l":
            xa = _as_pairs_l(x)
            single = xa.ndim == 2
            xa = xa.reshape(-1, self.n, 2)
        else:
            dt = np.complex64 if self.prec == "f" else np.complex128
            xa = np.ascontiguousarray(x, dtype=dt)
            single = xa.ndim == 1
            xa = xa.reshape(-1, self.n)
        y = np.empty_like(xa)
        getattr(lib(), f"orc_exec_batch_{self.prec}")(self.h, _ptr(xa), _ptr(y), xa.shape[0],
                                                      direction, int(normalize), threads)
        return y[0] if single else y


def fft(x, prec: str = "d", direction: int = -1, normalize: bool = False, threads: int = 1):
    x = np.asarray(x)
    n = x.shape[-2] if prec == "l" and x.dtype == np.longdouble else x.shape[-1]
    return Plan(n, prec).exec(x, direction, normalize, threads)


def fft_l(x: np.ndarray, direction: int = -1, normalize: bool = False) -> np.ndarray:
    """The restated algorithm in long double; returns complex128 rounded from long double."""
    y = fft(x, "l", direction, normalize)
    return (y[..., 0].astype(np.float64) + 1j * y[..., 1].astype(np.float64))


def dft_l_pairs(x, sign: int = -1) -> np.ndarray:
    """Naive long-double DFT of one vector; returns (n, 2) longdouble."""
    xa = _as_pairs_l(np.asarray(x))
    n = xa.shape[0]
    y = np.empty_like(xa)
    lib().orc_dft_l(n, _ptr(xa), _ptr(y), sign)
    return y


def dft_l(x, sign: int = -1) -> np.ndarray:
    y = dft_l_pairs(x, sign)
    return y[:, 0].astype(np.float64) + 1j * y[:, 1].astype(np.float64)


def stockham_pass(x: np.ndarray, s: int, R: int, prec: str = "d") -> np.ndarray:
    n = x.shape[0]
    if prec == "l":
        xa = _as_pairs_l(x)
    else:
        xa = np.ascontiguousarray(x, dtype=np.complex64 if prec == "f" else np.complex128)
    y = np.empty_like(xa)
    getattr(lib(), f"orc_stockham_pass_{prec}")(_ptr(xa), _ptr(y), n, s, R)
    return y


def codelet(v: np.ndarray, R: int, prec: str = "d") -> np.ndarray:
    if prec == "l":
        va = _as_pairs_l(v).copy()
    else:
        va = np.array(v, dtype=np.complex64 if prec == "f" else np.complex128)
    getattr(lib(), f"orc_codelet_{prec}")(_ptr(va), R)
    return va


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a - b||_2 / ||b||_2 computed in float64 (b is the reference)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
