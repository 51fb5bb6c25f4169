"""catabrafft -- CatabraFFT.jl's user API on the MI355X HIP library.

Same names, argument meaning and error behaviour as the reference's module (CatabraFFT.jl):

    fft(x)            CatabraFFT.jl:65-72    forward DFT, returns a new array, x untouched
    ifft(x)           CatabraFFT.jl:88-101   conj(fft(conj x)) / n
    bfft(x)           CatabraFFT.jl:103-115  conj(fft(conj x))
    plan_fft(x, flags) CatabraFFT.jl:149-171 -> Spell (spells.jl:18-37)
    p * x, p(x), mul(y, p, x)  CatabraFFT.jl:186-196, 224-226
    empty_cache()     CatabraFFT.jl:39-43
    NO_FLAG, MEASURE, ENCHANT  spells.jl:4-8

x is a complex64 / complex128 / complex32 numpy array (host, staged through the device) or a
complex torch tensor on a GPU (executed in place on the device, on torch's current stream).
complex32 is Julia's ComplexF16: numpy has no complex half type, so it is the structured dtype
`complex32` below ((re, im) float16 pairs), and torch's complex32 on the device; it is computed
in float32 and rounded to nearest even on output.  A 1-D x is one
transform; a 2-D x is a batch of rows (an extension: the reference has no batch API).

Differences from the reference, all fixes: a plan's flag is honoured by p * x (the reference
always runs the NO_FLAG function, CatabraFFT.jl:186-189 -> kernel.jl:97); mul(y, p, x) never
overwrites x (radix_exec.jl:63-68 does for plans of >= 3 passes); plan_bfft / plan_ifft work
(their reference stubs reference fields Spell does not have, CatabraFFT.jl:173-183).
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import (BACKWARD, C16, C32, C64, FORWARD, CatabraFFTArgumentError, CatabraFFTError,
                   check, lib)

NO_FLAG, MEASURE, ENCHANT = 0, 1, 2

__all__ = ["fft", "ifft", "bfft", "plan_fft", "plan_bfft", "plan_ifft", "mul", "empty_cache",
           "Spell", "Plan", "UnevenPlan", "Server", "Context", "context", "NO_FLAG", "MEASURE",
           "ENCHANT", "CatabraFFTError", "CatabraFFTArgumentError", "export_wisdom",
           "import_wisdom", "rfft", "irfft", "brfft", "RealPlan", "NDPlan", "fftn", "ifftn",
           "bfftn", "complex32", "to_complex32", "from_complex32"]


# ------------------------------------------------------------------------------ context
class Context:
    """A cfft_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().cfft_init(device, ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().cfft_finalize(self.h)
            self.h = None

    def empty_cache(self):
        check(lib().cfft_empty_cache(self.h))

    def synchronize(self):
        check(lib().cfft_synchronize(self.h))


_ctx_lock = threading.Lock()
_contexts: dict[int, Context] = {}


def context(device: int = 0) -> Context:
    with _ctx_lock:
        c = _contexts.get(device)
        if c is None:
            c = _contexts[device] = Context(device)
        return c


# ------------------------------------------------------------------------------ dtypes
# Julia's ComplexF16 on the host: (re, im) float16 pairs, 4 bytes, as the library stores them.
complex32 = np.dtype([("re", np.float16), ("im", np.float16)])


def to_complex32(z) -> np.ndarray:
    """A complex array rounded to ComplexF16 (each part to the nearest float16)."""
    z = np.asarray(z)
    out = np.empty(z.shape, dtype=complex32)
    out["re"] = z.real.astype(np.float16)
    out["im"] = z.imag.astype(np.float16)
    return out


def from_complex32(h) -> np.ndarray:
    """A ComplexF16 array widened (exactly) to complex64."""
    h = np.asarray(h)
    return h["re"].astype(np.float32) + 1j * h["im"].astype(np.float32)


_CODE_DTYPE = {C32: np.dtype(np.complex64), C64: np.dtype(np.complex128), C16: complex32}


def _dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.complex64:
        return C32
    if dt == np.complex128:
        return C64
    if dt == complex32:
        return C16
    raise TypeError("CatabraFFT transforms complex vectors (ComplexF32 / ComplexF64 / "
                    f"ComplexF16), got {dt}")


def _is_torch(x) -> bool:
    t = _lib.torch
    return t is not None and isinstance(x, t.Tensor)


def _torch_dtype_code(x) -> int:
    t = _lib.torch
    if x.dtype == t.complex64:
        return C32
    if x.dtype == t.complex128:
        return C64
    if x.dtype == t.complex32:
        return C16
    raise TypeError("CatabraFFT transforms complex vectors (ComplexF32 / ComplexF64 / "
                    f"ComplexF16), got {x.dtype}")


# ------------------------------------------------------------------------------ plans
class Plan:
    """A cfft_plan: `batch` transforms of length n, rows at idist / odist elements."""

    def __init__(self, n: int, batch: int = 1, dtype=np.complex64, flags: int = NO_FLAG,
                 idist: int | None = None, odist: int | None = None, device: int = 0):
        self.ctx = context(device)
        self.n, self.batch = int(n), int(batch)
        self.idist = int(idist if idist is not None else n)
        self.odist = int(odist if odist is not None else n)
        self.dtype_code = dtype if isinstance(dtype, int) else _dtype_code(dtype)
        self.dtype = _CODE_DTYPE[self.dtype_code]
        self.flags = int(flags)
        h = ctypes.c_void_p()
        check(lib().cfft_plan_1d(self.ctx.h, self.n, self.batch, self.idist, self.odist,
                                 self.dtype_code, self.flags, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        # at interpreter shutdown the module globals may already be gone: then the process exit
        # releases the plan
        h = getattr(self, "h", None)
        lib_mod = globals().get("_lib")
        if h is None or not h.value or lib_mod is None or getattr(lib_mod, "_lib", None) is None:
            return
        lib_mod._lib.cfft_plan_destroy(h)
        self.h = None

    # device pointers
    def exec_ptr(self, in_ptr: int, out_ptr: int, direction: int = FORWARD,
                 normalize: bool = False, stream: int | None = None):
        check(lib().cfft_exec(self.h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                              direction, int(normalize), ctypes.c_void_p(stream or 0)))

    def exec_torch(self, x, out=None, direction: int = FORWARD, normalize: bool = False):
        t = _lib.torch
        if not x.is_contiguous():
            x = x.contiguous()
        if out is None:
            if x.dtype == t.complex32:  # float16 pairs viewed as complex32 (no ComplexHalf alloc)
                out = t.empty(x.shape[:-1] + (2 * x.shape[-1],), dtype=t.float16,
                              device=x.device).view(t.complex32)
            else:
                out = t.empty_like(x)
        stream = t.cuda.current_stream(x.device).cuda_stream
        self.exec_ptr(x.data_ptr(), out.data_ptr(), direction, normalize, stream)
        return out

    def exec_host(self, x: np.ndarray, out: np.ndarray | None = None, direction: int = FORWARD,
                  normalize: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if out is None:
            out = np.empty_like(x)
        check(lib().cfft_exec_host(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                   out.ctypes.data_as(ctypes.c_void_p), direction,
                                   int(normalize)))
        return out

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        check(lib().cfft_plan_describe(self.h, buf, len(buf)))
        return buf.value.decode()

    def device_bytes(self) -> int:
        v = ctypes.c_int64()
        check(lib().cfft_plan_device_bytes(self.h, ctypes.byref(v)))
        return v.value


class RealPlan(Plan):
    """cfft_plan_1d_real: rows of n reals <-> half spectra of n/2+1 complex values."""

    def __init__(self, n: int, batch: int = 1, dtype=np.float32, flags: int = NO_FLAG,
                 rdist: int | None = None, cdist: int | None = None, device: int = 0):
        self.ctx = context(device)
        self.n, self.batch = int(n), int(batch)
        self.idist = int(rdist if rdist is not None else n)
        self.odist = int(cdist if cdist is not None else n // 2 + 1)
        dt = np.dtype(dtype)
        if dt not in (np.dtype(np.float32), np.dtype(np.float64)):
            raise TypeError(f"real transforms take float32 / float64 rows, got {dt}")
        self.dtype_code = C32 if dt == np.float32 else C64
        self.rdtype = dt.type
        self.dtype = np.complex64 if self.dtype_code == C32 else np.complex128
        self.flags = int(flags)
        h = ctypes.c_void_p()
        check(lib().cfft_plan_1d_real(self.ctx.h, self.n, self.batch, self.idist, self.odist,
                                      self.dtype_code, self.flags, ctypes.byref(h)))
        self.h = h

    def forward_host(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=self.rdtype)
        out = np.empty(x.shape[:-1] + (self.odist,), dtype=self.dtype)
        check(lib().cfft_exec_host(self.h, x.ctypes.data_as(ctypes.c_void_p),
                                   out.ctypes.data_as(ctypes.c_void_p), FORWARD, 0))
        return out[..., :self.n // 2 + 1]

    def backward_host(self, X: np.ndarray, normalize: bool) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=self.dtype)
        out = np.empty(X.shape[:-1] + (self.idist,), dtype=self.rdtype)
        check(lib().cfft_exec_host(self.h, X.ctypes.data_as(ctypes.c_void_p),
                                   out.ctypes.data_as(ctypes.c_void_p), BACKWARD,
                                   int(normalize)))
        return out[..., :self.n]

    def forward_torch(self, x):
        t = _lib.torch
        x = x.contiguous()
        out = t.empty(x.shape[:-1] + (self.odist,), dtype=t.complex64 if self.dtype_code == C32
                      else t.complex128, device=x.device)
        self.exec_ptr(x.data_ptr(), out.data_ptr(), FORWARD, False,
                      t.cuda.current_stream(x.device).cuda_stream)
        return out

    def backward_torch(self, X, normalize: bool):
        t = _lib.torch
        X = X.contiguous()
        out = t.empty(X.shape[:-1] + (self.idist,), dtype=t.float32 if self.dtype_code == C32
                      else t.float64, device=X.device)
        self.exec_ptr(X.data_ptr(), out.data_ptr(), BACKWARD, normalize,
                      t.cuda.current_stream(X.device).cuda_stream)
        return out


class NDPlan(Plan):
    """cfft_plan_nd: a multi-dimensional transform of one row-major array over `axes`."""

    def __init__(self, shape, axes=None, dtype=np.complex64, flags: int = NO_FLAG,
                 device: int = 0):
        self.ctx = context(device)
        self.shape = tuple(int(d) for d in shape)
        rank = len(self.shape)
        axes = range(rank) if axes is None else [a % rank for a in axes]
        mask = 0
        for a in axes:
            mask |= 1 << a
        self.dtype_code = dtype if isinstance(dtype, int) else _dtype_code(dtype)
        self.dtype = _CODE_DTYPE[self.dtype_code]
        self.n = int(np.prod([self.shape[a] for a in set(axes)])) if axes else 1
        self.batch, self.flags = 1, int(flags)
        dims = (ctypes.c_int64 * rank)(*self.shape)
        h = ctypes.c_void_p()
        check(lib().cfft_plan_nd(self.ctx.h, rank, dims, mask, self.dtype_code, self.flags,
                                 ctypes.byref(h)))
        self.h = h


class UnevenPlan:
    """cfft_plan_many_uneven: transform i has lengths[i] elements at offsets[i]."""

    def __init__(self, lengths, offsets, dtype=np.complex64, flags: int = NO_FLAG,
                 device: int = 0):
        self.ctx = context(device)
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int64)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.dtype_code = dtype if isinstance(dtype, int) else _dtype_code(dtype)
        P = ctypes.POINTER(ctypes.c_int64)
        h = ctypes.c_void_p()
        check(lib().cfft_plan_many_uneven(self.ctx.h, len(self.lengths),
                                          self.lengths.ctypes.data_as(P),
                                          self.offsets.ctypes.data_as(P), self.dtype_code,
                                          int(flags), ctypes.byref(h)))
        self.h = h

    __del__ = Plan.__del__
    exec_ptr = Plan.exec_ptr
    describe = Plan.describe
    device_bytes = Plan.device_bytes

    def exec_torch(self, x, out=None, direction: int = FORWARD, normalize: bool = False):
        t = _lib.torch
        if out is None:
            out = x.clone()
        stream = t.cuda.current_stream(x.device).cuda_stream
        self.exec_ptr(x.data_ptr(), out.data_ptr(), direction, normalize, stream)
        return out


class Server:
    """The continuous-batching serving loop (cfft_server_*).

    submit() takes a 1-D or 2-D (batch, n) complex array: a numpy array (host; staged by the
    server) or a GPU torch tensor (device-resident).  The output buffer is allocated here unless
    given; it is valid once wait(ticket) returns.
    """

    def __init__(self, hbm_cap_bytes: int = 200 * 10 ** 9, nstreams: int = 4, device: int = 0):
        self.ctx = context(device)
        h = ctypes.c_void_p()
        check(lib().cfft_server_create(self.ctx.h, int(hbm_cap_bytes), int(nstreams),
                                       ctypes.byref(h)))
        self.h = h
        self._live: dict[int, tuple] = {}

    def submit(self, x, out=None, direction: int = FORWARD, normalize: bool = False) -> int:
        n, batch = _shape(x)
        if _is_torch(x):
            code = _torch_dtype_code(x)
            if not x.is_contiguous():
                x = x.contiguous()
            if out is None:
                out = _lib.torch.empty_like(x)
            pin, pout, dev = x.data_ptr(), out.data_ptr(), 1
        else:
            x = np.ascontiguousarray(x)
            code = _dtype_code(x.dtype)
            if out is None:
                out = np.empty_like(x)
            pin, pout, dev = x.ctypes.data, out.ctypes.data, 0
        t = ctypes.c_int64()
        check(lib().cfft_server_submit(self.h, n, batch, code, direction, int(normalize),
                                       ctypes.c_void_p(pin), ctypes.c_void_p(pout), dev,
                                       ctypes.byref(t)))
        self._live[t.value] = (x, out)
        return t.value

    def poll(self, ticket: int) -> bool:
        rc = lib().cfft_server_poll(self.h, ticket)
        if rc == _lib.CFFT_ERR_NOTREADY:
            return False
        check(rc)
        return True

    def wait(self, ticket: int):
        check(lib().cfft_server_wait(self.h, ticket))
        return self._live.pop(ticket)[1]

    def stats(self) -> dict:
        v = [ctypes.c_double() for _ in range(4)]
        check(lib().cfft_server_stats(self.h, *[ctypes.byref(a) for a in v]))
        return {"p50_s": v[0].value, "p99_s": v[1].value, "transforms_per_s": v[2].value,
                "gflops": v[3].value}

    def reset_stats(self):
        check(lib().cfft_server_reset_stats(self.h))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            check(lib().cfft_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------ user API
_plan_lock = threading.Lock()
_plans: dict = {}


def _cached_plan(n: int, batch: int, code: int, flags: int, device: int) -> Plan:
    key = (n, batch, code, flags, device)
    with _plan_lock:
        p = _plans.get(key)
        if p is None:
            p = _plans[key] = Plan(n, batch, code, flags, device=device)
        return p


def _shape(x) -> tuple[int, int]:
    if x.ndim == 1:
        return x.shape[0], 1
    if x.ndim == 2:
        return x.shape[1], x.shape[0]
    raise ValueError("x must be a vector (one transform) or a 2-D array of rows")


def _run(x, direction: int, normalize: bool, flags: int = NO_FLAG):
    if _is_torch(x):
        code = _torch_dtype_code(x)
        if x.device.type != "cuda":
            raise TypeError("torch tensors must live on a GPU; pass numpy arrays for host data")
        n, batch = _shape(x)
        if n == 0 or batch == 0:
            return x.clone()
        p = _cached_plan(n, batch, code, flags, x.device.index or 0)
        return p.exec_torch(x, None, direction, normalize)
    x = np.asarray(x)
    code = _dtype_code(x.dtype)
    n, batch = _shape(x)
    if n == 0 or batch == 0:
        return x.copy()
    p = _cached_plan(n, batch, code, flags, 0)
    return p.exec_host(x, None, direction, normalize)


def fft(x):
    """Forward DFT y[k] = sum_j x[j] e^{-2 pi i jk/n} (CatabraFFT.jl:65-72)."""
    return _run(x, FORWARD, False)


def ifft(x):
    """Inverse DFT conj(fft(conj x)) / n (CatabraFFT.jl:88-101)."""
    return _run(x, BACKWARD, True)


def bfft(x):
    """Unnormalised backward DFT conj(fft(conj x)) (CatabraFFT.jl:103-115)."""
    return _run(x, BACKWARD, False)


class Spell:
    """A plan handle with the fields of spells.jl:18-28 (n, datatype, flag, inverse)."""

    def __init__(self, n: int, datatype, flag: int = NO_FLAG, direction: int = FORWARD,
                 normalize: bool = False, batch: int = 1):
        if flag not in (NO_FLAG, MEASURE, ENCHANT):
            raise ValueError(f"unknown flag {flag}")
        self.n, self.datatype, self.flag = int(n), np.dtype(datatype), int(flag)
        self.batch = int(batch)
        self.direction, self.normalize = direction, normalize
        self.inverse = None
        self._plan = _cached_plan(self.n, self.batch, _dtype_code(self.datatype), self.flag, 0)

    def __repr__(self):
        return f"Spell{{{self.datatype}}}({self.n}, flag={self.flag})"

    def __eq__(self, other):
        return (isinstance(other, Spell) and (self.n, self.datatype, self.flag, self.direction,
                                               self.normalize, self.batch) ==
                (other.n, other.datatype, other.flag, other.direction, other.normalize,
                 other.batch))

    def __hash__(self):
        return hash((self.n, self.datatype, self.flag, self.direction, self.normalize))

    def _check(self, x):
        n, batch = _shape(x)
        if n != self.n or batch != self.batch:
            raise ValueError(f"plan is for n={self.n} (batch {self.batch}), got n={n} "
                             f"(batch {batch})")

    def __mul__(self, x):
        """p * x (CatabraFFT.jl:192-196): a new output; this plan (and its flag) runs."""
        self._check(x)
        if _is_torch(x):
            return self._plan.exec_torch(x, None, self.direction, self.normalize)
        return self._plan.exec_host(np.asarray(x), None, self.direction, self.normalize)

    __rmul__ = None

    def __call__(self, x):
        """p(x) (CatabraFFT.jl:224-226) computes fft(x)."""
        return fft(x)

    def describe(self) -> str:
        return self._plan.describe()


def plan_fft(x, flags: int = NO_FLAG) -> Spell:
    """A forward plan for x's length and element type (CatabraFFT.jl:149-171)."""
    n, batch = _shape(x)
    dt = _CODE_DTYPE[_torch_dtype_code(x)] if _is_torch(x) else np.asarray(x).dtype
    return Spell(n, dt, flags, FORWARD, False, batch)


def plan_bfft(x, flags: int = NO_FLAG) -> Spell:
    n, batch = _shape(x)
    return Spell(n, np.asarray(x).dtype if not _is_torch(x) else plan_fft(x).datatype, flags,
                 BACKWARD, False, batch)


def plan_ifft(x, flags: int = NO_FLAG) -> Spell:
    n, batch = _shape(x)
    return Spell(n, np.asarray(x).dtype if not _is_torch(x) else plan_fft(x).datatype, flags,
                 BACKWARD, True, batch)


def mul(y, p: Spell, x):
    """mul!(y, p, x) (CatabraFFT.jl:186-189): writes p * x into y; x is left untouched."""
    p._check(x)
    if _is_torch(x):
        p._plan.exec_torch(x, y, p.direction, p.normalize)
        return y
    out = p._plan.exec_host(np.asarray(x), None, p.direction, p.normalize)
    y[...] = out
    return y


_rplans: dict = {}


def _real_plan(n: int, batch: int, rdt, device: int = 0) -> RealPlan:
    key = (n, batch, np.dtype(rdt).name, device)
    with _plan_lock:
        p = _rplans.get(key)
        if p is None:
            p = _rplans[key] = RealPlan(n, batch, rdt, device=device)
        return p


def rfft(x):
    """Half spectrum (n//2 + 1 values) of real rows (AbstractFFTs.rfft)."""
    if _is_torch(x):
        rdt = np.float32 if x.dtype == _lib.torch.float32 else np.float64
        n, batch = _shape(x)
        return _real_plan(n, batch, rdt, x.device.index or 0).forward_torch(x)
    x = np.asarray(x)
    if x.dtype not in (np.float32, np.float64):
        raise TypeError(f"rfft takes real float32 / float64 input, got {x.dtype}")
    n, batch = _shape(x)
    return _real_plan(n, batch, x.dtype).forward_host(x)


def brfft(X, n: int):
    """Unnormalised inverse of rfft to n reals (AbstractFFTs.brfft)."""
    return _c2r(X, n, False)


def irfft(X, n: int):
    """Inverse of rfft to n reals (AbstractFFTs.irfft)."""
    return _c2r(X, n, True)


def _c2r(X, n: int, normalize: bool):
    if _is_torch(X):
        code = _torch_dtype_code(X)
        m, batch = _shape(X)
        if m != n // 2 + 1:
            raise ValueError(f"a half spectrum of {m} values does not match n={n}")
        p = _real_plan(n, batch, np.float32 if code == C32 else np.float64, X.device.index or 0)
        return p.backward_torch(X, normalize)
    X = np.asarray(X)
    code = _dtype_code(X.dtype)
    m, batch = _shape(X)
    if m != n // 2 + 1:
        raise ValueError(f"a half spectrum of {m} values does not match n={n}")
    return _real_plan(n, batch, np.float32 if code == C32 else np.float64).backward_host(X, normalize)


_ndplans: dict = {}


def _nd(x, axes, direction: int, normalize: bool):
    rank = x.ndim
    ax = tuple(range(rank)) if axes is None else tuple(sorted(a % rank for a in axes))
    if _is_torch(x):
        code = _torch_dtype_code(x)
        key = (tuple(x.shape), ax, code, x.device.index or 0)
    else:
        x = np.asarray(x)
        code = _dtype_code(x.dtype)
        key = (x.shape, ax, code, 0)
    with _plan_lock:
        p = _ndplans.get(key)
        if p is None:
            p = _ndplans[key] = NDPlan(key[0], ax, code, device=key[3])
    if _is_torch(x):
        return p.exec_torch(x, None, direction, normalize)
    return p.exec_host(x, None, direction, normalize)


def fftn(x, axes=None):
    """Multi-dimensional forward DFT over `axes` (all by default) of a row-major array."""
    return _nd(x, axes, FORWARD, False)


def ifftn(x, axes=None):
    """Multi-dimensional inverse DFT (normalised by the transformed lengths)."""
    return _nd(x, axes, BACKWARD, True)


def bfftn(x, axes=None):
    """Multi-dimensional unnormalised backward DFT."""
    return _nd(x, axes, BACKWARD, False)


def export_wisdom(path: str, device: int = 0):
    """Writes the MEASURE / ENCHANT choices made so far to `path`."""
    check(lib().cfft_wisdom_export(context(device).h, str(path).encode()))


def import_wisdom(path: str, device: int = 0):
    """Loads choices written by export_wisdom; later MEASURE / ENCHANT plans reuse them."""
    check(lib().cfft_wisdom_import(context(device).h, str(path).encode()))


def empty_cache():
    """Drops cached plans and staging buffers (CatabraFFT.jl:39-43)."""
    with _plan_lock:
        _plans.clear()
        _rplans.clear()
        _ndplans.clear()
    with _ctx_lock:
        for c in _contexts.values():
            c.empty_cache()
