"""cfft_exec_noalloc, the entry point that allocates nothing, against cfft_exec on the GPU.

Both go through the C ABI on the same plans.  A four-step executed in place needs a work buffer
for its middle: cfft_exec allocates it (once, lazily, per plan), cfft_exec_noalloc takes it from
the library's static 64 MiB device work area.  The kernels are the same, so the results must be
bitwise equal; two calls in a row (the second smaller, on a work area that still holds the first
call's middle) check that nothing depends on what the area held before.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

CFFT_ERR_UNSUPPORTED = -4


def _input(torch, n, batch, dt, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    rdt = torch.float32 if dt == np.complex64 else torch.float64
    return torch.complex(torch.randn(batch, n, generator=g, device="cuda", dtype=rdt),
                         torch.randn(batch, n, generator=g, device="cuda", dtype=rdt))


def _noalloc(lib, plan, src, dst, direction, normalize, stream):
    return lib.cfft_exec_noalloc(plan.h, ctypes.c_void_p(src.data_ptr()),
                                 ctypes.c_void_p(dst.data_ptr()), direction, normalize,
                                 ctypes.c_void_p(stream))


@pytest.mark.gpu
def test_noalloc_equals_exec_on_two_calls(gpu_ready):
    torch = gpu_ready
    import catabrafft as C
    from catabrafft import _lib as L

    lib = L.lib()
    stream = torch.cuda.current_stream().cuda_stream
    # first call: the middle fills the whole work area (2^20 x 8 x 8 B = 64 MiB); second call:
    # a smaller in-place four-step; then an F64 one and out-of-place / direct plans
    cases = [((1 << 20), 8, np.complex64, True), ((1 << 16), 4, np.complex64, True),
             ((1 << 17), 2, np.complex128, True), (2520, 64, np.complex64, False),
             ((1 << 18), 3, np.complex128, False), (257, 16, np.complex64, False)]
    for i, (n, batch, dt, inplace) in enumerate(cases):
        print(f"noalloc case n={n} batch={batch} {np.dtype(dt).name} inplace={inplace}",
              flush=True)
        plan = C.Plan(n, batch, dt)
        x = _input(torch, n, batch, dt, seed=100 + i)
        for direction, norm in ((-1, 0), (1, 1)):
            ref = x.clone()
            got = x.clone()
            if inplace:
                plan.exec_ptr(ref.data_ptr(), ref.data_ptr(), direction, bool(norm), stream)
                rc = _noalloc(lib, plan, got, got, direction, norm, stream)
            else:
                ref = torch.empty_like(x)
                got = torch.full_like(x, 7 + 7j)
                plan.exec_ptr(x.data_ptr(), ref.data_ptr(), direction, bool(norm), stream)
                rc = _noalloc(lib, plan, x, got, direction, norm, stream)
            assert rc == 0, lib.cfft_last_error()
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (n, batch, direction)
            rows = [0, batch - 1]
            xr = x[rows].cpu().numpy().astype(np.complex128)
            want = O.fft(xr, "d", direction, bool(norm), threads=8) if n > 4096 else \
                O.fft_l(xr, direction, bool(norm))
            tol = 1e-6 if dt == np.complex64 else 4 * np.finfo(float).eps * np.log2(n)
            assert rel_l2(got[rows].cpu().numpy(), want) < tol, (n, direction)


@pytest.mark.gpu
def test_noalloc_refuses_what_it_cannot_hold(gpu_ready):
    torch = gpu_ready
    import catabrafft as C
    from catabrafft import _lib as L

    lib = L.lib()
    stream = torch.cuda.current_stream().cuda_stream
    # an in-place middle of 128 MiB exceeds the 64 MiB work area
    plan = C.Plan(1 << 20, 16, np.complex64)
    x = _input(torch, 1 << 20, 16, np.complex64, seed=7)
    before = x.clone()
    assert _noalloc(lib, plan, x, x, -1, 0, stream) == CFFT_ERR_UNSUPPORTED
    assert b"work area" in lib.cfft_last_error()
    # a plan kind with work buffers of its own (large Rader convolution)
    q = C.Plan(65537, 2, np.complex64)
    y = _input(torch, 65537, 2, np.complex64, seed=8)
    z = torch.empty_like(y)
    assert _noalloc(lib, q, y, z, -1, 0, stream) == CFFT_ERR_UNSUPPORTED
    torch.cuda.synchronize()
    assert torch.equal(x, before)  # nothing ran
