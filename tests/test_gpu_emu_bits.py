"""The GPU build and the CPU emulation build of the same kernel sources give the same bits.

Both builds perform the same IEEE operations in the same order (the library is compiled without
multiply-add contraction; the emulation with -ffp-contract=off), so a transform computed on the
MI355X equals, word for word, the one computed by tests/emu's CPU build, which the CPU suite
checks against the oracle.  Every plan kind is covered: direct rows (pipelined and not), the
four-step, Rader and Bluestein, fused and large (with the FP64 convolution of F32 input).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT

EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libcatabrafft_emu.so")

CASES = [  # (n, batch, dtype)
    (2520, 8, np.complex64), (2520, 8, np.complex128), (1024, 4, np.complex128),
    (720, 3, np.complex64), (64, 200, np.complex64), (1 << 16, 2, np.complex64),
    (1 << 17, 1, np.complex128), (17, 50, np.complex64), (257, 4, np.complex64),
    (4099, 2, np.complex64), (65537, 1, np.complex64), (17 * 19, 3, np.complex64),
    (12345, 2, np.complex128), (1, 3, np.complex64),
]


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU_LIB):
        pytest.skip("the CPU emulation build is absent (python tests/emu/build_emu.py)")
    import catabrafft._lib as L

    lib = L._sig(ctypes.CDLL(EMU_LIB))
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    yield lib, ctx
    lib.cfft_finalize(ctx)


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch,dt", CASES)
def test_gpu_bits_equal_cpu_emulation(gpu_ready, emu, n, batch, dt):
    import torch

    import catabrafft as C

    lib, ctx = emu
    rng = np.random.Generator(np.random.PCG64(n))
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(dt)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, n, n, 0 if dt == np.complex64 else 1, 0,
                            ctypes.byref(p)) == 0
    gp = C.Plan(n, batch, dt)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    word = np.uint32 if dt == np.complex64 else np.uint64
    for direction, norm in ((-1, 0), (1, 1)):
        ye = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, ye.ctypes.data, direction, norm, None) == 0
        gp.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, bool(norm), s)
        torch.cuda.synchronize()
        yg = yd.cpu().numpy()
        differing = int(np.count_nonzero(ye.view(word) != yg.view(word)))
        assert differing == 0, (gp.describe(), direction, differing)
    lib.cfft_plan_destroy(p)


@pytest.mark.gpu
@pytest.mark.parametrize("n,batch,dt", [(2520, 9, np.complex64), (1024, 5, np.complex128),
                                        (64, 300, np.complex64), (720, 3, np.complex64),
                                        (1 << 16, 3, np.complex64), (1 << 17, 1, np.complex128),
                                        (3 << 15, 2, np.complex64)])
def test_gpu_pipelined_bits_equal_cpu_emulation(gpu_ready, emu, tmp_path, n, batch, dt):
    """The persistent pipelined kernels (rows, four-step passes A and B; chosen through wisdom)
    give the bits of the emulation's plain kernels."""
    import torch

    import catabrafft as C

    lib, ctx = emu
    dcode = 0 if dt == np.complex64 else 1
    w = tmp_path / "wisdom.txt"
    w.write_text(f"{n} {batch} {n} {n} {dcode} 1 : 16 256 0 0 0 -1 1\n")
    C.import_wisdom(str(w))
    gp = C.Plan(n, batch, dt, flags=C.MEASURE)
    assert "pipe" in gp.describe(), gp.describe()
    rng = np.random.Generator(np.random.PCG64(n + 1))
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(dt)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, n, n, dcode, 0, ctypes.byref(p)) == 0
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    word = np.uint32 if dt == np.complex64 else np.uint64
    for direction, norm in ((-1, 0), (1, 1)):
        ye = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, ye.ctypes.data, direction, norm, None) == 0
        gp.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, bool(norm), s)
        torch.cuda.synchronize()
        assert np.array_equal(ye.view(word), yd.cpu().numpy().view(word)), direction
    lib.cfft_plan_destroy(p)
