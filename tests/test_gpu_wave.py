"""Lines whose last pass runs across the lanes of a wavefront, on the GPU.

A line whose last radix is a power of two with its digit inside one wavefront keeps the
penultimate pass's outputs in registers and computes the last pass with cross-lane exchanges
(DPP quad_perm / row_ror / row_shl / row_shr, v_permlane16_swap, v_permlane32_swap) instead of an
LDS write, a barrier and an LDS read (engine.h, wave_tail_pass).  The pass semantics are the
reference's layered Stockham pass (radix_2_codelets.jl:402-447) followed by its final pass.

Checked, through wisdom-forced plans and the C ABI:
  * against the oracle (ComplexF32 1e-6, ComplexF64 4 eps log2 n relative L2), forward,
    backward normalised, in place;
  * bitwise against the CPU emulation build of the same sources, which emulates the lane
    exchanges generically (so a wrong lane mapping of a DPP or permlane control shows up).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2, seeded_complex
from test_gpu_parity import oracle_ref, run_dev, tol

pytestmark = pytest.mark.gpu

EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libcatabrafft_emu.so")

# n, batch, dtype, wisdom fields (E target_threads n1 ca cb chunk pipe large ...)
WAVE = [
    (64, 300, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),      # lanes 1 apart, RL = 4
    (128, 40, np.complex128, "16 64 0 0 0 -1 0 0 wv=1"),     # RL = 8
    (256, 33, np.complex64, "16 256 0 0 0 -1 0 0 wv=1"),     # RL = 16, every DPP distance
    (512, 17, np.complex64, "8 64 0 0 0 -1 0 0 wv=1"),       # lanes 16 apart: permlane16
    (1024, 9, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),      # lanes 16, 32 apart
    (1024, 9, np.complex128, "8 64 0 0 0 -1 0 0 wv=1"),
    (1024, 5, np.complex64, "16 64 0 0 0 -1 0 0 ra=4,16,16 wv=1"),
    (1024, 5, np.complex128, "16 64 0 0 0 -1 0 0 ra=8,16,8 wv=1"),
    (257, 64, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),      # fused Rader, L = 256
    (253, 16, np.complex128, "16 64 0 0 0 -1 0 0 wv=1"),     # fused Bluestein, L = 512
]


def _forced(C, tmp_path, n, batch, dt, fields):
    code = 0 if dt == np.complex64 else 1
    w = tmp_path / "wisdom.txt"
    w.write_text(f"{n} {batch} {n} {n} {code} 1 : {fields}\n")
    C.import_wisdom(str(w))
    p = C.Plan(n, batch, dt, flags=C.MEASURE)
    d = p.describe()
    print(f"wave n={n} batch={batch} {np.dtype(dt).name}: {d}", flush=True)
    assert "(wisdom)" in d and " wave" in d, d
    return p, w


@pytest.mark.parametrize("n,batch,dt,fields", WAVE)
def test_wave_tail_matches_oracle(gpu_ready, tmp_path, n, batch, dt, fields):
    import catabrafft as C

    p, _ = _forced(C, tmp_path, n, batch, dt, fields)
    x = seeded_complex(n, batch, seed=n + batch, dtype=dt)
    for direction, norm in ((-1, False), (1, True)):
        y = run_dev(lambda xd: p.exec_torch(xd, direction=direction, normalize=norm), x)
        assert rel_l2(y, oracle_ref(x, direction, norm)) <= tol(n, dt), direction
    y = run_dev(lambda xd: p.exec_torch(xd, out=xd), x)  # in place
    assert rel_l2(y, oracle_ref(x)) <= tol(n, dt)


@pytest.mark.parametrize("n,batch,dt,fields", WAVE)
def test_wave_tail_bits_equal_cpu_emulation(gpu_ready, tmp_path, n, batch, dt, fields):
    import torch

    import catabrafft as C

    if not os.path.exists(EMU_LIB):
        pytest.skip("the CPU emulation build is absent (python tests/emu/build_emu.py)")
    import catabrafft._lib as L

    lib = L._sig(ctypes.CDLL(EMU_LIB))
    p, w = _forced(C, tmp_path, n, batch, dt, fields)
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    assert lib.cfft_wisdom_import(ctx, str(w).encode()) == 0
    ep = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, n, n, 0 if dt == np.complex64 else 1, 1,
                            ctypes.byref(ep)) == 0
    x = seeded_complex(n, batch, seed=7 * n, dtype=dt)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    word = np.uint32 if dt == np.complex64 else np.uint64
    for direction, norm in ((-1, 0), (1, 1)):
        ye = np.zeros_like(x)
        assert lib.cfft_exec(ep, x.ctypes.data, ye.ctypes.data, direction, norm, None) == 0
        p.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, bool(norm), s)
        torch.cuda.synchronize()
        differing = int(np.count_nonzero(ye.view(word) != yd.cpu().numpy().view(word)))
        assert differing == 0, (p.describe(), direction, differing)
    lib.cfft_plan_destroy(ep)
    lib.cfft_finalize(ctx)
