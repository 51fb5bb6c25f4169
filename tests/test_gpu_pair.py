"""ComplexF32 four-step passes with two adjacent lines per thread, on the GPU.

Each thread of kernels_pair_f32.hip computes two adjacent lines of a four-step pass in lockstep,
with 16-byte loads and stores, on the same tiles and LDS bytes as the float kernels.  The
arithmetic per line is the float kernels' (mixed_radix.jl:44-78 four-step, the layered Stockham
passes of radix_2_codelets.jl:402-447), so the checks are, through wisdom-forced plans and the C
ABI:

  * against the oracle (relative L2 <= 1e-6), forward and backward normalised, in place;
  * bit for bit against the same plan without pairs, aligned and with the buffers one element
    off 16-byte alignment (where the float kernels take the tiles);
  * bit for bit against the CPU emulation build of the same sources.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT, rel_l2, seeded_complex
from test_gpu_parity import oracle_ref

pytestmark = pytest.mark.gpu

EMU_LIB = os.path.join(ROOT, "tests", "emu", "_build", "libcatabrafft_emu.so")

PAIR = [  # n, batch, wisdom knobs, the description's pair marker
    (1 << 20, 4, "16 64 0 4 8 0 0 0 pr=3", "pair=AB"),
    (1 << 20, 3, "16 64 2048 8 8 0 0 0 ra=8,16,16 rb=4,8,16 pr=3", "pair=AB"),
    (1 << 16, 33, "16 64 0 8 16 0 0 0 pr=1", "pair=A"),
    (60 * 84, 70, "16 64 60 8 4 0 0 0 pr=3", "pair=AB"),     # ragged tiles
    (64 * 81, 40, "16 64 64 4 4 0 0 0 pr=3", "pair=B"),      # N2 odd: pass A stays float
    (10 ** 6, 2, "16 64 1000 4 8 0 0 0 pr=3", "pair=AB"),
    (1 << 20, 4, "8 64 2048 4 16 0 0 0 pr=3", "pair=AB"),    # E = 8: the ComplexF32 default
    (1 << 18, 9, "8 64 0 2 16 0 0 0 pr=3", "pair=AB"),       # E = 8, one line pair per A tile
    (48 * 96, 50, "8 64 48 4 16 0 0 0 pr=3", "pair=AB"),     # E = 8 lines of radix 3
]


def _plan(C, tmp_path, n, batch, line, tag):
    w = tmp_path / f"w_{tag}.txt"
    w.write_text(f"{n} {batch} {n} {n} 0 1 : {line}\n")
    C.import_wisdom(str(w))
    p = C.Plan(n, batch, np.complex64, flags=C.MEASURE)
    print(f"pair n={n} batch={batch} {tag}: {p.describe()}", flush=True)
    return p, w


@pytest.mark.parametrize("n,batch,fields,expect", PAIR)
def test_pair_fourstep_oracle_and_float_bits(gpu_ready, tmp_path, n, batch, fields, expect):
    torch = gpu_ready
    import catabrafft as C

    x = seeded_complex(n, batch, seed=n % 9973 + batch, dtype=np.complex64)
    s = torch.cuda.current_stream().cuda_stream
    outs = {}
    for tag, line in (("pair", fields), ("float", fields.rsplit(" pr=", 1)[0] + " pr=0")):
        p, _ = _plan(C, tmp_path, n, batch, line, tag)
        d = p.describe()
        assert "(wisdom)" in d and ((expect in d) == (tag == "pair")), d
        xd = torch.from_numpy(x).cuda()
        res = []
        for direction, norm in ((-1, False), (1, True)):
            yd = torch.empty_like(xd)
            p.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, norm, s)
            torch.cuda.synchronize()
            y = yd.cpu().numpy()
            if tag == "pair":
                assert rel_l2(y, oracle_ref(x, direction, norm)) <= 1e-6, direction
            res.append(y)
        zd = xd.clone()
        p.exec_ptr(zd.data_ptr(), zd.data_ptr(), -1, False, s)  # in place
        torch.cuda.synchronize()
        res.append(zd.cpu().numpy())
        # one element off 16-byte alignment (the float kernels take these calls)
        xb = torch.zeros(n * batch + 1, dtype=torch.complex64, device="cuda")
        yb = torch.zeros_like(xb)
        xb[1:] = xd.reshape(-1)
        p.exec_ptr(xb.data_ptr() + 8, yb.data_ptr() + 8, -1, False, s)
        torch.cuda.synchronize()
        res.append(yb[1:].reshape(batch, n).cpu().numpy())
        outs[tag] = res
    for a, b in zip(outs["pair"], outs["float"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(outs["pair"][0].view(np.uint32), outs["pair"][3].view(np.uint32))


@pytest.mark.parametrize("n,batch,fields,expect", PAIR[2:5] + PAIR[7:])
def test_pair_fourstep_bits_equal_cpu_emulation(gpu_ready, tmp_path, n, batch, fields, expect):
    torch = gpu_ready
    import catabrafft as C
    import catabrafft._lib as L

    if not os.path.exists(EMU_LIB):
        pytest.skip("the CPU emulation build is absent (python tests/emu/build_emu.py)")
    lib = L._sig(ctypes.CDLL(EMU_LIB))
    p, w = _plan(C, tmp_path, n, batch, fields, "pair")
    assert expect in p.describe()
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    assert lib.cfft_wisdom_import(ctx, str(w).encode()) == 0
    ep = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, n, n, 0, 1, ctypes.byref(ep)) == 0
    x = seeded_complex(n, batch, seed=5 * n, dtype=np.complex64)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    for direction, norm in ((-1, 0), (1, 1)):
        ye = np.zeros_like(x)
        assert lib.cfft_exec(ep, x.ctypes.data, ye.ctypes.data, direction, norm, None) == 0
        p.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, bool(norm), s)
        torch.cuda.synchronize()
        assert np.array_equal(ye.view(np.uint32), yd.cpu().numpy().view(np.uint32)), direction
    lib.cfft_plan_destroy(ep)
    lib.cfft_finalize(ctx)
