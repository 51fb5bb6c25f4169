"""ComplexF16 data (CFFT_C16): half-precision storage, float arithmetic.

The reference round-trips ComplexF16 through fft / ifft at n in [89, 121, 2^10, 2^11, 3^7, 5^6,
10^4] and asks for x ≈ ifft(fft(x)) (runtests.jl:39-46; Julia's isapprox default for Float16 is
rtol = sqrt(eps(Float16)), compared in norm).  Here, in addition:

  * the library's C16 result is, bit for bit, its C32 result on the widened input rounded to the
    nearest float16 (the half kernels run the float kernels' arithmetic; only the global loads
    and stores convert) -- this ties every C16 plan kind to the C32 paths the oracle checks;
  * the result is within half-precision rounding of the oracle's float64 DFT of the same input;
  * the emulation's software binary16 rounding equals numpy's float16 conversion on every half
    value, on every midpoint between neighbours and on random floats, so the GPU test that
    compares device bits with emulation bits pins the device's conversion too.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, rel_l2, seeded_complex

EMU = os.path.join(ROOT, "tests", "emu")
EMU_LIB = os.path.join(EMU, "_build", "libcatabrafft_emu.so")
C32, C64, C16 = 0, 1, 2
SELFTEST_N = [89, 121, 1 << 10, 1 << 11, 3 ** 7, 5 ** 6, 10 ** 4]  # runtests.jl:42
RTOL_F16 = float(np.sqrt(np.finfo(np.float16).eps))               # Julia isapprox, Float16


def to_h(z):
    """complex -> (re, im) float16 pairs, shape (..., n, 2)."""
    return np.stack([z.real.astype(np.float16), z.imag.astype(np.float16)], axis=-1)


def from_h(h):
    return h[..., 0].astype(np.float32) + 1j * h[..., 1].astype(np.float32)


def roundtrip_ok(x, y) -> bool:  # Julia's x ≈ y for arrays: norm(x - y) <= rtol * max(norms)
    x = np.asarray(x, np.complex128).ravel()
    y = np.asarray(y, np.complex128).ravel()
    return np.linalg.norm(x - y) <= RTOL_F16 * max(np.linalg.norm(x), np.linalg.norm(y))


# ------------------------------------------------------------------ software binary16 rounding
HALF_PROG = r"""
#include <cstdio>
#include <vector>
#include "hip/hip_runtime.h"
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    std::vector<float> in;
    float v;
    while (std::fread(&v, 4, 1, f) == 1) in.push_back(v);
    std::fclose(f);
    FILE* o = std::fopen(argv[2], "wb");
    for (float x : in) { uint16_t b = emu_half(x).bits; std::fwrite(&b, 2, 1, o); }
    std::fclose(o);
    FILE* w = std::fopen(argv[3], "wb");
    for (uint32_t h = 0; h < 65536; ++h) {
        emu_half e; e.bits = (uint16_t)h; float x = (float)e; std::fwrite(&x, 4, 1, w);
    }
    std::fclose(w);
    return 0;
}
"""


def test_software_half_matches_numpy(tmp_path):
    src = tmp_path / "half.cpp"
    src.write_text(HALF_PROG)
    exe = tmp_path / "half"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(EMU, "include"), str(src),
                    "-o", str(exe)], check=True)
    allh = np.arange(65536, dtype=np.uint16).view(np.float16)
    finite = allh[np.isfinite(allh)].astype(np.float32)
    pos = np.unique(np.abs(finite))
    mids = ((pos[:-1].astype(np.float64) + pos[1:]) / 2).astype(np.float32)  # exact ties
    near = np.concatenate([np.nextafter(mids, np.float32(0)), np.nextafter(mids, np.float32(1e9))])
    rng = np.random.Generator(np.random.PCG64(16))
    rnd = rng.integers(0, 2 ** 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    rnd = rnd[~np.isnan(rnd)]
    special = np.array([0.0, -0.0, np.inf, -np.inf, 65504, 65519.99, 65520, 1e-8, 2 ** -25,
                        2 ** -24, 3 * 2 ** -26, 6e-5, 6.1e-5], np.float32)
    xs = np.concatenate([finite, -finite, mids, -mids, near, -near, rnd, special])
    (tmp_path / "in.bin").write_bytes(xs.astype(np.float32).tobytes())
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "h.bin"),
                    str(tmp_path / "w.bin")], check=True)
    got = np.frombuffer((tmp_path / "h.bin").read_bytes(), np.uint16)
    with np.errstate(over="ignore"):
        want = xs.astype(np.float16).view(np.uint16)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(xs[i]), hex(got[i]), hex(want[i])) for i in bad[:5]]
    widened = np.frombuffer((tmp_path / "w.bin").read_bytes(), np.float32)
    ok = ~np.isnan(allh)
    assert np.array_equal(widened[ok].view(np.uint32), allh[ok].astype(np.float32).view(np.uint32))


# ------------------------------------------------------------------ emulation build (CPU)
@pytest.fixture(scope="module")
def E():
    subprocess.run([sys.executable, os.path.join(EMU, "build_emu.py")], check=True,
                   capture_output=True)
    import catabrafft._lib as L

    lib = L._sig(ctypes.CDLL(EMU_LIB))
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    return lib, ctx


def _plan1d(lib, ctx, n, batch, code):
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, n, n, code, 0, ctypes.byref(p)) == 0, \
        lib.cfft_last_error()
    return p


def _desc(lib, p):
    buf = ctypes.create_string_buffer(1024)
    lib.cfft_plan_describe(p, buf, len(buf))
    return buf.value.decode()


EMU_CASES = [  # (n, batch, strategy)
    (1, 3, "trivial"), (16, 5, "direct"), (2520, 2, "direct"), (1000, 3, "direct"),
    (1 << 16, 1, "fourstep"), (17, 7, "rader-fused"), (257, 2, "rader-fused"),
    (17 * 19, 2, "bluestein-fused"), (65537, 1, "rader-large"), (3 * 8191, 1, "bluestein-large"),
]


@pytest.mark.parametrize("n,batch,kind", EMU_CASES)
def test_emulated_f16_equals_rounded_f32(E, n, batch, kind):
    lib, ctx = E
    xh = to_h(seeded_complex(n, batch=batch, seed=n + 16))
    x32 = from_h(xh).astype(np.complex64)
    p16 = _plan1d(lib, ctx, n, batch, C16)
    p32 = _plan1d(lib, ctx, n, batch, C32)
    d = _desc(lib, p16)
    assert kind in d and "io=f16" in d, d
    for direction, norm in ((-1, 0), (1, 1), (1, 0)):
        y16 = np.zeros_like(xh)
        y32 = np.zeros_like(x32)
        assert lib.cfft_exec(p16, xh.ctypes.data, y16.ctypes.data, direction, norm, None) == 0
        assert lib.cfft_exec(p32, x32.ctypes.data, y32.ctypes.data, direction, norm, None) == 0
        assert np.array_equal(y16.view(np.uint16), to_h(y32).view(np.uint16)), (direction, norm)
        ref = O.fft(x32.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(from_h(y16), ref) < 2e-3, (direction, norm)
    for p in (p16, p32):
        lib.cfft_plan_destroy(p)


@pytest.mark.parametrize("n", SELFTEST_N[:5])
def test_emulated_f16_selftest_roundtrip(E, n):
    lib, ctx = E
    xh = to_h(seeded_complex(n, seed=n))
    p = _plan1d(lib, ctx, n, 1, C16)
    y = np.zeros_like(xh)
    z = np.zeros_like(xh)
    assert lib.cfft_exec(p, xh.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    assert lib.cfft_exec(p, y.ctypes.data, z.ctypes.data, 1, 1, None) == 0
    assert roundtrip_ok(from_h(xh), from_h(z))
    lib.cfft_plan_destroy(p)


def test_emulated_f16_in_place_and_nd(E):
    lib, ctx = E
    n, batch = 1 << 16, 1
    xh = to_h(seeded_complex(n, batch=batch, seed=3))
    p16 = _plan1d(lib, ctx, n, batch, C16)
    y = np.zeros_like(xh)
    assert lib.cfft_exec(p16, xh.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    xi = xh.copy()
    assert lib.cfft_exec(p16, xi.ctypes.data, xi.ctypes.data, -1, 0, None) == 0
    assert np.array_equal(xi.view(np.uint16), y.view(np.uint16))
    # N-D: 12 x 40 over both axes, against the C32 plan rounded
    dims = (ctypes.c_int64 * 2)(12, 40)
    ph, pf = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.cfft_plan_nd(ctx, 2, dims, 3, C16, 0, ctypes.byref(ph)) == 0, lib.cfft_last_error()
    assert lib.cfft_plan_nd(ctx, 2, dims, 3, C32, 0, ctypes.byref(pf)) == 0
    assert "io=f16" in _desc(lib, ph)
    xh = to_h(seeded_complex(40, batch=12, seed=5))
    x32 = from_h(xh).astype(np.complex64)
    yh, yf = np.zeros_like(xh), np.zeros_like(x32)
    assert lib.cfft_exec(ph, xh.ctypes.data, yh.ctypes.data, -1, 0, None) == 0
    assert lib.cfft_exec(pf, x32.ctypes.data, yf.ctypes.data, -1, 0, None) == 0
    assert np.array_equal(yh.view(np.uint16), to_h(yf).view(np.uint16))
    ref = O.fft(x32.astype(np.complex128), "d", -1, False)          # axis 1
    ref = O.fft(np.ascontiguousarray(ref.T), "d", -1, False).T       # axis 0
    assert rel_l2(from_h(yh), ref) < 2e-3
    for p in (p16, ph, pf):
        lib.cfft_plan_destroy(p)


def test_f16_rejected_where_unsupported(E):
    lib, ctx = E
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d_real(ctx, 64, 1, 64, 33, C16, 0, ctypes.byref(p)) != 0
    lengths = (ctypes.c_int64 * 2)(8, 9)
    offsets = (ctypes.c_int64 * 2)(0, 8)
    assert lib.cfft_plan_many_uneven(ctx, 2, lengths, offsets, C16, 0, ctypes.byref(p)) != 0
    assert lib.cfft_plan_1d(ctx, 64, 1, 64, 64, 3, 0, ctypes.byref(p)) != 0


def test_python_complex32_helpers():
    import catabrafft as C

    z = seeded_complex(33, seed=1)
    h = C.to_complex32(z)
    assert h.dtype == C.complex32 and h.itemsize == 4
    assert np.array_equal(C.from_complex32(h), from_h(to_h(z)))
    assert C._dtype_code(C.complex32) == C16


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("n,batch", [(2520, 64), (1 << 20, 2), (17, 300), (65537, 2),
                                     (3 * 8191, 2), (1000, 17), (1, 5)])
def test_gpu_f16_equals_rounded_f32(gpu_ready, n, batch):
    import torch

    import catabrafft as C

    x = seeded_complex(n, batch=batch, seed=n)
    xh = to_h(x)
    x32 = from_h(xh).astype(np.complex64)
    p16 = C.Plan(n, batch, C16)
    p32 = C.Plan(n, batch, np.complex64)
    assert "io=f16" in p16.describe()
    xhd = torch.from_numpy(xh).cuda()
    x32d = torch.from_numpy(x32).cuda()
    s = torch.cuda.current_stream().cuda_stream
    for direction, norm in ((-1, 0), (1, 1)):
        yhd = torch.empty_like(xhd)
        y32d = torch.empty_like(x32d)
        p16.exec_ptr(xhd.data_ptr(), yhd.data_ptr(), direction, norm, s)
        p32.exec_ptr(x32d.data_ptr(), y32d.data_ptr(), direction, norm, s)
        torch.cuda.synchronize()
        yh, y32 = yhd.cpu().numpy(), y32d.cpu().numpy()
        assert np.array_equal(yh.view(np.uint16), to_h(y32).view(np.uint16)), (direction, norm)
        if n <= 1 << 16:
            ref = O.fft(x32.astype(np.complex128), "d", direction, bool(norm), threads=8)
            assert rel_l2(from_h(yh), ref) < 2e-3


@pytest.mark.gpu
@pytest.mark.parametrize("n", SELFTEST_N)
def test_gpu_f16_selftest_roundtrip_torch_complex32(gpu_ready, n):
    import torch

    import catabrafft as C

    xh = to_h(seeded_complex(n, seed=n))
    xc = torch.from_numpy(xh).cuda().reshape(-1).view(torch.complex32)
    assert xc.shape == (n,)
    y = C.fft(xc)
    z = C.ifft(y)
    assert y.dtype == torch.complex32 and z.dtype == torch.complex32
    zh = z.view(torch.float16).reshape(n, 2).cpu().numpy()
    assert roundtrip_ok(from_h(xh), from_h(zh))
    ref = O.fft(from_h(xh).astype(np.complex128), "d", -1, False)
    yh = y.view(torch.float16).reshape(n, 2).cpu().numpy()
    assert rel_l2(from_h(yh), ref) < 2e-3


@pytest.mark.gpu
def test_gpu_f16_bits_equal_cpu_emulation(gpu_ready):
    if not os.path.exists(EMU_LIB):
        pytest.skip("the CPU emulation build is absent (python tests/emu/build_emu.py)")
    import torch

    import catabrafft as C
    import catabrafft._lib as L

    lib = L._sig(ctypes.CDLL(EMU_LIB))
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    for n, batch in ((2520, 4), (1 << 16, 1), (17, 9), (17 * 19, 2)):
        xh = to_h(seeded_complex(n, batch=batch, seed=7 * n))
        pe = _plan1d(lib, ctx, n, batch, C16)
        ye = np.zeros_like(xh)
        assert lib.cfft_exec(pe, xh.ctypes.data, ye.ctypes.data, -1, 0, None) == 0
        lib.cfft_plan_destroy(pe)
        pg = C.Plan(n, batch, C16)
        xd = torch.from_numpy(xh).cuda()
        yd = torch.empty_like(xd)
        pg.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, 0, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert np.array_equal(yd.cpu().numpy().view(np.uint16), ye.view(np.uint16)), n
    lib.cfft_finalize(ctx)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [89, 1 << 10, 10 ** 4])
def test_gpu_f16_host_arrays(gpu_ready, n):
    """numpy ComplexF16 arrays (catabrafft.complex32) through fft / ifft: the host-pointer path
    (cfft_exec_host) the Julia shim uses; the reference's round trip and the oracle."""
    import catabrafft as C

    x = C.to_complex32(seeded_complex(n, seed=2 * n))
    y = C.fft(x)
    z = C.ifft(y)
    assert y.dtype == C.complex32 and z.dtype == C.complex32
    xw = C.from_complex32(x)
    assert roundtrip_ok(xw, C.from_complex32(z))
    ref = O.fft(xw.astype(np.complex128), "d", -1, False)
    assert rel_l2(C.from_complex32(y), ref) < 2e-3
    p = C.plan_fft(x)
    assert np.array_equal((p * x).view(np.uint16), y.view(np.uint16))


@pytest.mark.parametrize("n,batch,kind", [(2520, 3, "direct"), (1 << 16, 2, "fourstep"),
                                          (17, 5, "rader-fused")])
def test_emulated_f16_strided_rows_and_in_place(E, n, batch, kind):
    """Rows at idist / odist > n, and in == out, against the C32 plan on the same layout."""
    lib, ctx = E
    idist, odist = n + 3, n + 5
    x = seeded_complex(idist, batch=batch, seed=9)
    xh = to_h(x)
    x32 = from_h(xh).astype(np.complex64)
    p16, p32 = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, idist, odist, C16, 0, ctypes.byref(p16)) == 0
    assert lib.cfft_plan_1d(ctx, n, batch, idist, odist, C32, 0, ctypes.byref(p32)) == 0
    assert kind in _desc(lib, p16)
    yh = np.zeros((batch, odist, 2), np.float16)
    y32 = np.zeros((batch, odist), np.complex64)
    assert lib.cfft_exec(p16, xh.ctypes.data, yh.ctypes.data, -1, 0, None) == 0
    assert lib.cfft_exec(p32, x32.ctypes.data, y32.ctypes.data, -1, 0, None) == 0
    assert np.array_equal(yh[:, :n].view(np.uint16), to_h(y32[:, :n]).view(np.uint16))
    assert not yh[:, n:].any()  # the gaps between output rows are not written
    q16, q32 = ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, batch, idist, idist, C16, 0, ctypes.byref(q16)) == 0
    assert lib.cfft_plan_1d(ctx, n, batch, idist, idist, C32, 0, ctypes.byref(q32)) == 0
    zh, z32 = xh.copy(), x32.copy()
    assert lib.cfft_exec(q16, zh.ctypes.data, zh.ctypes.data, 1, 1, None) == 0
    assert lib.cfft_exec(q32, z32.ctypes.data, z32.ctypes.data, 1, 1, None) == 0
    assert np.array_equal(zh[:, :n].view(np.uint16), to_h(z32[:, :n]).view(np.uint16))
    for p in (p16, p32, q16, q32):
        lib.cfft_plan_destroy(p)


@pytest.mark.timeout(120)
def test_emulated_f16_sixstep_through_wisdom(E, tmp_path):
    """A ComplexF16 plan forced onto the six-step path by wisdom: its staged float transform
    takes the plan's scratch while the half-precision staging buffers are held.  The result is
    bit-equal to the C32 six-step result rounded to float16 (and it returns at all)."""
    lib, ctx = E
    n, batch = 1 << 14, 2
    path = tmp_path / "w.txt"
    path.write_text(f"{n} {batch} {n} {n} {C16} 1 : 0 64 64 0 0 -1 -1 1\n"
                    f"{n} {batch} {n} {n} {C32} 1 : 0 64 64 0 0 -1 -1 1\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
    ps = {}
    for code in (C16, C32):
        p = ctypes.c_void_p()
        assert lib.cfft_plan_1d(c2, n, batch, n, n, code, 1, ctypes.byref(p)) == 0, \
            lib.cfft_last_error()
        assert "sixstep" in _desc(lib, p), _desc(lib, p)
        ps[code] = p
    xh = to_h(seeded_complex(n, batch=batch, seed=61))
    x32 = from_h(xh).astype(np.complex64)
    for direction, norm in ((-1, 0), (1, 1)):
        y16 = np.zeros_like(xh)
        y32 = np.zeros_like(x32)
        assert lib.cfft_exec(ps[C16], xh.ctypes.data, y16.ctypes.data, direction, norm, None) == 0
        assert lib.cfft_exec(ps[C32], x32.ctypes.data, y32.ctypes.data, direction, norm, None) == 0
        assert np.array_equal(y16.view(np.uint16), to_h(y32).view(np.uint16)), (direction, norm)
        ref = O.fft(x32.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(from_h(y16), ref) < 2e-3
    for p in ps.values():
        lib.cfft_plan_destroy(p)
    lib.cfft_finalize(c2)
