"""The library's own sources (planner, C ABI, HIP kernel bodies) built for the CPU emulation
runtime in tests/emu and checked against the oracle on small cases.

This runs the same kernel code paths as the GPU build (pass sequences, LDS layouts, four-step
tiling, Rader / Bluestein, work lists, hipGraph capture and the server's rounds) without a GPU, so
indexing and planning mistakes are caught on the CPU.  It says nothing about GPU timing or about
compiler behaviour on gfx950; the GPU suite covers that.
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


@pytest.fixture(scope="module")
def E():
    subprocess.run([sys.executable, os.path.join(EMU, "build_emu.py")], check=True,
                   capture_output=True)
    import catabrafft._lib as L

    lib = L._sig(ctypes.CDLL(os.path.join(EMU, "_build", "libcatabrafft_emu.so")))
    ctx = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
    return lib, ctx


def _plan(lib, ctx, n, batch, dt, idist=None, odist=None, flags=0):
    p = ctypes.c_void_p()
    rc = lib.cfft_plan_1d(ctx, n, batch, idist or n, odist or n, 0 if dt == np.complex64 else 1,
                          flags, ctypes.byref(p))
    assert rc == 0, lib.cfft_last_error()
    return p


def _desc(lib, p):
    buf = ctypes.create_string_buffer(1024)
    lib.cfft_plan_describe(p, buf, len(buf))
    return buf.value.decode()


def _tol(n, dt):
    return 1e-6 if dt == np.complex64 else 4 * np.finfo(float).eps * max(1, np.log2(n))


CASES = [  # (n, batch, dtype, expected strategy)
    (1, 3, np.complex64, "trivial"), (2, 4, np.complex128, "direct"), (16, 5, np.complex64, "direct"),
    (60, 5, np.complex64, "direct"), (1000, 3, np.complex64, "direct"),
    (1024, 2, np.complex128, "direct"), (2520, 2, np.complex64, "direct"),
    (2520, 2, np.complex128, "direct"), (11 * 13 * 8, 2, np.complex64, "direct"),
    (17, 7, np.complex64, "rader-fused"), (97, 3, np.complex128, "rader-fused"),
    (257, 2, np.complex64, "rader-fused"), (4099, 1, np.complex64, "rader-fused"),
    (17 * 19, 2, np.complex64, "bluestein-fused"), (1 << 16, 1, np.complex64, "fourstep"),
    (1 << 17, 1, np.complex128, "fourstep"), (65537, 1, np.complex64, "rader-large"),
    (3 * 8191, 1, np.complex64, "bluestein-large"),
]


@pytest.mark.parametrize("n,batch,dt,kind", CASES)
def test_emulated_kernels_match_oracle(E, n, batch, dt, kind):
    lib, ctx = E
    x = seeded_complex(n, batch=batch, seed=n, dtype=dt)
    p = _plan(lib, ctx, n, batch, dt)
    assert kind in _desc(lib, p)
    for direction, norm in ((-1, 0), (1, 1), (1, 0)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(y, ref) < _tol(n, dt), (direction, norm)
    lib.cfft_plan_destroy(p)


@pytest.mark.parametrize("n", [2520, 4099, 1 << 16, 65537])
def test_emulated_in_place_and_strided(E, n):
    lib, ctx = E
    batch, idist, odist = 2, n + 3, n + 7
    x = seeded_complex(idist * batch, seed=3, dtype=np.complex64)
    y = np.zeros(odist * batch, dtype=np.complex64)
    p = _plan(lib, ctx, n, batch, np.complex64, idist, odist)
    assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    for b in range(batch):
        ref = O.fft(x[b * idist:b * idist + n].astype(np.complex128), "d")
        assert rel_l2(y[b * odist:b * odist + n], ref) < 1e-6
        assert np.all(y[b * odist + n:(b + 1) * odist] == 0)
    z = seeded_complex(n, batch=2, seed=4, dtype=np.complex64)
    z0 = z.copy()
    q = _plan(lib, ctx, n, 2, np.complex64)
    assert lib.cfft_exec(q, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    assert rel_l2(z, O.fft(z0.astype(np.complex128), "d")) < 1e-6


def test_emulated_uneven_plan_and_alone_equality(E):
    lib, ctx = E
    lengths = [64, 1000, 4099, 2520, 1 << 16, 12345, 7, 1, 2, 17, 17, 60, 60, 60, 3 * 8191, 65537]
    offs = np.cumsum([0] + [n + 3 for n in lengths[:-1]]).astype(np.int64)
    span = int(offs[-1] + lengths[-1])
    le = np.array(lengths, dtype=np.int64)
    P = ctypes.POINTER(ctypes.c_int64)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_many_uneven(ctx, len(lengths), le.ctypes.data_as(P),
                                     offs.ctypes.data_as(P), 0, 0, ctypes.byref(p)) == 0
    x = seeded_complex(span, seed=1, dtype=np.complex64)
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        for n, o in zip(lengths, offs):
            ref = O.fft(x[o:o + n].astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(y[o:o + n], ref) < 1e-6, n
            # alone (single-length plan) and in the mix give identical bits
            q = _plan(lib, ctx, n, 1, np.complex64)
            alone = np.zeros(n, dtype=np.complex64)
            assert lib.cfft_exec(q, x[o:].ctypes.data, alone.ctypes.data, direction, norm, None) == 0
            assert np.array_equal(alone, y[o:o + n]), n
        assert np.all(y[offs[0] + lengths[0]:offs[1]] == 0)


def test_emulated_server_rounds(E):
    lib, ctx = E
    srv = ctypes.c_void_p()
    assert lib.cfft_server_create(ctx, 1 << 30, 2, ctypes.byref(srv)) == 0
    reqs = []
    for i, (n, b) in enumerate([(60, 3), (1024, 2), (17, 4), (4099, 1), (1 << 16, 1), (60, 2),
                                 (3 * 8191, 1)]):
        dt = np.complex128 if i == 1 else np.complex64
        x = seeded_complex(n, batch=b, seed=i, dtype=dt)
        y = np.zeros_like(x)
        t = ctypes.c_int64()
        direction = -1 if i % 2 == 0 else 1
        assert lib.cfft_server_submit(srv, n, b, 0 if dt == np.complex64 else 1, direction,
                                      int(direction > 0), x.ctypes.data, y.ctypes.data, i % 2,
                                      ctypes.byref(t)) == 0
        reqs.append((t.value, x, y, direction))
    for t, x, y, direction in reqs:
        assert lib.cfft_server_wait(srv, t) == 0
        ref = O.fft(x.astype(np.complex128), "d", direction, direction > 0)
        assert rel_l2(y, ref) < _tol(x.shape[1], x.dtype)
    v = [ctypes.c_double() for _ in range(4)]
    assert lib.cfft_server_stats(srv, *[ctypes.byref(a) for a in v]) == 0
    assert v[0].value > 0 and v[1].value >= v[0].value and v[2].value > 0
    # a request above the cap is refused
    big = seeded_complex(1 << 16, batch=8, seed=9, dtype=np.complex64)
    out = np.zeros_like(big)
    srv2 = ctypes.c_void_p()
    assert lib.cfft_server_create(ctx, big.nbytes, 1, ctypes.byref(srv2)) == 0
    t = ctypes.c_int64()
    assert lib.cfft_server_submit(srv2, 1 << 16, 8, 0, -1, 0, big.ctypes.data, out.ctypes.data, 0,
                                  ctypes.byref(t)) == 0
    assert lib.cfft_server_wait(srv2, t.value) == -5
    assert lib.cfft_server_destroy(srv2) == 0
    assert lib.cfft_server_destroy(srv) == 0


def test_emulated_server_in_place_requests(E):
    """On-device requests with in == out: a four-step length (its plan runs out of place, so the
    server stages its result through a pooled buffer) and a work-list length, in one round with
    an ordinary request; each result matches the oracle."""
    lib, ctx = E
    srv = ctypes.c_void_p()
    assert lib.cfft_server_create(ctx, 1 << 30, 1, ctypes.byref(srv)) == 0
    cases = [(1 << 16, 2, True), (1024, 3, False), (2520, 2, True)]
    reqs = []
    for i, (n, b, inplace) in enumerate(cases):
        x = seeded_complex(n, batch=b, seed=40 + i, dtype=np.complex64)
        buf = x.copy()
        y = buf if inplace else np.zeros_like(x)
        t = ctypes.c_int64()
        assert lib.cfft_server_submit(srv, n, b, 0, -1, 0, buf.ctypes.data, y.ctypes.data, 1,
                                      ctypes.byref(t)) == 0
        reqs.append((t.value, x, y, buf))
    for t, x, y, _ in reqs:
        assert lib.cfft_server_wait(srv, t) == 0, lib.cfft_last_error()
        assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6
    assert lib.cfft_server_destroy(srv) == 0


def test_emulated_measure_flags(E):
    lib, ctx = E
    x = seeded_complex(1024, batch=2, seed=2, dtype=np.complex64)
    for flag in (1, 2):
        p = _plan(lib, ctx, 1024, 2, np.complex64, flags=flag)
        assert "tuned" in _desc(lib, p)
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
        assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6


def test_emulated_argument_errors(E):
    lib, ctx = E
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, 0, 1, 1, 1, 0, 0, ctypes.byref(p)) == -1
    assert lib.cfft_plan_1d(ctx, 16, 1, 8, 16, 0, 0, ctypes.byref(p)) == -1
    assert lib.cfft_plan_1d(ctx, 16, 1, 16, 16, 5, 0, ctypes.byref(p)) == -1
    assert lib.cfft_plan_1d(ctx, 16, 1, 16, 16, 0, 9, ctypes.byref(p)) == -1
    q = _plan(lib, ctx, 16, 1, np.complex64)
    assert lib.cfft_exec(q, None, None, -1, 0, None) == -1
    assert lib.cfft_exec(q, 1, 1, 0, 0, None) == -1
    le = np.array([8, 8], dtype=np.int64)
    of = np.array([0, 4], dtype=np.int64)
    P = ctypes.POINTER(ctypes.c_int64)
    assert lib.cfft_plan_many_uneven(ctx, 2, le.ctypes.data_as(P), of.ctypes.data_as(P), 0, 0,
                                     ctypes.byref(p)) == -1


def test_emulated_wisdom_round_trip(E, tmp_path):
    lib, ctx = E
    p = _plan(lib, ctx, 2048, 4, np.complex64, flags=1)
    assert "tuned" in _desc(lib, p)
    path = str(tmp_path / "wisdom.txt").encode()
    assert lib.cfft_wisdom_export(ctx, path) == 0
    text = open(path.decode()).read()
    assert "2048 4 2048 2048 0 1 :" in text
    ctx2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(ctx2)) == 0
    assert lib.cfft_wisdom_import(ctx2, path) == 0
    q = _plan(lib, ctx2, 2048, 4, np.complex64, flags=1)
    assert "(wisdom)" in _desc(lib, q)
    x = seeded_complex(2048, batch=4, seed=8, dtype=np.complex64)
    y = np.zeros_like(x)
    assert lib.cfft_exec(q, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6
    assert lib.cfft_wisdom_import(ctx2, b"/nonexistent/wisdom") == -1
    # an import replaces the cached plan of its key: the next request follows the new choice
    forced = tmp_path / "forced.txt"
    forced.write_text("2048 4 2048 2048 0 1 : 16 64 32 2 16 -1 0 0\n")
    assert lib.cfft_wisdom_import(ctx2, str(forced).encode()) == 0
    r = _plan(lib, ctx2, 2048, 4, np.complex64, flags=1)
    d = _desc(lib, r)
    assert "(wisdom)" in d and "N1=32" in d and "CA=2" in d and "CB=16" in d, d
    y2 = np.zeros_like(x)
    assert lib.cfft_exec(q, x.ctypes.data, y2.ctypes.data, -1, 0, None) == 0  # q still valid
    assert np.array_equal(y, y2)
    assert lib.cfft_exec(r, x.ctypes.data, y2.ctypes.data, -1, 0, None) == 0
    assert rel_l2(y2, O.fft(x.astype(np.complex128), "d")) < 1e-6
    for h in (p, q, r):
        lib.cfft_plan_destroy(h)
    lib.cfft_finalize(ctx2)


@pytest.mark.parametrize("line", [
    "4096 2 4096 4096 0 1 : 32 64 0 0 0 -1 0 0",    # E the kernels are not built for
    "4096 2 4096 4096 0 1 : 16 0 0 0 0 -1 0 0",     # no threads
    "4096 2 4096 4096 0 1 : 16 64 0 3 4 -1 0 0",    # tile not a power of two
    "4096 2 4096 4096 0 1 : 16 64 -5 0 0 -1 0 0",   # negative split
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -2 0 0",    # chunk below -1
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 3 0",    # pipe out of range
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 0 2",    # large out of range
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 0 0 pm=4",   # pipelined-pass mask out of range
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 0 0 eb=12",  # pass B E not built
    "4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 0 0 pd=2"])   # LDS padding flag out of range
def test_emulated_wisdom_rejects_out_of_range_knobs(E, tmp_path, line):
    """A wisdom line whose knobs no kernel instance implements is refused (the import fails and
    the key keeps its tuned or default plan), instead of building a plan that skips butterflies."""
    lib, _ = E
    path = tmp_path / "bad.txt"
    path.write_text(line + "\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == -1
    assert b"malformed" in lib.cfft_last_error()
    path.write_text("4096 2 4096 4096 0 1 : 16 64 0 0 0 -1 0 0\n")
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
    p = _plan(lib, c2, 4096, 2, np.complex64, flags=1)
    x = seeded_complex(4096, batch=2, seed=9, dtype=np.complex64)
    y = np.zeros_like(x)
    assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6
    lib.cfft_plan_destroy(p)
    lib.cfft_finalize(c2)


@pytest.mark.parametrize("n,batch,dt", [(2520, 11, np.complex64), (2520, 9, np.complex128),
                                        (64, 150, np.complex64), (12, 500, np.complex128),
                                        (11 * 13 * 8, 7, np.complex64), (1 << 16, 3, np.complex64),
                                        (1 << 17, 1, np.complex128), (3 << 15, 2, np.complex64),
                                        (7 << 14, 1, np.complex128)])
def test_emulated_pipelined_match_plain(E, tmp_path, n, batch, dt):
    """The persistent pipelined kernels (rows, and four-step passes A and B; their few
    workgroups each loop over several line groups or tiles here) against the plain kernels,
    chosen through wisdom: identical bits, and the oracle, forward and backward, out of place and
    in place."""
    lib, ctx = E
    dcode = 0 if dt == np.complex64 else 1
    outs = {}
    for pipe in (0, 1):
        path = tmp_path / f"w{pipe}.txt"
        path.write_text(f"{n} {batch} {n} {n} {dcode} 1 : 16 256 0 0 0 -1 {pipe}\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
        p = _plan(lib, c2, n, batch, dt, flags=1)
        assert ("pipe" in _desc(lib, p)) == bool(pipe), _desc(lib, p)
        x = seeded_complex(n, batch=batch, seed=5, dtype=dt)
        res = []
        for direction, norm in ((-1, 0), (1, 1)):
            y = np.zeros_like(x)
            assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
            ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(y, ref) < _tol(n, dt), (pipe, direction)
            res.append(y)
        z = x.copy()
        assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
        assert np.array_equal(z, res[0])
        outs[pipe] = res
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,batch,dt,fields", [
    (2520, 5, np.complex128, "16 64 0 0 0 -1 0 0"),
    (2520, 9, np.complex64, "8 512 0 0 0 -1 0 0"),
    (4096, 3, np.complex128, "8 64 0 0 0 -1 0 0"),
    (1024, 7, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),
    (3 ** 7, 4, np.complex64, "16 128 0 0 0 -1 0 0"),
    (12, 40, np.complex128, "8 64 0 0 0 -1 0 0")])
def test_emulated_unpadded_lds_lines(E, tmp_path, n, batch, dt, fields):
    """Direct lines whose LDS lines carry no pad words (wisdom `pd=0`): the same bits as the
    padded plan, the oracle both ways, in place, and the knob kept through a wisdom export."""
    lib, ctx = E
    dcode = 0 if dt == np.complex64 else 1
    key = f"{n} {batch} {n} {n} {dcode} 1"
    x = seeded_complex(n, batch=batch, seed=8, dtype=dt)
    outs = {}
    for pd in (1, 0):
        path = tmp_path / f"pd{pd}.txt"
        path.write_text(f"{key} : {fields} pd={pd}\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
        p = _plan(lib, c2, n, batch, dt, flags=1)
        d = _desc(lib, p)
        assert "direct" in d and ("nopad" in d) == (pd == 0), d
        res = []
        for direction, norm in ((-1, 0), (1, 1)):
            y = np.zeros_like(x)
            assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
            ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(y, ref) < _tol(n, dt), (pd, direction)
            res.append(y)
        z = x.copy()
        assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
        assert np.array_equal(z, res[0])
        out = tmp_path / f"pd{pd}_out.txt"
        assert lib.cfft_wisdom_export(c2, str(out).encode()) == 0
        assert f" pd={pd}" in out.read_text()
        outs[pd] = res
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,batch,dt,n1", [(1 << 16, 3, np.complex64, 256),
                                           (1 << 17, 1, np.complex128, 256),
                                           (7 << 14, 1, np.complex128, 448)])
def test_emulated_per_pass_pipeline_and_capacity(E, tmp_path, n, batch, dt, n1):
    """Four-step plans whose passes differ in kernel and per-thread capacity (wisdom `pm=` runs
    one pass on the pipelined kernel, `eb=` gives pass B its own E): the same bits as the plain
    plan of that split, the oracle, and the knobs kept through a wisdom export."""
    lib, ctx = E
    dcode = 0 if dt == np.complex64 else 1
    key = f"{n} {batch} {n} {n} {dcode} 1"
    lines = {"plain": f"8 128 {n1} 4 4 -1 0 0 pr=0",
             "pipeB": f"8 128 {n1} 4 8 -1 0 0 pr=0 pm=2 eb=16",
             "pipeA": f"16 64 {n1} 8 4 -1 0 0 pr=0 pm=1 eb=8"}
    want = {"plain": ("pipe=", "E=8", "E=8"), "pipeB": ("pipe=B", "E=8", "E=16"),
            "pipeA": ("pipe=A", "E=16", "E=8")}
    x = seeded_complex(n, batch=batch, seed=6, dtype=dt)
    ref = O.fft(x.astype(np.complex128), "d")
    outs = {}
    for name, fields in lines.items():
        path = tmp_path / f"{name}.txt"
        path.write_text(f"{key} : {fields}\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
        p = _plan(lib, c2, n, batch, dt, flags=1)
        d = _desc(lib, p)
        tag, ea, eb = want[name]
        line_a, line_b = d.split(" N2=")
        assert (tag in d) == (name != "plain") and ea in line_a and eb in line_b, d
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
        assert rel_l2(y, ref) < _tol(n, dt), name
        outs[name] = y
        out = tmp_path / f"{name}_out.txt"
        assert lib.cfft_wisdom_export(c2, str(out).encode()) == 0
        text = out.read_text()
        assert (" pm=" in text) == (name != "plain") and (" eb=" in text) == (name != "plain"), text
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)
    assert np.array_equal(outs["plain"], outs["pipeB"])
    assert np.array_equal(outs["plain"], outs["pipeA"])


@pytest.mark.parametrize("n,batch,rt", [(16, 3, np.float32), (2520, 2, np.float64), (1000, 2, np.float32),
                                         (17, 3, np.float32), (1 << 16, 1, np.float32),
                                         (2 * 4099, 1, np.float32), (99, 2, np.float64), (2, 2, np.float32),
                                         (3, 2, np.float64)])
def test_emulated_real_transforms(E, n, batch, rt):
    lib, ctx = E
    rng = np.random.Generator(np.random.PCG64(n))
    rdist, cdist = n + 2, n // 2 + 3
    x = rng.standard_normal((batch, rdist)).astype(rt)
    X = np.zeros((batch, cdist), dtype=np.complex64 if rt == np.float32 else np.complex128)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d_real(ctx, n, batch, rdist, cdist, 0 if rt == np.float32 else 1, 0,
                                 ctypes.byref(p)) == 0, lib.cfft_last_error()
    assert lib.cfft_exec(p, x.ctypes.data, X.ctypes.data, -1, 0, None) == 0, lib.cfft_last_error()
    tol = 1e-6 if rt == np.float32 else 4 * np.finfo(float).eps * max(1, np.log2(n))
    ref = O.fft(x[:, :n].astype(np.complex128), "d")[:, :n // 2 + 1]
    assert rel_l2(X[:, :n // 2 + 1], ref) < tol
    assert np.all(X[:, n // 2 + 1:] == 0)
    for norm in (0, 1):
        y = np.zeros((batch, rdist), dtype=rt)
        assert lib.cfft_exec(p, X.ctypes.data, y.ctypes.data, 1, norm, None) == 0
        expect = x[:, :n] * (1.0 if norm else n)
        assert rel_l2(y[:, :n], expect) < 2 * tol
        assert np.all(y[:, n:] == 0)


@pytest.mark.parametrize("shape,axes", [((8, 12), (0, 1)), ((16, 60), (0,)), ((16, 60), (1,)),
                                        ((6, 10, 14), (0, 1, 2)), ((6, 10, 14), (0, 2)),
                                        ((5, 1, 97), (0, 1, 2)), ((4, 2520), (0, 1)),
                                        ((3, 4, 5), ())])
def test_emulated_nd(E, shape, axes):
    lib, ctx = E
    x = seeded_complex(int(np.prod(shape)), seed=len(shape), dtype=np.complex64).reshape(shape)
    mask = 0
    for a in axes:
        mask |= 1 << a
    dims = (ctypes.c_int64 * len(shape))(*shape)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_nd(ctx, len(shape), dims, mask, 0, 0, ctypes.byref(p)) == 0, lib.cfft_last_error()
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        xr = x.astype(np.complex128)
        ref = np.fft.fftn(xr, axes=axes) if direction < 0 else np.fft.ifftn(xr, axes=axes)
        if not axes:
            ref = xr
        assert rel_l2(y, ref) < 1e-6, (direction, norm)
    # in place
    z = x.copy()
    assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    ref = np.fft.fftn(x.astype(np.complex128), axes=axes) if axes else x
    assert rel_l2(z, ref) < 1e-6


def test_emulated_fourstep_chunks(E, monkeypatch):
    """Four-step in rounds of 2 transforms (the Infinity-Cache chunking) equals one round."""
    lib, ctx = E
    n, batch = 1 << 15, 5
    x = seeded_complex(n, batch=batch, seed=12, dtype=np.complex64)
    monkeypatch.setenv("CFFT_FOURSTEP_CHUNK", "2")
    p = _plan(lib, ctx, n, batch, np.complex64, idist=n + 1, odist=n)
    assert "chunk=2" in _desc(lib, p)
    xs = np.zeros((batch, n + 1), dtype=np.complex64)
    xs[:, :n] = x
    y = np.zeros_like(x)
    assert lib.cfft_exec(p, xs.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6
    monkeypatch.setenv("CFFT_FOURSTEP_CHUNK", "0")
    assert lib.cfft_empty_cache(ctx) == 0  # the cached plan was made with chunks of 2
    q = _plan(lib, ctx, n, batch, np.complex64, idist=n + 1, odist=n)
    assert "chunk" not in _desc(lib, q)
    z = np.zeros_like(x)
    assert lib.cfft_exec(q, xs.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    assert np.array_equal(y, z)


@pytest.mark.parametrize("n,batch,dt", [(60, 3, np.complex64), (1 << 16, 2, np.complex64),
                                        (17, 4, np.complex128), (2520, 2, np.complex128)])
def test_emulated_host_path_keeps_row_gaps(E, n, batch, dt):
    """cfft_exec_host (zero-copy for small single-kernel plans, pinned staging otherwise) writes
    the output rows only: the gaps between rows at odist > n keep the caller's contents."""
    lib, ctx = E
    idist, odist = n + 2, n + 7
    x = seeded_complex(idist, batch=batch, seed=11, dtype=dt)
    p = _plan(lib, ctx, n, batch, dt, idist, odist)
    out = np.full((batch, odist), 7 + 7j, dtype=dt)
    assert lib.cfft_exec_host(p, x.ctypes.data, out.ctypes.data, -1, 0) == 0
    ref = O.fft(np.ascontiguousarray(x[:, :n]).astype(np.complex128), "d", -1, False)
    assert rel_l2(out[:, :n], ref) < _tol(n, dt)
    assert np.all(out[:, n:] == 7 + 7j)
    lib.cfft_plan_destroy(p)


@pytest.mark.parametrize("n,n1,batch,dt", [(1 << 16, 256, 2, np.complex64),
                                           (1 << 17, 128, 1, np.complex128),
                                           (3 ** 10, 243, 2, np.complex64),
                                           (2520 * 16, 2520, 1, np.complex128),
                                           (1 << 14, 64, 3, np.complex64)])
def test_emulated_sixstep_matches_oracle(E, tmp_path, n, n1, batch, dt):
    """The six-step path (forced through wisdom at sizes the four-step also covers): pass A,
    twiddled transpose, M-point row FFTs, transpose; forward, backward normalised, strided rows
    and in place, against the oracle."""
    lib, ctx = E
    dcode = 0 if dt == np.complex64 else 1
    idist, odist = n + 4, n + 6
    path = tmp_path / "w.txt"
    path.write_text(f"{n} {batch} {idist} {odist} {dcode} 1 : 0 64 {n1} 0 0 -1 -1 1\n"
                    f"{n} {batch} {n} {n} {dcode} 1 : 0 64 {n1} 0 0 -1 -1 1\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
    p = _plan(lib, c2, n, batch, dt, idist, odist, flags=1)
    d = _desc(lib, p)
    assert d.startswith(f"n={n}") and "sixstep" in d and f"N1={n1}" in d, d
    x = seeded_complex(idist, batch=batch, seed=n1, dtype=dt)
    xr = np.ascontiguousarray(x[:, :n]).astype(np.complex128)
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.full((batch, odist), 3 + 3j, dtype=dt)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = O.fft(xr, "d", direction, bool(norm))
        assert rel_l2(y[:, :n], ref) < _tol(n, dt), (direction, norm)
        assert np.all(y[:, n:] == 3 + 3j)
    q = _plan(lib, c2, n, batch, dt, flags=1)
    assert "sixstep" in _desc(lib, q)
    z = np.ascontiguousarray(x[:, :n])
    assert lib.cfft_exec(q, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    assert rel_l2(z, O.fft(xr, "d", -1, False)) < _tol(n, dt)
    lib.cfft_plan_destroy(p)
    lib.cfft_plan_destroy(q)
    lib.cfft_finalize(c2)


@pytest.mark.parametrize("n,batch,dt,kind", [(10007, 2, np.complex64, "rader-large"),
                                             (3 * 10007, 1, np.complex128, "bluestein-large"),
                                             (65537, 1, np.complex64, "rader-large")])
def test_emulated_convolution_over_sixstep(E, tmp_path, n, batch, dt, kind):
    """A large convolution whose length-L FFT runs as six steps (forced through wisdom): the
    convolution epilogue (K, the A^[0] slots) is applied by the last transpose."""
    lib, ctx = E
    dcode = 0 if dt == np.complex64 else 1
    path = tmp_path / "w.txt"
    path.write_text(f"{n} {batch} {n} {n} {dcode} 1 : 0 64 0 0 0 -1 -1 1\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
    p = _plan(lib, c2, n, batch, dt, flags=1)
    d = _desc(lib, p)
    assert kind in d and "sixstep" in d, d
    x = seeded_complex(n, batch=batch, seed=3, dtype=dt)
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(y, ref) < (2e-6 if dt == np.complex64 else 1e-13), (direction, norm)
    lib.cfft_plan_destroy(p)
    lib.cfft_finalize(c2)


def _oracle_nd(x, axes, direction, normalize):
    y = x.astype(np.complex128)
    for a in axes:
        y = np.moveaxis(y, a, -1)
        shp = y.shape
        y = O.fft(np.ascontiguousarray(y.reshape(-1, shp[-1])), "d", direction,
                  normalize).reshape(shp)
        y = np.moveaxis(y, -1, a)
    return y


@pytest.mark.parametrize("shape,axes,dt", [((17, 100), (0, 1), np.complex64),
                                           ((13, 17, 40), (0, 1, 2), np.complex128),
                                           ((97, 6, 5), (0,), np.complex64),
                                           ((20000, 3), (0, 1), np.complex64),
                                           ((4, 12000, 2), (1,), np.complex128)])
def test_emulated_nd_transposed_axes(E, shape, axes, dt):
    """N-D axes one workgroup cannot take as strided lines (not smooth: 17, 97; too long: 20000
    in F32, 12000 in F64) run as transpose, batched 1-D plan, transpose; against the oracle."""
    lib, ctx = E
    rank = len(shape)
    dims = (ctypes.c_int64 * rank)(*shape)
    mask = sum(1 << (a % rank) for a in axes)
    p = ctypes.c_void_p()
    assert lib.cfft_plan_nd(ctx, rank, dims, mask, 0 if dt == np.complex64 else 1, 0,
                            ctypes.byref(p)) == 0, lib.cfft_last_error()
    assert "transposed" in _desc(lib, p), _desc(lib, p)
    x = seeded_complex(int(np.prod(shape)), seed=rank, dtype=dt).reshape(shape)
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = _oracle_nd(x, [a % rank for a in axes], direction, bool(norm))
        n_all = max(int(np.prod([shape[a % rank] for a in axes])), 2)
        assert rel_l2(y, ref) < _tol(n_all, dt), (direction, norm)
    z = x.copy()
    assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    assert rel_l2(z, _oracle_nd(x, [a % rank for a in axes], -1, False)) < _tol(max(n_all, 2), dt)
    lib.cfft_plan_destroy(p)


def test_emulated_exec_noalloc_matches_exec(E):
    """cfft_exec_noalloc on the CPU build: in-place four-step plans take their middle from the
    static work area (two calls in a row, the second smaller), other plans run as cfft_exec;
    bitwise equal to cfft_exec.  Plans that own work buffers and middles beyond the work area are
    refused."""
    lib, ctx = E
    for n, batch, dt, inplace in ((1 << 16, 2, np.complex64, True),
                                  (1 << 16, 1, np.complex64, True),
                                  (1 << 17, 1, np.complex128, True),
                                  (2520, 3, np.complex64, False),
                                  (17, 5, np.complex128, False)):
        p = _plan(lib, ctx, n, batch, dt)
        x = seeded_complex(n, batch=batch, seed=n + batch, dtype=dt)
        for direction, norm in ((-1, 0), (1, 1)):
            ref = x.copy()
            got = x.copy()
            if inplace:
                assert lib.cfft_exec(p, ref.ctypes.data, ref.ctypes.data, direction, norm,
                                     None) == 0
                assert lib.cfft_exec_noalloc(p, got.ctypes.data, got.ctypes.data, direction, norm,
                                             None) == 0, lib.cfft_last_error()
            else:
                assert lib.cfft_exec(p, x.ctypes.data, ref.ctypes.data, direction, norm, None) == 0
                assert lib.cfft_exec_noalloc(p, x.ctypes.data, got.ctypes.data, direction, norm,
                                             None) == 0, lib.cfft_last_error()
            assert np.array_equal(got, ref), (n, direction)
            want = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(got, want) < _tol(n, dt)
        lib.cfft_plan_destroy(p)
    big = _plan(lib, ctx, 1 << 20, 9, np.complex64)  # 72 MiB middle
    assert lib.cfft_exec_noalloc(big, 16, 16, -1, 0, None) == -4
    assert b"work area" in lib.cfft_last_error()
    rl = _plan(lib, ctx, 65537, 1, np.complex64)
    assert "rader-large" in _desc(lib, rl)
    assert lib.cfft_exec_noalloc(rl, 16, 32, -1, 0, None) == -4
    for h in (big, rl):
        lib.cfft_plan_destroy(h)


@pytest.mark.parametrize("n,batch,dt,fields,expect", [
    (2520, 3, np.complex64, "ra=5,7,8,9", "[5,7,8,9]"),
    (2520, 2, np.complex128, "ra=8,9,7,5", "[8,9,7,5]"),
    (2520, 2, np.complex64, "ra=3,3,2,4,7,5", "[3,3,2,4,7,5]"),
    (1024, 4, np.complex64, "ra=4,16,16", "[4,16,16]"),
    (1024, 2, np.complex128, "ra=8,8,16", "[8,8,16]"),
    (1 << 16, 1, np.complex64, "ra=4,16,4 rb=16,4,4", "N1=256 [4,16,4]"),
    (1 << 16, 2, np.complex128, "ra=2,8,16 rb=8,8,4", "[8,8,4]"),
    (3 ** 10, 1, np.complex64, "ra=3,3,3,9 rb=9,9,3", "[3,3,3,9]")])
def test_emulated_forced_radix_sequences(E, tmp_path, n, batch, dt, fields, expect):
    """Radix sequences other than the default (the search space of MEASURE / ENCHANT, the
    reference's radix_plan.jl:23-68), forced through wisdom: the plan runs them (its description
    shows them), the result matches the oracle forward and backward, and the wisdom file
    round-trips the sequences."""
    lib, _ = E
    dcode = 0 if dt == np.complex64 else 1
    path = tmp_path / "w.txt"
    path.write_text(f"{n} {batch} {n} {n} {dcode} 1 : 0 64 0 0 0 -1 0 0 {fields}\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0, lib.cfft_last_error()
    p = _plan(lib, c2, n, batch, dt, flags=1)
    d = _desc(lib, p)
    assert "(wisdom)" in d and expect in d, d
    x = seeded_complex(n, batch=batch, seed=11, dtype=dt)
    for direction, norm in ((-1, 0), (1, 1)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(y, ref) < _tol(n, dt), (direction, norm)
    out = tmp_path / "out.txt"
    assert lib.cfft_wisdom_export(c2, str(out).encode()) == 0
    assert fields in out.read_text(), out.read_text()
    lib.cfft_plan_destroy(p)
    lib.cfft_finalize(c2)


@pytest.mark.parametrize("line", [
    "2520 2 2520 2520 0 1 : 0 64 0 0 0 -1 0 0 ra=9,8,7,6",      # radix 6 is not a codelet
    "2520 2 2520 2520 0 1 : 0 64 0 0 0 -1 0 0 ra=9,8,7,",       # truncated list
    "2520 2 2520 2520 0 1 : 0 64 0 0 0 -1 0 0 ra=9,8,7,17"])
def test_emulated_wisdom_rejects_bad_radix_lists(E, tmp_path, line):
    lib, _ = E
    path = tmp_path / "bad.txt"
    path.write_text(line + "\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    rc = lib.cfft_wisdom_import(c2, str(path).encode())
    if rc == 0:  # a list of known radices whose product is not n: the plan falls back
        p = _plan(lib, c2, 2520, 2, np.complex64, flags=1)
        assert "(wisdom)" not in _desc(lib, p)
        lib.cfft_plan_destroy(p)
    else:
        assert rc == -1
    lib.cfft_finalize(c2)


def test_emulated_measure_searches_radix_sequences(E):
    """MEASURE's candidate set includes radix sequences other than the default: the tuned plan's
    description counts them, and whatever wins matches the oracle."""
    lib, ctx = E
    p = _plan(lib, ctx, 1024, 8, np.complex64, flags=1)
    d = _desc(lib, p)
    assert "tuned(measure" in d, d
    ncand = int(d.split("tuned(measure, ")[1].split(" candidates")[0])
    assert ncand >= 4 * 2 + 5 * 2, d  # shapes x pipe, plus radix sequences x threads
    x = seeded_complex(1024, batch=8, seed=12, dtype=np.complex64)
    y = np.zeros_like(x)
    assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, -1, 0, None) == 0
    assert rel_l2(y, O.fft(x.astype(np.complex128), "d")) < 1e-6
    lib.cfft_plan_destroy(p)


@pytest.mark.parametrize("n,batch,dt,fields", [
    (64, 37, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),
    (128, 9, np.complex128, "16 64 0 0 0 -1 0 0 wv=1"),
    (256, 5, np.complex64, "16 256 0 0 0 -1 0 0 wv=1"),
    (512, 3, np.complex64, "8 64 0 0 0 -1 0 0 wv=1"),
    (1024, 3, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),
    (1024, 2, np.complex128, "8 64 0 0 0 -1 0 0 wv=1"),
    (1024, 2, np.complex64, "16 64 0 0 0 -1 0 0 ra=4,16,16 wv=1"),
    (1024, 2, np.complex64, "16 64 0 0 0 -1 0 0 ra=8,16,8 wv=1"),
    (257, 6, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),
    (11 * 23, 3, np.complex128, "16 64 0 0 0 -1 0 0 wv=1")])
def test_emulated_wave_tail_matches_oracle(E, tmp_path, n, batch, dt, fields):
    """Lines whose last pass runs across wavefront lanes (the cross-lane exchanges emulated):
    direct lines of 64..1024 points, several lines per workgroup, other radix orders, and the
    fused Rader / Bluestein convolutions (L = 256 / 512); forward and backward, in place, against
    the oracle.  The plan describes itself with "wave"."""
    lib, _ = E
    dcode = 0 if dt == np.complex64 else 1
    path = tmp_path / "w.txt"
    path.write_text(f"{n} {batch} {n} {n} {dcode} 1 : {fields}\n")
    c2 = ctypes.c_void_p()
    assert lib.cfft_init(0, ctypes.byref(c2)) == 0
    assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0, lib.cfft_last_error()
    p = _plan(lib, c2, n, batch, dt, flags=1)
    d = _desc(lib, p)
    assert "(wisdom)" in d and " wave" in d, d
    x = seeded_complex(n, batch=batch, seed=21, dtype=dt)
    for direction, norm in ((-1, 0), (1, 1), (1, 0)):
        y = np.zeros_like(x)
        assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
        ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
        assert rel_l2(y, ref) < _tol(n, dt), (direction, norm)
    z = x.copy()
    assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
    assert rel_l2(z, O.fft(x.astype(np.complex128), "d")) < _tol(n, dt)
    lib.cfft_plan_destroy(p)
    lib.cfft_finalize(c2)


def test_emulated_wave_tail_refused_where_lanes_do_not_hold_the_digit(E, tmp_path):
    """A line whose last digit spans more than one wavefront (2048 = [16,16,8]: lanes 16 apart,
    8 of them) or whose last radix is not a power of two (2520) keeps its LDS last pass."""
    lib, _ = E
    for n in (2048, 2520):
        path = tmp_path / "w.txt"
        path.write_text(f"{n} 2 {n} {n} 0 1 : 16 64 0 0 0 -1 0 0 wv=1\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
        p = _plan(lib, c2, n, 2, np.complex64, flags=1)
        assert " wave" not in _desc(lib, p), _desc(lib, p)
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)


@pytest.mark.parametrize("n,batch,dt,kind", [(65537, 2, np.complex64, "rader-large"),
                                             (65537, 1, np.complex128, "rader-large"),
                                             (3 * 8191, 2, np.complex64, "bluestein-large"),
                                             (100003, 1, np.complex128, "rader-large"),
                                             (40009, 2, np.complex64, "rader-large"),
                                             (17 * 2357, 2, np.complex64, "bluestein-large")])
def test_emulated_large_convolution_fused_and_unfused(E, tmp_path, n, batch, dt, kind):
    """A large convolution over a four-step FFT_L runs as three fused launches by default (gather
    fused into the first pass A, the first pass B, the spectrum product and the second pass A in
    one workgroup, the second pass B fused with the scatter); forced through wisdom, the four-launch
    path (gather, two four-steps, scatter) is kept.  Both against the oracle, forward and
    backward, in place."""
    lib, _ = E
    dcode = 0 if dt == np.complex64 else 1
    x = seeded_complex(n, batch=batch, seed=31, dtype=dt)
    for fc in (1, 0):
        path = tmp_path / f"w{fc}.txt"
        path.write_text(f"{n} {batch} {n} {n} {dcode} 1 : 0 64 0 0 0 -1 -1 0 fc={fc}\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0
        p = _plan(lib, c2, n, batch, dt, flags=1)
        d = _desc(lib, p)
        assert kind in d and ("fused3" in d) == bool(fc), d
        for direction, norm in ((-1, 0), (1, 1)):
            y = np.zeros_like(x)
            assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
            ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(y, ref) < (1e-6 if dt == np.complex64 else 1e-13), (fc, direction)
        z = x.copy()
        assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
        assert rel_l2(z, O.fft(x.astype(np.complex128), "d")) < (1e-6 if dt == np.complex64 else 1e-13)
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)


@pytest.mark.parametrize("n,batch,fields,expect", [
    (1 << 16, 3, "16 64 0 4 8 0 0 0 pr=3", "pair=AB"),
    (60 * 84, 5, "16 64 60 8 4 0 0 0 pr=3", "pair=AB"),      # ragged tiles: 84 = 10.5 x 8
    (60 * 84, 2, "16 64 60 16 2 0 0 0 pr=1", "pair=A"),      # last tile 4 of 16 lines
    (64 * 81, 3, "16 64 64 4 4 0 0 0 pr=3", "pair=B"),       # N2 odd: pass A keeps one line
    (1 << 14, 2, "8 64 0 2 16 0 0 0 pr=2", "pair=B"),        # E = 8, one pair per A tile
    (1 << 16, 2, "8 64 0 4 8 0 0 0 pr=3", "pair=AB"),         # E = 8, both passes (the default)
    (48 * 96, 3, "8 64 48 4 16 0 0 0 pr=3", "pair=AB"),      # E = 8 lines of radix 3
])
def test_emulated_pair_fourstep_matches_oracle_and_float_bits(E, tmp_path, n, batch, fields,
                                                              expect):
    """ComplexF32 four-step passes with two adjacent lines per thread, forced through wisdom
    (pr=): against the oracle forward and backward, bit for bit against the same plan without
    pairs, in place, and with input and output addresses off 16-byte alignment (which run the
    float kernels on the same tiles)."""
    lib, _ = E
    x = seeded_complex(n, batch=batch, seed=31, dtype=np.complex64)
    outs = {}
    for tag, line in (("pair", fields), ("float", fields.rsplit(" pr=", 1)[0] + " pr=0")):
        path = tmp_path / f"w_{tag}.txt"
        path.write_text(f"{n} {batch} {n} {n} 0 1 : {line}\n")
        c2 = ctypes.c_void_p()
        assert lib.cfft_init(0, ctypes.byref(c2)) == 0
        assert lib.cfft_wisdom_import(c2, str(path).encode()) == 0, lib.cfft_last_error()
        p = _plan(lib, c2, n, batch, np.complex64, flags=1)
        d = _desc(lib, p)
        assert "(wisdom)" in d and "fourstep" in d and ((expect in d) == (tag == "pair")), d
        res = []
        for direction, norm in ((-1, 0), (1, 1)):
            y = np.zeros_like(x)
            assert lib.cfft_exec(p, x.ctypes.data, y.ctypes.data, direction, norm, None) == 0
            ref = O.fft(x.astype(np.complex128), "d", direction, bool(norm))
            assert rel_l2(y, ref) < _tol(n, np.complex64), (tag, direction)
            res.append(y)
        z = x.copy()
        assert lib.cfft_exec(p, z.ctypes.data, z.ctypes.data, -1, 0, None) == 0
        res.append(z)
        # one element off 16-byte alignment, in and out
        xb = np.zeros(n * batch + 1, np.complex64)
        yb = np.zeros(n * batch + 1, np.complex64)
        base = 1 if xb.ctypes.data % 16 == 0 else 0
        xb[base:base + n * batch] = x.ravel()
        assert (xb.ctypes.data + 8 * base) % 16 == 8
        assert lib.cfft_exec(p, xb.ctypes.data + 8 * base, yb.ctypes.data + 8 * base, -1, 0,
                             None) == 0
        res.append(yb[base:base + n * batch].reshape(batch, n).copy())
        outs[tag] = res
        lib.cfft_plan_destroy(p)
        lib.cfft_finalize(c2)
    for a, b in zip(outs["pair"], outs["float"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(outs["pair"][0].view(np.uint32), outs["pair"][2].view(np.uint32))
    assert np.array_equal(outs["pair"][0].view(np.uint32), outs["pair"][3].view(np.uint32))


_STRIDE_SCRIPT = r"""
import ctypes, sys
sys.path.insert(0, sys.argv[1])
import catabrafft._lib as L
lib = L._sig(ctypes.CDLL(sys.argv[2]))
ctx = ctypes.c_void_p()
assert lib.cfft_init(0, ctypes.byref(ctx)) == 0
assert lib.cfft_wisdom_import(ctx, sys.argv[3].encode()) == 0
out = []
for n in (1 << 21, 1 << 20):
    p = ctypes.c_void_p()
    assert lib.cfft_plan_1d(ctx, n, 4, n, n, 0, 1, ctypes.byref(p)) == 0
    buf = ctypes.create_string_buffer(1024)
    lib.cfft_plan_describe(p, buf, len(buf))
    out.append(buf.value.decode())
print("\n".join(out))
"""


def test_emulated_lds_stride_does_not_depend_on_earlier_plans(E, tmp_path):
    """The LDS line stride of a four-step line is chosen by a bank-conflict model that replays
    the line's passes; a plan built after others with the same line length but other radix
    sequences gets the stride it gets alone (a fresh process), so a tuned plan and its wisdom
    line plan the same kernels."""
    w = tmp_path / "w.txt"
    w.write_text("2097152 4 2097152 2097152 0 1 : 16 64 2048 4 8 0 0 0 ra=8,16,16 rb=16,16,4\n"
                 "1048576 4 1048576 1048576 0 1 : 16 64 2048 4 8 0 0 0 ra=16,16,8 rb=4,8,16\n")
    w2 = tmp_path / "w2.txt"
    w2.write_text(w.read_text().splitlines()[1] + "\n")
    lib_path = os.path.join(EMU, "_build", "libcatabrafft_emu.so")
    pkg = os.path.join(ROOT, "catabrafft.jl_amd")
    both = subprocess.run([sys.executable, "-c", _STRIDE_SCRIPT, pkg, lib_path, str(w)],
                          check=True, capture_output=True, text=True).stdout.split("\n")
    alone = subprocess.run([sys.executable, "-c", _STRIDE_SCRIPT.replace("(1 << 21, 1 << 20)",
                                                                        "(1 << 20,)"),
                            pkg, lib_path, str(w2)],
                           check=True, capture_output=True, text=True).stdout.split("\n")
    assert "ra=" not in both[1] and "(wisdom)" in both[1], both
    assert both[1] == alone[0], (both[1], alone[0])
