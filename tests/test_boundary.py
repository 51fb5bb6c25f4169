"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every declared entry
point, argument errors are reported without a GPU, and the product never touches the oracle."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "catabrafft.jl_amd")


@pytest.fixture(scope="module")
def L():
    import catabrafft._lib as lib

    if not os.path.exists(lib.LIB_PATH):  # the CPU suite may run before build()
        subprocess.run([sys.executable, os.path.join(PKG, "build.py")], check=True)
    return lib


def test_library_exports_every_declared_symbol(L):
    names = L.declared_functions()
    assert len(names) >= 23
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(cfft_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = L.lib()
    for n in names:
        getattr(lib, n)


def test_python_bindings_cover_the_header(L):
    src = open(os.path.join(PKG, "catabrafft", "_lib.py")).read()
    for n in L.declared_functions():
        assert f'"{n}"' in src, n


def test_errors_without_a_gpu(L):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible; this checks the no-device error path")
    lib = L.lib()
    h = ctypes.c_void_p()
    rc = lib.cfft_init(0, ctypes.byref(h))
    assert rc == L.CFFT_ERR_HIP
    assert b"device" in lib.cfft_last_error()
    assert lib.cfft_plan_1d(None, 16, 1, 16, 16, 0, 0, ctypes.byref(h)) == L.CFFT_ERR_INVALID
    assert lib.cfft_exec(None, None, None, -1, 0, None) == L.CFFT_ERR_INVALID
    assert lib.cfft_server_poll(None, 1) == L.CFFT_ERR_INVALID
    assert lib.cfft_version().startswith(b"catabrafft")


def test_python_api_rejects_bad_arguments_before_any_device_call():
    import catabrafft as C

    with pytest.raises(TypeError):
        C.fft(np.zeros(8))  # real input: CatabraFFT transforms complex vectors
    with pytest.raises(TypeError):
        C.fft(np.zeros(8, dtype=np.int32))
    with pytest.raises(ValueError):
        C.fft(np.zeros((2, 2, 2), dtype=np.complex64))
    with pytest.raises(ValueError):
        C.Spell(8, np.complex64, flag=7)


def test_flags_match_the_reference():
    import catabrafft as C

    assert (C.NO_FLAG, C.MEASURE, C.ENCHANT) == (0, 1, 2)  # spells.jl:4-8
    hdr = open(os.path.join(ROOT, "include", "catabrafft.h")).read()
    assert "#define CFFT_NO_FLAG 0" in hdr and "#define CFFT_ENCHANT 2" in hdr


def test_generated_codelets_are_current(tmp_path):
    sys.path.insert(0, os.path.join(PKG, "gen"))
    import gen_codelets as G

    G.self_check()
    G.check_mul_diag()
    G.check_lane_w16()
    tmp = str(tmp_path / "codelets_check.h")
    subprocess.run([sys.executable, os.path.join(PKG, "gen", "gen_codelets.py"), tmp], check=True,
                   capture_output=True)
    assert open(tmp).read() == open(os.path.join(PKG, "csrc", "codelets.h")).read()


def test_product_never_uses_the_oracle():
    """No CPU fallback: the package neither imports nor links the oracle."""
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".jl")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liborc" not in text, f
    import catabrafft._lib as lib

    libs = subprocess.run(["readelf", "-d", lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "liborc" not in libs


def test_julia_shim_binds_the_c_abi():
    shim = open(os.path.join(PKG, "julia", "CatabraFFT.jl")).read()
    for name in ("cfft_init", "cfft_plan_1d", "cfft_exec_host", "cfft_empty_cache"):
        assert name in shim
    for api in ("function fft(", "function ifft(", "function bfft(", "function plan_fft(",
                "function empty_cache("):
        assert api in shim
