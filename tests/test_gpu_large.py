"""Transforms longer than the four-step covers (no split into two factors that fit a workgroup:
above about 9800^2 in F64 and 16384^2 in F32, or such as 3^17), and convolutions over such
lengths: the six-step path on the GPU.

At these sizes the checks are size-independent closed forms and invariants rather than a full
oracle transform: the DFT of an impulse at j is e^{-2 pi i jk/n} (checked at sampled k, and at
k = 0 for a random input: X[0] = sum x), and ifft(fft(x)) = x.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(1 << 27, np.complex128), (1 << 29, np.complex64), (3 ** 17, np.complex64),
         (3 ** 17, np.complex128)]


def _tol(dt, n):
    return 4e-6 if dt == np.complex64 else 4 * np.finfo(float).eps * np.log2(n)


@pytest.mark.parametrize("n,dt", CASES)
def test_sixstep_impulse_and_round_trip(gpu_ready, n, dt):
    torch = gpu_ready
    import catabrafft as C

    tdt = torch.complex64 if dt == np.complex64 else torch.complex128
    p = C.Plan(n, 1, dt)
    assert "sixstep" in p.describe(), p.describe()
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.Generator(np.random.PCG64(n % 1000003))
    j = int(rng.integers(1, n))
    x = torch.zeros(n, dtype=tdt, device="cuda")
    x[j] = 1
    y = torch.empty_like(x)
    p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    ks = np.concatenate([np.arange(64), rng.integers(0, n, 4096), n - 1 - np.arange(64)])
    got = y[torch.from_numpy(ks).cuda()].cpu().numpy().astype(np.complex128)
    want = np.exp(-2j * np.pi * ((j * ks.astype(object)) % n).astype(np.float64) / n)
    assert np.max(np.abs(got - want)) < 50 * _tol(dt, n), (j, np.max(np.abs(got - want)))
    del x, y
    # random input: X[0] = sum x, and the round trip
    g = torch.Generator(device="cuda").manual_seed(7)
    re = torch.randn(n, generator=g, device="cuda", dtype=torch.float64)
    im = torch.randn(n, generator=g, device="cuda", dtype=torch.float64)
    x = torch.complex(re, im).to(tdt)
    del re, im
    y = torch.empty_like(x)
    p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    x0 = complex(x.to(torch.complex128).sum().item())
    assert abs(complex(y[0].item()) - x0) < 1e-2 * np.sqrt(n) * (1 if dt == np.complex64 else 1e-6)
    p.exec_ptr(y.data_ptr(), y.data_ptr(), 1, True, s)  # in place
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm((y - x).to(torch.complex128)) /
           torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert err < _tol(dt, n), err


@pytest.mark.parametrize("n,dt,kind", [(50000017, np.complex64, "rader-large"),
                                       (17 << 22, np.complex128, "bluestein-large")])
def test_convolution_over_sixstep_impulse_and_round_trip(gpu_ready, n, dt, kind):
    """Non-smooth lengths whose convolution FFT (L >= 2n - 1) exceeds the four-step, whose F64
    lines reach about 9800 points: a prime above 5e7 (FP64 convolution of F32 data, L >= 1e8)
    and 17 * 2^22 in F64 (L >= 1.4e8)."""
    torch = gpu_ready
    import catabrafft as C

    tdt = torch.complex64 if dt == np.complex64 else torch.complex128
    p = C.Plan(n, 1, dt)
    d = p.describe()
    assert kind in d and "sixstep" in d, d
    s = torch.cuda.current_stream().cuda_stream
    rng = np.random.Generator(np.random.PCG64(n % 1000003))
    j = int(rng.integers(1, n))
    x = torch.zeros(n, dtype=tdt, device="cuda")
    x[j] = 1
    y = torch.empty_like(x)
    p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    ks = np.concatenate([np.arange(64), rng.integers(0, n, 4096), n - 1 - np.arange(64)])
    got = y[torch.from_numpy(ks).cuda()].cpu().numpy().astype(np.complex128)
    want = np.exp(-2j * np.pi * ((j * ks.astype(object)) % n).astype(np.float64) / n)
    assert np.max(np.abs(got - want)) < 50 * _tol(dt, n), np.max(np.abs(got - want))
    del x, y
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.complex(torch.randn(n, generator=g, device="cuda", dtype=torch.float64),
                      torch.randn(n, generator=g, device="cuda", dtype=torch.float64)).to(tdt)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s)
    p.exec_ptr(y.data_ptr(), z.data_ptr(), 1, True, s)
    torch.cuda.synchronize()
    err = (torch.linalg.vector_norm((z - x).to(torch.complex128)) /
           torch.linalg.vector_norm(x.to(torch.complex128))).item()
    assert err < _tol(dt, n), err

