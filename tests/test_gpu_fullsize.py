"""The workloads at their full sizes on the GPU, through the C ABI.

The batched cases (n = 2520 x 4096 and n = 2^20 x 64, ComplexF32 and ComplexF64) and the Rader
primes (p x floor(2^24 / p)) run at the batch sizes the benchmark times, so the launch geometry
the benchmark uses (grid sizes, XCD remap, tile and line counts) is the one checked:

  * 8 rows spread across the batch against the oracle (relative L2: 1e-6 for ComplexF32,
    4 eps log2 n for ComplexF64);
  * over the whole batch, on the device: the round trip ifft(fft(x)) = x and Parseval
    ||X||^2 = n ||x||^2 per row (size-independent properties of the DFT).

The serving loop is checked with requests of up to 256 transforms and lengths up to 2^18: every
result against the oracle, and bitwise against the same transform planned alone.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, seeded_complex

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps


def _tol(n, dt):
    return 4 * EPS64 * max(1.0, np.log2(n)) if np.dtype(dt) == np.complex128 else 1e-6


def _device_input(torch, n, batch, dt, seed, dist="normal"):
    g = torch.Generator(device="cuda").manual_seed(seed)
    rdt = torch.float32 if np.dtype(dt) == np.complex64 else torch.float64
    if dist == "normal":
        re = torch.randn(batch, n, generator=g, device="cuda", dtype=rdt)
        im = torch.randn(batch, n, generator=g, device="cuda", dtype=rdt)
    else:
        re = torch.rand(batch, n, generator=g, device="cuda", dtype=rdt) * 2 - 1
        im = torch.rand(batch, n, generator=g, device="cuda", dtype=rdt) * 2 - 1
    return torch.complex(re, im)


def _spread_rows(batch, k=8):
    return sorted(set(int(round(i * (batch - 1) / (k - 1))) for i in range(k)))


def _check_full(torch, C, x, n, batch, dt, direction_tol_scale=1.0):
    plan = C.Plan(n, batch, dt)
    stream = torch.cuda.current_stream().cuda_stream
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    plan.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, stream)
    plan.exec_ptr(y.data_ptr(), z.data_ptr(), 1, True, stream)
    torch.cuda.synchronize()
    rows = _spread_rows(batch)
    xr = x[rows].cpu().numpy().astype(np.complex128)
    ref = O.fft(xr, "d", threads=8) if n > 4096 else O.fft_l(xr)
    yr = y[rows].cpu().numpy()
    for i, r in enumerate(rows):
        assert rel_l2(yr[i], ref[i]) < _tol(n, dt) * direction_tol_scale, (n, batch, r)
    # whole batch on the device: round trip and Parseval, per row
    acc = torch.float64
    rt = (torch.linalg.vector_norm((z - x).to(torch.complex128), dim=1) /
          torch.linalg.vector_norm(x.to(torch.complex128), dim=1))
    assert float(rt.max()) < 2 * _tol(n, dt) * direction_tol_scale, float(rt.max())
    ex = torch.linalg.vector_norm(x.to(torch.complex128), dim=1).to(acc) ** 2
    ey = torch.linalg.vector_norm(y.to(torch.complex128), dim=1).to(acc) ** 2
    parseval = (ey / (n * ex) - 1).abs()
    assert float(parseval.max()) < 4 * _tol(n, dt) * direction_tol_scale, float(parseval.max())
    return plan.describe()


@pytest.mark.parametrize("n,batch", [(2520, 4096), (1 << 20, 64)])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_batched_full_size(gpu_ready, n, batch, dt):
    torch = gpu_ready
    import catabrafft as C

    print(f"full-size batched n={n} batch={batch} {np.dtype(dt).name}", flush=True)
    x = _device_input(torch, n, batch, dt, seed=2)
    d = _check_full(torch, C, x, n, batch, dt)
    print(d, flush=True)


@pytest.mark.parametrize("p", [17, 257, 4099, 65537, 1048573])
def test_rader_full_size(gpu_ready, p):
    torch = gpu_ready
    import catabrafft as C

    batch = (1 << 24) // p
    print(f"full-size Rader p={p} batch={batch}", flush=True)
    x = _device_input(torch, p, batch, np.complex64, seed=3, dist="uniform")
    d = _check_full(torch, C, x, p, batch, np.complex64)
    print(d, flush=True)


def test_server_large_requests_match_alone(gpu_ready):
    """Requests of up to 256 transforms and lengths up to 2^18 (the serving configuration's
    extremes), device-resident and host, ComplexF32 and ComplexF64, several in flight at once."""
    torch = gpu_ready
    import catabrafft as C

    srv = C.Server(hbm_cap_bytes=64 << 30, nstreams=4)
    cases = [(1 << 18, 4, np.complex64, True), (1 << 18, 1, np.complex128, False),
             (262139, 2, np.complex64, True),                # prime near 2^18
             (3 * 5 * 7 * 11 * 13 * 17, 3, np.complex64, True),
             (4096, 256, np.complex64, True), (2520, 256, np.complex128, True),
             (1021, 256, np.complex64, False), (64, 256, np.complex64, True),
             (100000, 16, np.complex64, True), (131072, 32, np.complex64, True)]
    reqs = []
    for i, (n, batch, dt, dev) in enumerate(cases):
        print(f"server request n={n} batch={batch} {np.dtype(dt).name} device={dev}", flush=True)
        x = seeded_complex(n, batch=batch, seed=200 + i, dtype=dt)
        t = srv.submit(torch.from_numpy(x).cuda() if dev else x)
        reqs.append((t, x))
    for t, x in reqs:
        y = srv.wait(t)
        y = y.cpu().numpy() if hasattr(y, "cpu") else y
        n, batch = x.shape[1], x.shape[0]
        rows = _spread_rows(batch)
        ref = O.fft(x[rows].astype(np.complex128), "d", threads=8)
        for i, r in enumerate(rows):
            assert rel_l2(y[r], ref[i]) < _tol(n, x.dtype), (n, batch, r)
        alone = C.fft(x)
        assert np.array_equal(alone, y), (n, batch)
    srv.close()
