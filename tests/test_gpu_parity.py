"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Tolerances (relative L2 against the oracle's long-double or double restatement):
  ComplexF64: 4 * eps64 * log2(n)   ComplexF32: 1e-6
The size matrix replays the reference's test/runtests.jl:4-46 (mixed 10^1..10^7, Rader
[11,13,17,19,43,89,1721], 2^1..2^20, 3^1..3^13, 5^1..5^10, 7^1..7^8, round trips), whose own
comparator is the much looser sqrt(eps) (runtests.jl:13).
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, seeded_complex

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps


def tol(n, dt):
    if np.dtype(dt) == np.complex128:
        return 4 * EPS64 * max(1.0, np.log2(n))
    return 1e-6


def oracle_ref(x, direction=-1, normalize=False):
    """Long-double restatement for n <= 2^16, double restatement above (cost)."""
    x = np.asarray(x, dtype=np.complex128)
    n = x.shape[-1]
    if n <= (1 << 16):
        return O.fft_l(x, direction, normalize)
    return O.fft(x, "d", direction, normalize, threads=8)


def run_dev(fn, x):
    import torch

    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    y = fn(xd)
    torch.cuda.synchronize()
    return y.cpu().numpy()


SIZE_MATRIX = (
    [10 ** k for k in range(1, 7)]
    + [11, 13, 17, 19, 43, 89, 1721]
    + [2 ** k for k in range(1, 21)]
    + [3 ** k for k in range(1, 14)]
    + [5 ** k for k in range(1, 11)]
    + [7 ** k for k in range(1, 9)]
)


@pytest.mark.parametrize("n", SIZE_MATRIX)
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_reference_size_matrix(gpu_ready, n, dt):
    import catabrafft as C

    x = seeded_complex(n, seed=n % 1000 + 7, dtype=dt)
    print(f"size-matrix n={n} dtype={np.dtype(dt)}", flush=True)
    y = run_dev(C.fft, x)
    assert rel_l2(y, oracle_ref(x)) < tol(n, dt)
    yi = run_dev(C.ifft, x)
    assert rel_l2(yi, oracle_ref(x, 1, True)) < tol(n, dt)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_ten_million_mixed(gpu_ready, dt):
    import catabrafft as C

    n = 10 ** 7
    x = seeded_complex(n, seed=77, dtype=dt)
    y = run_dev(C.fft, x)
    assert rel_l2(y, O.fft(x.astype(np.complex128), "d", threads=8)) < tol(n, dt)


@pytest.mark.parametrize("n", [89, 121, 2 ** 10, 2 ** 11, 3 ** 7, 5 ** 6, 10 ** 4])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_round_trip(gpu_ready, n, dt):
    """runtests.jl:39-46."""
    import catabrafft as C

    x = seeded_complex(n, seed=3, dtype=dt)
    back = run_dev(lambda t: C.ifft(C.fft(t)), x)
    assert rel_l2(back, x) < 2 * tol(n, dt)


def test_cfg1_single_1024_f64_host_path(gpu_ready):
    """n=1024 ComplexF64 forward + inverse through the host-pointer ABI (the Julia drop-in)."""
    import catabrafft as C

    x = seeded_complex(1024, seed=1)
    y = C.fft(x)
    assert rel_l2(y, O.fft_l(x)) < tol(1024, np.complex128)
    back = C.ifft(y)
    assert rel_l2(back, x) < 2 * tol(1024, np.complex128)
    assert rel_l2(C.bfft(y) / 1024, x) < 2 * tol(1024, np.complex128)


@pytest.mark.parametrize("n,batch", [(2520, 64), (1 << 20, 2)])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_cfg2_batched_rows(gpu_ready, n, batch, dt):
    import catabrafft as C

    x = seeded_complex(n, batch=batch, seed=2, dtype=dt)
    y = run_dev(C.fft, x)
    ref = oracle_ref(x)
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) < tol(n, dt), b


@pytest.mark.parametrize("p,batch", [(17, 1000), (257, 100), (4099, 8), (65537, 2), (1048573, 1)])
def test_cfg3_rader_primes(gpu_ready, p, batch):
    import catabrafft as C

    x = seeded_complex(p, batch=batch, seed=3, dtype=np.complex64, dist="uniform")
    y = run_dev(C.fft, x)
    ref = oracle_ref(x)
    assert rel_l2(y, ref) < 1e-6
    yi = run_dev(C.ifft, x)
    assert rel_l2(yi, oracle_ref(x, 1, True)) < 1e-6


@pytest.mark.parametrize("p", [17, 257, 65537])
def test_rader_repeat_calls_identical(gpu_ready, p):
    """prime.jl:61 overwrites W on the first call for Fermat primes; here calls repeat exactly."""
    import catabrafft as C

    x = seeded_complex(p, seed=4, dtype=np.complex64)
    a = run_dev(C.fft, x)
    b = run_dev(C.fft, x)
    c = run_dev(C.ifft, x)
    d = run_dev(C.fft, x)
    assert np.array_equal(a, b) and np.array_equal(a, d)
    assert rel_l2(c, oracle_ref(x, 1, True)) < 1e-6


@pytest.mark.parametrize("n", [60, 1000, 2310, 4096, 17 * 19, 3 * 8191, 100003 * 2])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_bluestein_and_mixed_lengths(gpu_ready, n, dt):
    import catabrafft as C

    x = seeded_complex(n, seed=n, dtype=dt)
    y = run_dev(C.fft, x)
    assert rel_l2(y, oracle_ref(x)) < tol(n, dt)


# ------------------------------------------------------------------ layouts and edge cases
@pytest.mark.parametrize("n", [16, 1024, 2520, 17, 4099, 1 << 18, 1000])
def test_in_place_and_strided(gpu_ready, n):
    import torch

    import catabrafft as C

    batch, idist, odist = 3, n + 5, n + 9
    x = seeded_complex(idist * batch, seed=1, dtype=np.complex64)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(odist * batch, dtype=torch.complex64, device="cuda")
    p = C.Plan(n, batch, np.complex64, idist=idist, odist=odist)
    p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False,
               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    for b in range(batch):
        ref = oracle_ref(x[b * idist:b * idist + n])
        assert rel_l2(y[b * odist:b * odist + n], ref) < 1e-6
        assert np.all(y[b * odist + n:(b + 1) * odist] == 0)  # gaps untouched
    assert np.array_equal(xd.cpu().numpy(), x)  # input untouched
    # in place
    q = C.Plan(n, batch, np.complex64)
    z = torch.from_numpy(seeded_complex(n, batch=batch, seed=2, dtype=np.complex64)).cuda()
    z0 = z.cpu().numpy()
    if q.describe().find("large") >= 0:
        pytest.skip("large convolution plans are out of place only")
    q.exec_ptr(z.data_ptr(), z.data_ptr(), -1, False, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rel_l2(z.cpu().numpy(), oracle_ref(z0)) < 1e-6


def test_trivial_and_empty(gpu_ready):
    import catabrafft as C

    x = np.array([1.5 - 2j], dtype=np.complex128)
    assert np.array_equal(C.fft(x), x)
    assert np.array_equal(C.ifft(x), x)
    e = np.zeros((0, 8), dtype=np.complex64)
    assert C.fft(e).shape == (0, 8)


@pytest.mark.parametrize("flag", [1, 2])
def test_measure_and_enchant_plans_agree(gpu_ready, flag):
    import catabrafft as C

    for n in (1024, 2520, 1 << 16):
        x = seeded_complex(n, batch=4, seed=8, dtype=np.complex64)
        p = C.plan_fft(x, flag)
        y = p * x
        assert rel_l2(y, oracle_ref(x)) < 1e-6, (n, p.describe())


def test_plan_api(gpu_ready):
    import catabrafft as C

    x = seeded_complex(256, seed=6)
    p = C.plan_fft(x)
    y = np.empty_like(x)
    x0 = x.copy()
    C.mul(y, p, x)
    assert np.array_equal(x, x0)
    assert rel_l2(y, O.fft_l(x)) < tol(256, np.complex128)
    assert rel_l2(p(x), y) == 0.0
    assert rel_l2(C.plan_ifft(x) * y, x) < 1e-14
    with pytest.raises(ValueError):
        p * seeded_complex(128, seed=1)
    C.empty_cache()
    assert rel_l2(C.fft(x), y) == 0.0


@pytest.mark.parametrize("n,batch,dt,inplace,code", [
    (65537, 3, np.complex64, False, None),       # large convolution: plan-owned work lines
    (1 << 20, 3, np.complex128, True, None),     # four-step in place: plan-owned middle
    (17 * 19 * 23, 5, np.complex64, False, 2),   # ComplexF16 staged through plan-owned buffers
    (2520, 7, np.complex64, False, None),        # direct: no work buffers, runs concurrently
])
def test_one_plan_on_two_streams(gpu_ready, n, batch, dt, inplace, code):
    """Executions of one plan enqueued on two streams at once give the results of sequential
    executions, bit for bit (those that use the plan's own work buffers are ordered on the
    device)."""
    torch = gpu_ready
    import catabrafft as C

    if code == 2:
        p = C.Plan(n, batch, 2)
        mk = lambda s: torch.from_numpy(np.stack([seeded_complex(n, batch=batch, seed=s).real,
                                                  seeded_complex(n, batch=batch, seed=s).imag],
                                                 -1).astype(np.float16)).cuda()
    else:
        p = C.Plan(n, batch, dt)
        mk = lambda s: torch.from_numpy(seeded_complex(n, batch=batch, seed=s, dtype=dt)).cuda()
    xs = [mk(1), mk(2)]
    refs = []
    for x in xs:
        y = x.clone() if inplace else torch.empty_like(x)
        p.exec_ptr(x.data_ptr() if not inplace else y.data_ptr(), y.data_ptr(), -1, False,
                   torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        refs.append(y.cpu())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    outs = []
    for x, s in zip(xs, streams):
        y = x.clone() if inplace else torch.empty_like(x)
        torch.cuda.synchronize()
        outs.append(y)
    for x, y, s in zip(xs, outs, streams):
        p.exec_ptr(y.data_ptr() if inplace else x.data_ptr(), y.data_ptr(), -1, False,
                   s.cuda_stream)
    torch.cuda.synchronize()
    for y, r in zip(outs, refs):
        assert torch.equal(y.cpu().view(torch.int16) if code == 2 else y.cpu(), 
                           r.view(torch.int16) if code == 2 else r)


@pytest.mark.parametrize("n,batch,dt,kind", [(65537, 4, np.complex64, "rader-large"),
                                             (1048573, 1, np.complex64, "rader-large"),
                                             (65537, 2, np.complex128, "rader-large"),
                                             (3 * 8191, 3, np.complex64, "bluestein-large"),
                                             (40009, 2, np.complex64, "rader-large"),
                                             (17 * 2357, 2, np.complex128, "bluestein-large")])
def test_large_convolution_fused_and_unfused(gpu_ready, tmp_path, n, batch, dt, kind):
    """Large Rader / Bluestein convolutions: the three fused launches (default) and the
    four-launch path (gather, two four-steps, scatter; forced through wisdom) against the oracle,
    forward and backward normalised (prime.jl:5-82)."""
    import catabrafft as C

    code = 0 if dt == np.complex64 else 1
    x = seeded_complex(n, batch, seed=n % 101, dtype=dt)
    for fc in (1, 0):
        w = tmp_path / f"w{fc}.txt"
        w.write_text(f"{n} {batch} {n} {n} {code} 1 : 0 64 0 0 0 -1 -1 0 fc={fc}\n")
        C.import_wisdom(str(w))
        p = C.Plan(n, batch, dt, flags=C.MEASURE)
        d = p.describe()
        print(f"large conv fc={fc}: {d}", flush=True)
        assert kind in d and ("fused3" in d) == bool(fc), d
        for direction, norm in ((-1, False), (1, True)):
            y = run_dev(lambda xd: p.exec_torch(xd, direction=direction, normalize=norm), x)
            assert rel_l2(y, oracle_ref(x, direction, norm)) <= max(tol(n, dt), 1e-13), (fc, direction)
