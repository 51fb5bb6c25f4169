"""Seeded sweep of lengths against the oracle.

The size matrix of the reference's tests (test/runtests.jl:4-46) holds powers, a few primes and
powers of ten.  This sweep adds lengths drawn at random, so that every planning branch meets
lengths no fixed list names: 7-smooth lengths (direct, four-step), arbitrary integers (mostly
with a large prime factor: Rader for primes, Bluestein for the rest, fused or large, FP32 or FP64
convolution) and products of a smooth part with a prime.  Each case runs forward and normalised
backward against the oracle at the parity tolerances: ComplexF64 4 * eps64 * log2(n),
ComplexF32 1e-6 (relative L2).
"""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import oracle_ref, run_dev, tol

pytestmark = pytest.mark.gpu


def _smooth(rng, lo, hi):
    while True:
        n = 1
        while n < lo:
            n *= int(rng.choice([2, 3, 5, 7]))
        if n <= hi:
            return n


def _lengths():
    rng = np.random.Generator(np.random.PCG64(2026))
    out = []
    for i in range(40):
        lo = int(np.exp(rng.uniform(np.log(2), np.log(1 << 20))))
        out.append(_smooth(rng, lo, 1 << 21))
    for i in range(40):
        out.append(int(np.exp(rng.uniform(np.log(2), np.log(1 << 20)))))
    for i in range(20):
        p = int(rng.choice([11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71]))
        out.append(p * _smooth(rng, 1, 1 << 14))
    return out


CASES = [(n, np.complex64 if i % 2 == 0 else np.complex128) for i, n in enumerate(_lengths())]


@pytest.mark.parametrize("n,dt", CASES)
def test_sweep_matches_oracle(gpu_ready, n, dt):
    import catabrafft as C

    rng = np.random.Generator(np.random.PCG64(n))
    x = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(dt)
    p = C.Plan(n, 2, dt)
    print(f"sweep n={n} {np.dtype(dt)}: {p.describe()}", flush=True)
    y = run_dev(lambda xd: p.exec_torch(xd), x)
    err = rel_l2(y, oracle_ref(x))
    assert err <= tol(n, dt), (p.describe(), err)
    yi = run_dev(lambda xd: p.exec_torch(xd, direction=1, normalize=True), x)
    err = rel_l2(yi, oracle_ref(x, 1, True))
    assert err <= tol(n, dt), (p.describe(), "backward", err)


def _to_h(z):
    return np.stack([z.real.astype(np.float16), z.imag.astype(np.float16)], axis=-1)


def _from_h(h):
    return h[..., 0].astype(np.float32) + 1j * h[..., 1].astype(np.float32)


@pytest.mark.parametrize("n,dt", CASES[::4])
def test_sweep_layouts_match_oracle(gpu_ready, n, dt):
    """In place, padded row distances (only the n elements of each output row are written),
    and ComplexF16 storage (relative L2 2e-3 against the oracle of the rounded input)."""
    import torch

    import catabrafft as C

    rng = np.random.Generator(np.random.PCG64(n + 5))
    x = (rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))).astype(dt)
    ref = oracle_ref(x)
    s = torch.cuda.current_stream().cuda_stream
    print(f"layouts n={n} {np.dtype(dt)}", flush=True)
    # in place
    p = C.Plan(n, 2, dt)
    xd = torch.from_numpy(x).cuda()
    p.exec_ptr(xd.data_ptr(), xd.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    assert rel_l2(xd.cpu().numpy(), ref) <= tol(n, dt), ("in place", p.describe())
    # padded rows: idist = n + 3, odist = n + 5
    q = C.Plan(n, 2, dt, idist=n + 3, odist=n + 5)
    xp = np.zeros((2, n + 3), dtype=dt)
    xp[:, :n] = x
    xp[:, n:] = 9.0
    yp = np.full((2, n + 5), 7.0 - 7.0j, dtype=dt)
    xpd, ypd = torch.from_numpy(xp).cuda(), torch.from_numpy(yp).cuda()
    q.exec_ptr(xpd.data_ptr(), ypd.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    yp = ypd.cpu().numpy()
    assert rel_l2(yp[:, :n], ref) <= tol(n, dt), ("padded", q.describe())
    assert np.all(yp[:, n:] == 7.0 - 7.0j), ("padding written", q.describe())
    # ComplexF16 storage
    xh = _to_h(x.astype(np.complex64))
    h = C.Plan(n, 2, 2)
    xhd = torch.from_numpy(xh).cuda()
    yhd = torch.empty_like(xhd)
    h.exec_ptr(xhd.data_ptr(), yhd.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    refh = oracle_ref(_from_h(xh).astype(np.complex128))
    assert rel_l2(_from_h(yhd.cpu().numpy()), refh) < 2e-3, ("f16", h.describe())


@pytest.mark.parametrize("n,dt", [(n, dt) for n, dt in CASES[1::3] if n > 1])
def test_sweep_real_matches_oracle(gpu_ready, n, dt):
    """rfft against the oracle's complex transform of the real input; irfft and brfft / n back
    to the input (even and odd lengths)."""
    import catabrafft as C

    rt = np.float32 if dt == np.complex64 else np.float64
    rng = np.random.Generator(np.random.PCG64(n + 9))
    x = rng.standard_normal((2, n)).astype(rt)
    print(f"real n={n} {np.dtype(rt)}", flush=True)
    X = C.rfft(x)
    assert X.shape == (2, n // 2 + 1)
    ref = oracle_ref(x.astype(np.complex128))[:, :n // 2 + 1]
    t = tol(n, dt)
    assert rel_l2(X, ref) <= t
    assert rel_l2(C.irfft(X, n), x) <= 2 * t
    assert rel_l2(C.brfft(X, n) / n, x) <= 2 * t


def _nd_cases():
    rng = np.random.Generator(np.random.PCG64(77))
    out = []
    while len(out) < 24:
        rank = int(rng.integers(2, 4))
        shape = []
        for _ in range(rank):
            kind = rng.integers(0, 3)
            if kind == 0:
                shape.append(_smooth(rng, int(rng.integers(2, 600)), 4096))
            elif kind == 1:
                shape.append(int(rng.integers(1, 700)))
            else:
                shape.append(int(np.exp(rng.uniform(np.log(2), np.log(40000)))))
        if np.prod(shape) > (1 << 21):
            continue
        axes = tuple(a for a in range(rank) if rng.random() < 0.7) or (int(rng.integers(0, rank)),)
        out.append((tuple(shape), axes, np.complex64 if len(out) % 2 == 0 else np.complex128))
    return out


@pytest.mark.parametrize("shape,axes,dt", _nd_cases())
def test_sweep_nd_matches_oracle(gpu_ready, shape, axes, dt):
    """fftn / ifftn over seeded shapes and axis sets against the oracle along each axis."""
    import catabrafft as C
    from test_gpu_nd import oracle_nd

    rng = np.random.Generator(np.random.PCG64(int(np.prod(shape))))
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
    total = int(np.prod([shape[a] for a in axes]))
    t = 1e-6 if dt == np.complex64 else 4 * np.finfo(float).eps * max(1.0, np.log2(total))
    print(f"nd shape={shape} axes={axes} {np.dtype(dt)}", flush=True)
    y = C.fftn(x, axes)
    assert rel_l2(y, oracle_nd(x, axes)) <= t
    assert rel_l2(C.ifftn(y, axes), x) <= 2 * t


@pytest.mark.parametrize("seed,dt", [(5, np.complex64), (6, np.complex128), (7, np.complex64)])
def test_sweep_uneven_matches_oracle(gpu_ready, seed, dt):
    """Seeded uneven batches (the cost-sorted work lists of every plan kind in one launch
    sequence): each transform against the oracle, and bit-identical to its single plan."""
    import torch

    import catabrafft as C
    from test_gpu_uneven_server import layout, mixed_lengths, oracle_rows

    lengths = mixed_lengths(150, seed=seed)
    offs, span = layout(lengths)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = (rng.standard_normal(span) + 1j * rng.standard_normal(span)).astype(dt)
    p = C.UnevenPlan(lengths, offs, dt)
    print(f"uneven seed={seed} {np.dtype(dt)}: {p.describe()}", flush=True)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    ref = oracle_rows(x, lengths, offs, -1, False)
    for n, o in zip(lengths, offs):
        assert rel_l2(y[o:o + n], ref[o:o + n]) <= tol(n, dt), n
    for n, o in list(zip(lengths, offs))[::11]:
        alone = C.fft(xd[o:o + n].clone()).cpu().numpy()
        assert np.array_equal(alone, y[o:o + n]), n
