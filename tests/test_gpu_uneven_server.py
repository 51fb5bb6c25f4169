"""GPU tests of the uneven-batch plan (one work-list launch sequence, hipGraph replays) and of
the continuous-batching server.  Every transform is checked against the oracle, and against
the same transform computed alone (single-length plan): the work-list kernels run the same
arithmetic, so results must be bit-identical in and out of the mix."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, seeded_complex

pytestmark = pytest.mark.gpu


def mixed_lengths(count, seed, lo=64, hi=1 << 16):
    """Log-uniform lengths snapped to 7-smooth (70%), other composite (20%) or prime (10%)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        L = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        u = rng.uniform()
        if u < 0.7:
            out.append(nearest(L, smooth7))
        elif u < 0.9:
            out.append(nearest(L, lambda m: not smooth7(m) and not O.isprime(m)))
        else:
            out.append(nearest(L, O.isprime))
    return out


def smooth7(m):
    for p in (2, 3, 5, 7):
        while m % p == 0:
            m //= p
    return m == 1


def nearest(L, pred):
    for d in range(0, 1 << 20):
        for c in (L - d, L + d):
            if c >= 2 and pred(c):
                return c
    raise ValueError(L)


def oracle_rows(x, lengths, offsets, direction=-1, normalize=False):
    out = np.zeros_like(x, dtype=np.complex128)
    for n, o in zip(lengths, offsets):
        seg = x[o:o + n].astype(np.complex128)
        out[o:o + n] = O.fft(seg, "d", direction, normalize) if n > 4096 else O.fft_l(seg, direction, normalize)
    return out


def layout(lengths, gap=3):
    offs, o = [], 0
    for n in lengths:
        offs.append(o)
        o += n + gap
    return offs, o


@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_uneven_plan_matches_oracle_and_single_plans(gpu_ready, dt):
    import torch

    import catabrafft as C

    lengths = mixed_lengths(120, seed=4) + [1, 2, 17, 4099, 1 << 17, 3 * 65537]
    offs, span = layout(lengths)
    x = seeded_complex(span, seed=4, dtype=dt)
    p = C.UnevenPlan(lengths, offs, dt)
    print("uneven plan:", p.describe(), flush=True)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    stream = torch.cuda.current_stream().cuda_stream
    tol = 1e-6 if dt == np.complex64 else 4 * np.finfo(float).eps * 20
    for direction, normalize in ((-1, False), (1, True), (1, False)):
        p.exec_ptr(xd.data_ptr(), yd.data_ptr(), direction, normalize, stream)
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        ref = oracle_rows(x, lengths, offs, direction, normalize)
        for n, o in zip(lengths, offs):
            assert rel_l2(y[o:o + n], ref[o:o + n]) < tol, (n, direction, normalize)
        # gaps between transforms are not written
        assert np.all(y[offs[0] + lengths[0]:offs[1]] == 0)
    # alone vs in the mix: bit-identical to the single-length plans
    p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False, stream)
    torch.cuda.synchronize()
    y = yd.cpu().numpy()
    for n, o in list(zip(lengths, offs))[::7]:
        alone = C.fft(xd[o:o + n].clone()).cpu().numpy()
        assert np.array_equal(alone, y[o:o + n]), n


def test_uneven_graph_replays_are_stable(gpu_ready):
    """The captured graph is replayed with eager work and syncs in between; every replay
    must equal the first result exactly."""
    import torch

    import catabrafft as C

    lengths = mixed_lengths(60, seed=9, hi=1 << 18)
    offs, span = layout(lengths, gap=0)
    x = seeded_complex(span, seed=9, dtype=np.complex64)
    p = C.UnevenPlan(lengths, offs, np.complex64)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    s = torch.cuda.current_stream().cuda_stream
    p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False, s)
    torch.cuda.synchronize()
    first = yd.cpu().numpy()
    ref = oracle_rows(x, lengths, offs)
    assert rel_l2(first, ref) < 1e-6
    for k in range(4):
        junk = torch.randn(1 << 20, device="cuda") * k  # eager work between replays
        yd.zero_()
        p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False, s)
        torch.cuda.synchronize()
        assert np.array_equal(yd.cpu().numpy(), first), k
        del junk


def test_uneven_in_place(gpu_ready):
    import torch

    import catabrafft as C

    lengths = [64, 1000, 4099, 2520, 1 << 16, 12345, 1 << 18, 7]
    offs, span = layout(lengths, gap=0)
    x = seeded_complex(span, seed=2, dtype=np.complex64)
    p = C.UnevenPlan(lengths, offs, np.complex64)
    xd = torch.from_numpy(x).cuda()
    p.exec_ptr(xd.data_ptr(), xd.data_ptr(), -1, False, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    y = xd.cpu().numpy()
    ref = oracle_rows(x, lengths, offs)
    for n, o in zip(lengths, offs):
        assert rel_l2(y[o:o + n], ref[o:o + n]) < 1e-6, n


def test_uneven_rejects_overlaps():
    import catabrafft as C

    with pytest.raises(ValueError):
        C.UnevenPlan([8, 8], [0, 4], np.complex64)
    with pytest.raises(ValueError):
        C.UnevenPlan([0], [0], np.complex64)


# --------------------------------------------------------------------------------- server
def test_server_mixed_requests_match_alone(gpu_ready):
    import torch

    import catabrafft as C

    srv = C.Server(hbm_cap_bytes=8 << 30, nstreams=3)
    rng = np.random.Generator(np.random.PCG64(5))
    lengths = mixed_lengths(40, seed=5, hi=1 << 17)
    reqs = []
    for i, n in enumerate(lengths):
        batch = int(rng.integers(1, 6))
        dt = np.complex128 if i % 10 == 3 else np.complex64
        x = seeded_complex(n, batch=batch, seed=100 + i, dtype=dt)
        direction = -1 if i % 3 else 1
        if i % 2:
            t = srv.submit(torch.from_numpy(x).cuda(), direction=direction, normalize=direction > 0)
        else:
            t = srv.submit(x, direction=direction, normalize=direction > 0)
        reqs.append((t, x, direction))
    for t, x, direction in reqs:
        y = srv.wait(t)
        y = y.cpu().numpy() if hasattr(y, "cpu") else y
        ref = np.stack([O.fft(r.astype(np.complex128), "d", direction, direction > 0) for r in x])
        tol = 1e-6 if x.dtype == np.complex64 else 1e-13
        assert rel_l2(y, ref) < tol, (x.shape, direction)
        alone = (C.ifft(x) if direction > 0 else C.fft(x))
        assert np.array_equal(alone, y), (x.shape, direction)
    st = srv.stats()
    print("server stats:", st, flush=True)
    assert st["p50_s"] > 0 and st["p99_s"] >= st["p50_s"] and st["transforms_per_s"] > 0
    srv.close()


def test_server_admission_cap(gpu_ready):
    import catabrafft as C

    n = 1 << 14
    x = seeded_complex(n, batch=4, seed=1, dtype=np.complex64)
    per_request = 2 * x.nbytes
    srv = C.Server(hbm_cap_bytes=int(per_request * 1.5), nstreams=2)  # one request at a time
    ts = [srv.submit(x) for _ in range(6)]
    for t in ts:
        assert rel_l2(srv.wait(t), O.fft(x.astype(np.complex128), "d")) < 1e-6
    big = seeded_complex(n, batch=16, seed=2, dtype=np.complex64)
    t = srv.submit(big)
    with pytest.raises(C.CatabraFFTError) as e:
        srv.wait(t)
    assert e.value.code == -5  # CFFT_ERR_ADMISSION
    srv.close()
