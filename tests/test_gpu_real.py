"""GPU parity of the real-input transforms (rfft / brfft / irfft) against the oracle."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [16, 1024, 2520, 17, 99, 1 << 16, 2 * 4099, 1 << 20, 10 ** 5])
@pytest.mark.parametrize("rt", [np.float32, np.float64])
def test_rfft_irfft(gpu_ready, n, rt):
    import catabrafft as C

    rng = np.random.Generator(np.random.PCG64(n))
    x = rng.standard_normal((3, n)).astype(rt)
    X = C.rfft(x)
    ref = O.fft(x.astype(np.complex128), "d", threads=8)[:, :n // 2 + 1]
    tol = 1e-6 if rt == np.float32 else 4 * np.finfo(float).eps * np.log2(n)
    assert X.shape == (3, n // 2 + 1)
    assert rel_l2(X, ref) < tol
    assert rel_l2(C.irfft(X, n), x) < 2 * tol
    assert rel_l2(C.brfft(X, n) / n, x) < 2 * tol


def test_rfft_device_tensors(gpu_ready):
    import torch

    import catabrafft as C

    x = torch.randn(4, 2520, device="cuda")
    X = C.rfft(x)
    ref = O.fft(x.cpu().numpy().astype(np.complex128), "d")[:, :1261]
    assert rel_l2(X.cpu().numpy(), ref) < 1e-6
    back = C.irfft(X, 2520)
    torch.cuda.synchronize()
    assert rel_l2(back.cpu().numpy(), x.cpu().numpy()) < 2e-6
