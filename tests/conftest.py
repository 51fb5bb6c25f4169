"""Shared test setup: import paths, the `gpu` marker and common helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "catabrafft.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def seeded_complex(n, batch=None, seed=0, dtype=np.complex128, dist="normal"):
    """re, im i.i.d. N(0,1) (or U(-1,1)) from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (n,) if batch is None else (batch, n)
    if dist == "normal":
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    else:
        x = rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)
    return x.astype(dtype)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def golden():
    return np.load(os.path.join(GOLDEN, "dft_golden.npz"), allow_pickle=False)


@pytest.fixture(scope="session")
def gpu_ready():
    """Skips when no GPU is visible (the CPU suite deselects gpu tests with -m 'not gpu')."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    import catabrafft

    catabrafft.context(0)
    return torch
