"""GPU parity of multi-dimensional transforms (cfft_plan_nd) against the oracle applied along
each transformed axis."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2, seeded_complex

pytestmark = pytest.mark.gpu


def oracle_nd(x, axes, direction=-1, normalize=False):
    y = x.astype(np.complex128)
    for a in axes:
        y = np.moveaxis(y, a, -1)
        shp = y.shape
        y = O.fft(np.ascontiguousarray(y.reshape(-1, shp[-1])), "d", direction, normalize,
                  threads=8).reshape(shp)
        y = np.moveaxis(y, -1, a)
    return y


@pytest.mark.parametrize("shape,axes", [((64, 64), (0, 1)), ((256, 2520), (0, 1)),
                                        ((1000, 60), (0,)), ((60, 1000), (1,)),
                                        ((16, 32, 64), (0, 1, 2)), ((8, 12, 97), (0, 2)),
                                        ((128, 1 << 14), (0, 1)), ((4096, 17), (0, 1))])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_nd_matches_oracle(gpu_ready, shape, axes, dt):
    import catabrafft as C

    x = seeded_complex(int(np.prod(shape)), seed=11, dtype=dt).reshape(shape)
    tol = 1e-6 if dt == np.complex64 else 64 * np.finfo(float).eps
    y = C.fftn(x, axes)
    assert rel_l2(y, oracle_nd(x, axes)) < tol
    back = C.ifftn(y, axes)
    assert rel_l2(back, x) < 2 * tol


@pytest.mark.parametrize("shape,axes", [((17, 1000), (0, 1)), ((1009, 64), (0,)),
                                        ((20000, 8), (0, 1)), ((3, 40000, 2), (1,)),
                                        ((5, 17, 19, 8), (0, 1, 2, 3))])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
def test_nd_transposed_axes_match_oracle(gpu_ready, shape, axes, dt):
    """Strided axes one workgroup cannot take (not smooth, or longer than a workgroup's line)
    run as transpose, batched 1-D plan of any kind, transpose back."""
    import catabrafft as C

    x = seeded_complex(int(np.prod(shape)), seed=13, dtype=dt).reshape(shape)
    tol = 2e-6 if dt == np.complex64 else 64 * np.finfo(float).eps
    p = C.NDPlan(shape, axes, dt)
    assert "transposed" in p.describe(), p.describe()
    y = C.fftn(x, axes)
    assert rel_l2(y, oracle_nd(x, axes)) < tol
    back = C.ifftn(y, axes)
    assert rel_l2(back, x) < 2 * tol
