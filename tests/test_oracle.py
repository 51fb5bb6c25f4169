"""Pins the CPU oracle (the C restatement of CatabraFFT.jl's algorithm) before it is trusted.

Each check is independent of the oracle's own formulas: mpmath golden vectors, a naive
long-double DFT, closed forms and invariants of the DFT, sympy's primitive roots, and the plan
trees read off the reference's code (kernel.jl:72-191, radix_plan.jl:100-138).
"""
import numpy as np
import pytest
import sympy

import oracle as O
from conftest import golden, rel_l2, seeded_complex

EPS64 = np.finfo(np.float64).eps
EPSL = float(np.finfo(np.longdouble).eps)


def _golden_ld(g, n):
    hi, lo = g[f"y_hi_{n}"], g[f"y_lo_{n}"]
    out = np.empty((n, 2), dtype=np.longdouble)
    out[:, 0] = hi.real.astype(np.longdouble) + lo.real.astype(np.longdouble)
    out[:, 1] = hi.imag.astype(np.longdouble) + lo.imag.astype(np.longdouble)
    return out


def _rel_ld(a, b):
    a = np.asarray(a, dtype=np.longdouble)
    b = np.asarray(b, dtype=np.longdouble)
    return float(np.sqrt(((a - b) ** 2).sum()) / np.sqrt((b ** 2).sum()))


# ----------------------------------------------------------------- golden vectors (mpmath)
@pytest.mark.parametrize("n", list(range(1, 17)) + [17, 60, 257, 1024, 2520])
def test_restatement_long_double_matches_mpmath(n):
    g = golden()
    ref = _golden_ld(g, n)
    got = O.fft(g[f"x_{n}"], "l")
    assert _rel_ld(got, ref) < 64 * EPSL * max(1.0, np.log2(n))


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 17, 60, 257, 1024, 2520])
def test_naive_dft_matches_mpmath(n):
    g = golden()
    assert _rel_ld(O.dft_l_pairs(g[f"x_{n}"]), _golden_ld(g, n)) < 64 * EPSL * max(1, np.log2(n))


@pytest.mark.parametrize("n", [2, 16, 17, 60, 257, 1024, 2520])
@pytest.mark.parametrize("prec,bound", [("d", 4 * EPS64), ("f", 1e-6)])
def test_restatement_fp64_fp32_within_bound(n, prec, bound):
    g = golden()
    ref = g[f"y_hi_{n}"] + g[f"y_lo_{n}"]
    x = g[f"x_{n}"]
    if prec == "f":
        x = x.astype(np.complex64)
        ref = O.dft_l(x.astype(np.complex128))
        assert rel_l2(O.fft(x, "f"), ref) < bound
    else:
        assert rel_l2(O.fft(x, "d"), ref) < bound * np.log2(n)


# ----------------------------------------------------------------- a planted mistake is seen
def test_checks_detect_planted_mistakes():
    """A conjugated, index-shifted, transposed or dropped-term result fails the golden check."""
    g = golden()
    n = 60
    ref = g[f"y_hi_{n}"]
    good = O.fft(g[f"x_{n}"], "d")
    assert rel_l2(good, ref) < 1e-14
    wrong = {
        "conjugated": np.conj(good),
        "shifted": np.roll(good, 1),
        "transposed": good.reshape(4, 15).T.ravel(),
        "dropped_term": good - g[f"x_{n}"][7] * np.exp(-2j * np.pi * 7 * np.arange(n) / n),
        "sign_flip": good * np.where(np.arange(n) % 2, -1, 1),
    }
    for name, w in wrong.items():
        assert rel_l2(w, ref) > 1e-3, name


# ----------------------------------------------------------------- closed forms / invariants
SIZES_ALL_STRATEGIES = [1, 2, 8, 27, 125, 343, 1024, 60, 97, 257, 1721, 2520, 4099, 10000]


@pytest.mark.parametrize("n", SIZES_ALL_STRATEGIES)
def test_delta_and_constant(n):
    d0 = np.zeros(n, complex)
    d0[0] = 1
    assert np.allclose(O.fft(d0, "d"), np.ones(n), atol=1e-13)
    j = 3 % n
    dj = np.zeros(n, complex)
    dj[j] = 1
    assert np.allclose(O.fft(dj, "d"), np.exp(-2j * np.pi * j * np.arange(n) / n), atol=1e-12)
    c = O.fft(np.ones(n, complex), "d")
    expect = np.zeros(n, complex)
    expect[0] = n
    assert np.allclose(c, expect, atol=1e-10 * n)


@pytest.mark.parametrize("n", SIZES_ALL_STRATEGIES)
def test_parseval_linearity_shift(n):
    x = seeded_complex(n, seed=11)
    y = seeded_complex(n, seed=12)
    X, Y = O.fft(x, "d"), O.fft(y, "d")
    assert np.isclose(np.vdot(X, X).real, n * np.vdot(x, x).real, rtol=1e-12)
    a, b = 0.3 - 1.1j, -2.0 + 0.5j
    assert rel_l2(O.fft(a * x + b * y, "d"), a * X + b * Y) < 1e-13
    m = 5 % n
    assert rel_l2(O.fft(np.roll(x, m), "d"), X * np.exp(-2j * np.pi * m * np.arange(n) / n)) < 1e-12


@pytest.mark.parametrize("n", [89, 121, 1024, 2048, 3 ** 7, 5 ** 6, 10 ** 4])
@pytest.mark.parametrize("prec", ["d", "f"])
def test_round_trip(n, prec):
    """runtests.jl:39-46: x ~ ifft(fft(x))."""
    dt = np.complex128 if prec == "d" else np.complex64
    x = seeded_complex(n, seed=5, dtype=dt)
    back = O.fft(O.fft(x, prec), prec, direction=1, normalize=True)
    assert rel_l2(back, x) < (1e-13 if prec == "d" else 2e-6)


def test_bfft_is_unnormalised_ifft():
    x = seeded_complex(60, seed=3)
    b = O.fft(x, "d", direction=1, normalize=False)
    i = O.fft(x, "d", direction=1, normalize=True)
    assert rel_l2(b / 60, i) < 1e-15
    assert rel_l2(b, np.conj(O.fft(np.conj(x), "d"))) < 1e-15


# ----------------------------------------------------------------- planner integers
def test_factor_choice_matches_reference_reading():
    """kernel.jl:146-191 by hand: largest 2/3/5/7 power first, else divisor below sqrt."""
    assert O.closest_factors(2520) == (9, 280)
    assert O.closest_factors(280) == (8, 35)
    assert O.closest_factors(35) == (7, 5)
    assert O.closest_factors(10 ** 7) == (78125, 128)
    assert O.closest_factors(341) == (11, 31)
    assert O.closest_factors(1048572) == (27, 38836)
    assert O.closest_factors(17) == (1, 17)


def test_greedy_radix_plans():
    """radix_plan.jl:100-138 with [8,4,2] / [9,3] / [5] / [7] (kernel.jl:131-139)."""
    assert O.std_radix_plan(1024) == [8, 8, 8, 2]
    assert O.std_radix_plan(2 ** 20) == [8] * 6 + [4]
    assert O.std_radix_plan(2 ** 16) == [8] * 5 + [2]
    assert O.std_radix_plan(256) == [8, 8, 4]
    assert O.std_radix_plan(16) == [8, 2]
    assert O.std_radix_plan(3 ** 5) == [9, 9, 3]
    assert O.std_radix_plan(5 ** 3) == [5, 5, 5]
    assert O.std_radix_plan(7 ** 2) == [7, 7]
    assert O.std_radix_plan(60) == []


def test_plan_trees_of_the_configs():
    t = O.plan_tree(2520)
    assert t[:4] == ("mixed", 2520, 9, 280)
    assert t[4] == ("radix", 9, [9])
    assert t[5][:4] == ("mixed", 280, 8, 35)
    assert t[5][5][:4] == ("mixed", 35, 7, 5)
    assert O.plan_tree(17) == ("rader", 17, 3, ("radix", 16, [8, 2]))
    t = O.plan_tree(4099)
    assert t[:3] == ("rader", 4099, 2)
    assert t[3][:4] == ("mixed", 4098, 3, 1366)
    t = O.plan_tree(1048573)
    assert t[:3] == ("rader", 1048573, 2)
    assert t[3][:4] == ("mixed", 1048572, 27, 38836)


@pytest.mark.parametrize("p", [11, 13, 17, 19, 43, 89, 257, 1721, 4099, 65537, 1048573])
def test_primitive_roots_match_sympy(p):
    assert O.primitive_root(p) == sympy.primitive_root(p)


# ----------------------------------------------------------------- codelets and one pass
@pytest.mark.parametrize("R", [2, 3, 4, 5, 7, 8, 9, 16])
@pytest.mark.parametrize("prec", ["d", "l"])
def test_codelets_brute_force(R, prec):
    v = seeded_complex(R, seed=R)
    ref = O.dft_l(v)
    got = O.codelet(v, R, prec)
    if prec == "l":
        got = got[:, 0].astype(float) + 1j * got[:, 1].astype(float)
    assert rel_l2(got, ref) < 1e-15


@pytest.mark.parametrize("n,s,R", [(16, 1, 8), (16, 8, 2), (72, 1, 9), (72, 9, 8), (35, 1, 7),
                                   (35, 7, 5), (1024, 64, 8)])
def test_one_stockham_pass_brute_force(n, s, R):
    """out[q + s(Rp + k)] = w_g^{pk} sum_j in[q + s(p + j n1)] w_R^{jk} (layered codelets)."""
    x = seeded_complex(n, seed=n + s)
    g, n1 = n // s, n // s // R
    ref = np.zeros(n, complex)
    for p in range(n1):
        for q in range(s):
            for k in range(R):
                acc = sum(x[q + s * (p + j * n1)] * np.exp(-2j * np.pi * j * k / R) for j in range(R))
                ref[q + s * (R * p + k)] = acc * np.exp(-2j * np.pi * p * k / g)
    assert rel_l2(O.stockham_pass(x, s, R, "d"), ref) < 1e-14


def test_batched_rows_are_independent():
    x = seeded_complex(2520, batch=5, seed=9)
    y = O.fft(x, "d", threads=2)
    for b in range(5):
        assert rel_l2(y[b], O.fft(x[b], "d")) == 0.0
