#!/usr/bin/env python3
"""Writes dft_golden.npz: seeded inputs and their DFTs computed with mpmath at 30 digits.

The reference repository commits no golden vectors (its tests compare unseeded randn inputs with
FFTW, test/runtests.jl:4-46), so these fixtures are made here, independently of both the oracle
and the product: the DFT definition X_k = sum_j x_j e^{-2 pi i jk/n} evaluated term by term in
multiple precision.  Each output is stored as a double-double (hi + lo) so that the fixtures keep
~30 significant digits.

Sizes: n = 1..16, 17 (Rader), 60 (mixed), 257 (Rader, Fermat), 1024 (radix-2 path, cfg 1),
2520 (mixed radix, cfg 2).  Inputs: re, im ~ N(0, 1) from numpy PCG64(seed = 1000 + n), rounded
to float64 (the values the tests feed to the code under test).

Usage: python make_golden.py   (about a minute)
"""
from __future__ import annotations

import os

import mpmath
import numpy as np

SIZES = list(range(1, 17)) + [17, 60, 257, 1024, 2520]


def dft_mp(x: np.ndarray, dps: int = 30):
    mpmath.mp.dps = dps
    n = len(x)
    xs = [mpmath.mpc(float(v.real), float(v.imag)) for v in x]
    w = [mpmath.expjpi(mpmath.mpf(-2 * m) / n) for m in range(n)]  # e^{-2 pi i m/n}
    out = []
    for k in range(n):
        terms = [w[(j * k) % n] for j in range(n)]
        out.append(mpmath.fdot(xs, terms))
    return out


def split(vals):
    hi = np.array([complex(float(v.real), float(v.imag)) for v in vals])
    lo = np.array([complex(float(v.real - mpmath.mpf(h.real)), float(v.imag - mpmath.mpf(h.imag)))
                   for v, h in zip(vals, hi)])
    return hi, lo


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    data = {}
    for n in SIZES:
        rng = np.random.Generator(np.random.PCG64(1000 + n))
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        hi, lo = split(dft_mp(x))
        data[f"x_{n}"] = x
        data[f"y_hi_{n}"] = hi
        data[f"y_lo_{n}"] = lo
    np.savez_compressed(os.path.join(here, "dft_golden.npz"), sizes=np.array(SIZES), **data)
    print("wrote", os.path.join(here, "dft_golden.npz"))


if __name__ == "__main__":
    main()
