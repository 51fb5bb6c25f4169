#!/usr/bin/env python3
"""Generator of the radix-R butterflies ("codelets") used by the HIP kernels.

Emits csrc/codelets.h: one straight-line, in-place, forward DFT_R device function per radix,
for R in RADICES.  This plays the role of CatabraFFT.jl's offline codelet generators
(radix_factory.jl:129-172 `recfft2`, symbfftkernel.jl:17-60 `recfft`), but for any radix:

  * R prime (3, 5, 7, 11, 13): conjugate-pair form, the factoring the hand-written
    radix-5/7 codelets use (radix_5_codelets.jl:27-44, radix_7_codelets.jl:30-51):
        X_0 = v_0 + sum_j (v_j + v_{R-j})
        X_k, X_{R-k} = A_k -/+ i B_k,  A_k = v_0 + sum_j cos(2 pi jk/R)(v_j + v_{R-j}),
                                       B_k = sum_j sin(2 pi jk/R)(v_j - v_{R-j})
  * R composite: Cooley-Tukey R = R1 * R2 (DFT_R1 on R2 strided sub-vectors, internal
    twiddles w_R^{j2 k1}, DFT_R2), recursively; twiddles that are +-1, +-i or (+-1 +- i)/sqrt 2
    become sign swaps / one multiply, as in radix_2_codelets.jl:470-488.

Constants come from mpmath at 40 digits (not copied literals).  Before writing, every codelet
is executed by a small interpreter of the emitted statement list and compared with the DFT
definition; the header is only written when all match to 1e-13.

Usage: python gen_codelets.py [output_path]
"""
from __future__ import annotations

import os
import sys
from fractions import Fraction

import mpmath
import numpy as np

mpmath.mp.dps = 40

RADICES = [2, 3, 4, 5, 7, 8, 9, 11, 13, 16]


def _const(x) -> str:
    s = mpmath.nstr(mpmath.mpf(x), 36, min_fixed=-5, max_fixed=5)
    if "." not in s and "e" not in s:
        s += ".0"
    return s


def cos_sin(m: int, R: int):
    """cos and sin of 2 pi m / R to 40 digits."""
    a = 2 * mpmath.pi * mpmath.mpf(m) / R
    return mpmath.cos(a), mpmath.sin(a)


class Prog:
    """SSA statement list over complex temporaries.

    Statements: (dst, op, args) with ops
      add(a,b) sub(a,b) mul_mi(a) (times -i) mul_pi(a) (times +i) neg(a)
      cmul(a, cr, ci)  (times the constant cr + i ci)
      lin(a0, [(c, a_j)...], imag_only)   a0 + sum c a_j  (real coefficients; a0 may be None)
    """

    def __init__(self, R: int):
        self.R = R
        self.stmts = []
        self.n = 0

    def new(self, op, *args):
        name = f"t{self.n}"
        self.n += 1
        self.stmts.append((name, op, args))
        return name

    # complex helpers -------------------------------------------------------------
    def add(self, a, b):
        return self.new("add", a, b)

    def sub(self, a, b):
        return self.new("sub", a, b)

    def twiddle(self, a, m: int, R: int):
        """a * w_R^m, w_R = e^{-2 pi i / R}, with the trivial cases simplified."""
        f = Fraction(m % R, R)
        if f == 0:
            return a
        if f == Fraction(1, 2):
            return self.new("neg", a)
        if f == Fraction(1, 4):
            return self.new("mul_mi", a)
        if f == Fraction(3, 4):
            return self.new("mul_pi", a)
        c, s = cos_sin(f.numerator, f.denominator)
        # w^m = cos - i sin
        if f.denominator == 8:
            return self.new("mul_diag", a, _const(c), _const(-s))
        return self.new("cmul", a, _const(c), _const(-s))


def gen_dft(P: Prog, vals: list[str]) -> list[str]:
    R = len(vals)
    if R == 1:
        return list(vals)
    if R == 2:
        return [P.add(vals[0], vals[1]), P.sub(vals[0], vals[1])]
    if all(R % d for d in range(2, int(R ** 0.5) + 1)):  # prime
        return gen_prime(P, vals)
    # composite: choose R1 as the smallest factor for R = 4,8,16 -> 4 x 4 for 16, 2 x 4 for 8
    if R == 16:
        R1 = 4
    elif R == 8:
        R1 = 2
    else:
        R1 = min(d for d in range(2, R) if R % d == 0)
    R2 = R // R1
    # A[j2][k1] = DFT_R1 over j1 of x[R2 j1 + j2]
    A = []
    for j2 in range(R2):
        sub = gen_dft(P, [vals[R2 * j1 + j2] for j1 in range(R1)])
        A.append([P.twiddle(sub[k1], j2 * k1, R) for k1 in range(R1)])
    out = [None] * R
    for k1 in range(R1):
        col = gen_dft(P, [A[j2][k1] for j2 in range(R2)])
        for k2 in range(R2):
            out[k1 + R1 * k2] = col[k2]
    return out


def gen_prime(P: Prog, v: list[str]) -> list[str]:
    R = len(v)
    h = (R - 1) // 2
    sp = [None] + [P.add(v[j], v[R - j]) for j in range(1, h + 1)]
    sm = [None] + [P.sub(v[j], v[R - j]) for j in range(1, h + 1)]
    out = [None] * R
    out[0] = P.new("lin", v[0], [("1.0", sp[j]) for j in range(1, h + 1)])
    for k in range(1, h + 1):
        A = P.new("lin", v[0], [(_const(cos_sin(j * k, R)[0]), sp[j]) for j in range(1, h + 1)])
        B = P.new("lin", None, [(_const(cos_sin(j * k, R)[1]), sm[j]) for j in range(1, h + 1)])
        iB = P.new("mul_mi", B)            # -i B
        out[k] = P.add(A, iB)              # A - i B
        out[R - k] = P.sub(A, iB)          # A + i B
    return out


def build(R: int):
    P = Prog(R)
    ins = [f"v[{i}]" for i in range(R)]
    outs = gen_dft(P, ins)
    return P, outs


# ---------------------------------------------------------------------- self-check
def interpret(P: Prog, outs, x: np.ndarray) -> np.ndarray:
    env = {f"v[{i}]": complex(x[i]) for i in range(len(x))}
    for dst, op, args in P.stmts:
        if op == "add":
            env[dst] = env[args[0]] + env[args[1]]
        elif op == "sub":
            env[dst] = env[args[0]] - env[args[1]]
        elif op == "neg":
            env[dst] = -env[args[0]]
        elif op == "mul_mi":
            env[dst] = -1j * env[args[0]]
        elif op == "mul_pi":
            env[dst] = 1j * env[args[0]]
        elif op in ("cmul", "mul_diag"):
            env[dst] = env[args[0]] * complex(float(args[1]), float(args[2]))
        elif op == "lin":
            acc = env[args[0]] if args[0] is not None else 0j
            for c, a in args[1]:
                acc += float(c) * env[a]
            env[dst] = acc
        else:
            raise ValueError(op)
    return np.array([env[o] for o in outs])


def self_check():
    rng = np.random.default_rng(12345)
    for R in RADICES:
        P, outs = build(R)
        for _ in range(3):
            x = rng.standard_normal(R) + 1j * rng.standard_normal(R)
            ref = np.array([sum(x[j] * np.exp(-2j * np.pi * j * k / R) for j in range(R))
                            for k in range(R)])
            got = interpret(P, outs, x)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < 1e-13, (R, err)


# ---------------------------------------------------------------------- emission
def emit(R: int) -> str:
    P, outs = build(R)
    L = []
    L.append(f"// DFT_{R}, forward (w = e^{{-2 pi i/{R}}}), in place on v[0..{R - 1}].")
    L.append(f"template <typename V>\n__device__ __forceinline__ void dft{R}(V* v) {{")
    L.append("    using T = decltype(v[0].x);  // scalar type for the constants")
    for dst, op, args in P.stmts:
        if op == "add":
            a, b = args
            L.append(f"    const V {dst} = {{{a}.x + {b}.x, {a}.y + {b}.y}};")
        elif op == "sub":
            a, b = args
            L.append(f"    const V {dst} = {{{a}.x - {b}.x, {a}.y - {b}.y}};")
        elif op == "neg":
            a = args[0]
            L.append(f"    const V {dst} = {{-{a}.x, -{a}.y}};")
        elif op == "mul_mi":
            a = args[0]
            L.append(f"    const V {dst} = {{{a}.y, -{a}.x}};")
        elif op == "mul_pi":
            a = args[0]
            L.append(f"    const V {dst} = {{-{a}.y, {a}.x}};")
        elif op == "mul_diag":
            # (a.x + i a.y)(cr + i ci) with |cr| = |ci| = 1/sqrt 2
            a, cr, ci = args
            sgn = "+" if float(cr) * float(ci) > 0 else "-"
            r = _const(abs(mpmath.mpf(cr)))
            sr = "" if float(cr) > 0 else "-"
            L.append(f"    const V {dst} = {{T({sr}{r}) * ({a}.x {'-' if sgn == '+' else '+'} {a}.y), "
                     f"T({sr}{r}) * ({a}.y {sgn} {a}.x)}};")
        elif op == "cmul":
            # fused: re = fma(a.x, cr, -(a.y ci)), im = fma(a.x, ci, a.y cr)
            a, cr, ci = args
            L.append(f"    const V {dst} = {{fmadd({a}.x, T({cr}), -({a}.y * T({ci}))), "
                     f"fmadd({a}.x, T({ci}), {a}.y * T({cr}))}};")
        elif op == "lin":
            # a0 + c1 a1 + c2 a2 + ..., left to right, each c_j a_j + acc one fused multiply-add
            a0, terms = args
            comps = []
            for f in ("x", "y"):
                acc = None if a0 is None else f"{a0}.{f}"
                for c, a in terms:
                    if c == "1.0":
                        acc = f"{a}.{f}" if acc is None else f"{acc} + {a}.{f}"
                    elif acc is None:
                        acc = f"T({c}) * {a}.{f}"
                    else:
                        acc = f"fmadd(T({c}), {a}.{f}, {acc})"
                comps.append(acc)
            L.append(f"    const V {dst} = {{{comps[0]}, {comps[1]}}};")
    for k, o in enumerate(outs):
        L.append(f"    const V o{k} = {o};")
    for k in range(R):
        L.append(f"    v[{k}] = o{k};")
    L.append("}")
    if not any("T(" in l for l in L[3:]):
        L.pop(2)  # no constants: drop the scalar alias
    return "\n".join(L)


def emit_lane_w16() -> str:
    """w_16^m = e^{-2 pi i m / 16} for a lane-dependent m (0..15): the stage twiddles of the
    cross-lane radix-2 stages (engine.h, wave_tail_pass), by selects over three constants."""
    c, s = cos_sin(1, 16)
    r = mpmath.sqrt(mpmath.mpf(2)) / 2
    return "\n".join([
        "// w_16^m = e^{-2 pi i m/16} for a lane-dependent m in 0..15: w^(m & 3) from three",
        "// constants by selects, then times (-i)^(m >> 2).",
        "template <typename V>",
        "__device__ __forceinline__ V lane_w16(int m) {",
        "    using T = decltype(V{}.x);",
        f"    const T c = T({_const(c)}), s = T({_const(s)}), r = T({_const(r)});",
        "    const int q = m & 3;",
        "    const T x = q == 0 ? T(1) : (q == 1 ? c : (q == 2 ? r : s));",
        "    const T y = q == 0 ? T(0) : (q == 1 ? -s : (q == 2 ? -r : -c));",
        "    const bool a1 = (m & 4) != 0, a2 = (m & 8) != 0;",
        "    const T x1 = a1 ? y : x, y1 = a1 ? -x : y;",
        "    return V{a2 ? -x1 : x1, a2 ? -y1 : y1};",
        "}",
    ])


def check_lane_w16():
    """The select formula of lane_w16 equals e^{-2 pi i m / 16} for every m."""
    c, s = (float(v) for v in cos_sin(1, 16))
    r = float(mpmath.sqrt(2) / 2)
    for m in range(16):
        q = m & 3
        x = [1.0, c, r, s][q]
        y = [0.0, -s, -r, -c][q]
        x1, y1 = (y, -x) if m & 4 else (x, y)
        got = complex(-x1, -y1) if m & 8 else complex(x1, y1)
        want = complex(mpmath.cos(2 * mpmath.pi * m / 16), -mpmath.sin(2 * mpmath.pi * m / 16))
        assert abs(got - want) < 1e-15, (m, got, want)


def check_mul_diag():
    """The emitted mul_diag formula equals a * (cr + i ci) for the four diagonal constants."""
    for cr, ci in [(1, 1), (1, -1), (-1, 1), (-1, -1)]:
        a = complex(0.3, -0.7)
        c = complex(cr, ci) / np.sqrt(2)
        sgn = "+" if cr * ci > 0 else "-"
        s = 1 if cr > 0 else -1
        r = 1 / np.sqrt(2)
        if sgn == "+":
            got = complex(s * r * (a.real - a.imag), s * r * (a.imag + a.real))
        else:
            got = complex(s * r * (a.real + a.imag), s * r * (a.imag - a.real))
        assert abs(got - a * c) < 1e-15, (cr, ci)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "..", "csrc", "codelets.h")
    self_check()
    check_mul_diag()
    check_lane_w16()
    parts = [
        "// codelets.h -- GENERATED by gen/gen_codelets.py; do not edit by hand.",
        "// Straight-line forward DFT_R butterflies for R in {" + ", ".join(map(str, RADICES)) + "}.",
        "// V is a two-component vector type (float2 / double2) holding (re, im).",
        "#pragma once",
        "#include <hip/hip_runtime.h>",
        "namespace cfft {",
        "// a * b + c rounded once (v_fma_f32 / v_fma_f64; fma() on the host, also correctly",
        "// rounded, so a CPU build of these butterflies gives the same bits).",
        "__host__ __device__ __forceinline__ float fmadd(float a, float b, float c) {",
        "    return __builtin_fmaf(a, b, c);",
        "}",
        "__host__ __device__ __forceinline__ double fmadd(double a, double b, double c) {",
        "    return __builtin_fma(a, b, c);",
        "}",
        "",
    ]
    for R in RADICES:
        parts.append(emit(R))
        parts.append("")
    parts.append("template <int R> struct Dft;")
    for R in RADICES:
        parts.append(f"template <> struct Dft<{R}> {{ template <typename V> "
                     f"__device__ __forceinline__ static void run(V* v) {{ dft{R}(v); }} }};")
    parts.append("constexpr int kRadices[] = {" + ", ".join(map(str, RADICES)) + "};")
    parts.append("")
    parts.append(emit_lane_w16())
    parts.append("}  // namespace cfft")
    with open(out, "w") as f:
        f.write("\n".join(parts) + "\n")
    print(f"wrote {os.path.normpath(out)}")


if __name__ == "__main__":
    main()
