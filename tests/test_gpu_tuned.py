"""Tuned plans against the oracle: MEASURE / ENCHANT plans, and every family of configuration
the autotuner can pick, forced through wisdom.

The reference's MEASURE / ENCHANT planner (radix_plan.jl:23-68, radix_exec.jl:140-297) times
candidate radix decompositions and keeps the fastest; each candidate computes the same DFT.
Here the candidates are the radix sequences of each line and the kernel configurations of one
transform (threads per line, per-thread capacity E, the four-step split, its tile widths,
chunking, the pipelined kernels, the six-step path), so whichever one the timing picks must match the oracle to the same tolerance as the
default plan:  ComplexF64 4 * eps64 * log2(n), ComplexF32 1e-6 (relative L2).
"""
import numpy as np
import pytest

from conftest import rel_l2, seeded_complex
from test_gpu_parity import oracle_ref, run_dev, tol

pytestmark = pytest.mark.gpu

# wisdom fields after the key: E target_threads n1 ca cb chunk pipe large [ra=...] [rb=...]
# (0 = the planner's default for n1 / ca / cb, chunk -1 = default; ra / rb = radix sequences;
# pm = the four-step passes on the pipelined kernels, eb = pass B's E, pr = line-pair passes,
# pd = 0: direct lines without LDS pad words)
FORCED = [
    # n, batch, dtype, fields, what the description must show
    (1 << 16, 6, np.complex64, "16 64 256 2 32 -1 0 0", ["N1=256", "CA=2", "CB=32"]),
    (1 << 16, 6, np.complex128, "8 64 128 8 4 2 0 0", ["N1=128", "CA=8", "CB=4"]),
    (1 << 16, 5, np.complex64, "16 64 512 4 16 -1 1 0", ["N1=512", "pipe"]),
    (1 << 16, 5, np.complex128, "16 64 0 0 0 -1 1 0", ["fourstep", "pipe"]),
    # one pass on the pipelined kernel, with its own per-thread capacity (pm=, eb=)
    (1 << 20, 3, np.complex128, "8 128 1024 4 8 -1 0 0 pm=2 eb=16", ["pipe=B", "CB=8"]),
    (1 << 20, 3, np.complex64, "16 64 2048 4 16 -1 0 0 pm=1 eb=8 pr=2",
     ["pipe=A", "pair=B"]),
    (1 << 16, 3, np.complex64, "16 64 256 0 0 -1 0 1", ["sixstep"]),
    (2520, 7, np.complex128, "8 128 40 16 16 -1 0 0", ["N1=40", "CA=16", "CB=16"]),
    (2520, 7, np.complex64, "16 64 0 0 0 -1 1 0", ["direct", "pipe"]),
    (1000, 9, np.complex64, "8 512 0 0 0 -1 0 0", ["direct"]),
    # direct lines with unpadded LDS lines (pd=0)
    (2520, 4096, np.complex128, "16 64 0 0 0 -1 0 0 ra=8,7,5,3,3 pd=0", ["Tt=180", "nopad"]),
    (1024, 40, np.complex64, "16 64 0 0 0 -1 0 0 wv=1 pd=0", ["wave", "nopad"]),
    (4096, 16, np.complex128, "8 64 0 0 0 -1 0 0 pd=0", ["direct", "nopad"]),
    # radix sequences other than the default (the reference's MEASURE search space)
    (2520, 64, np.complex64, "8 64 0 0 0 -1 0 0 ra=5,7,8,9", ["direct [5,7,8,9]"]),
    (2520, 64, np.complex128, "8 64 0 0 0 -1 0 0 ra=8,9,7,5", ["direct [8,9,7,5]"]),
    (2520, 33, np.complex64, "16 256 0 0 0 -1 0 0 ra=3,3,8,7,5", ["[3,3,8,7,5]"]),
    (1024, 40, np.complex64, "16 64 0 0 0 -1 0 0 ra=4,16,16", ["direct [4,16,16]"]),
    (1024, 40, np.complex128, "8 64 0 0 0 -1 0 0 ra=8,8,16", ["direct [8,8,16]"]),
    (1024, 40, np.complex64, "16 64 0 0 0 -1 0 0 ra=16,4,16", ["[16,4,16]"]),
    (1 << 20, 2, np.complex64, "16 64 0 0 0 0 0 0 ra=8,8,16 rb=4,16,16",
     ["N1=1024 [8,8,16]", "[4,16,16]"]),
    (1 << 20, 2, np.complex128, "8 64 0 0 0 0 0 0 ra=16,4,16 rb=8,16,8",
     ["N1=1024 [16,4,16]", "[8,16,8]"]),
    (3 ** 12, 2, np.complex64, "16 64 0 0 0 0 0 0 ra=3,3,3,3,9 rb=3,3,9,9",
     ["N1=729 [3,3,3,3,9]", "[3,3,9,9]"]),
]


def _check_plan(p, x, dt):
    n = x.shape[-1]
    for direction, norm in ((-1, False), (1, True)):
        y = run_dev(lambda xd: p.exec_torch(xd, direction=direction, normalize=norm), x)
        err = rel_l2(y, oracle_ref(x, direction, norm))
        assert err <= tol(n, dt), (p.describe(), direction, err)


@pytest.mark.parametrize("n,batch,dt,fields,marks", FORCED)
def test_forced_configuration_matches_oracle(gpu_ready, tmp_path, n, batch, dt, fields, marks):
    import catabrafft as C

    code = 0 if dt == np.complex64 else 1
    w = tmp_path / "wisdom.txt"
    w.write_text(f"{n} {batch} {n} {n} {code} 1 : {fields}\n")
    C.import_wisdom(str(w))
    p = C.Plan(n, batch, dt, flags=C.MEASURE)
    d = p.describe()
    print(f"forced n={n} batch={batch} {np.dtype(dt)}: {d}", flush=True)
    assert "(wisdom)" in d, d
    for m in marks:
        assert m in d, (m, d)
    _check_plan(p, seeded_complex(n, batch, seed=n % 97 + batch, dtype=dt), dt)


@pytest.mark.parametrize("n,batch,dt,flags", [(1000, 8, np.complex64, 2),
                                              (2520, 16, np.complex128, 1),
                                              (2520, 512, np.complex64, 1),
                                              (1024, 256, np.complex64, 2),
                                              (1 << 16, 4, np.complex64, 1),
                                              (4099, 3, np.complex64, 2)])
def test_tuned_plan_matches_oracle(gpu_ready, n, batch, dt, flags):
    import catabrafft as C

    p = C.Plan(n, batch, dt, flags=flags)
    d = p.describe()
    print(f"tuned n={n} batch={batch} {np.dtype(dt)} flags={flags}: {d}", flush=True)
    assert "tuned(" in d or "(wisdom)" in d, d
    _check_plan(p, seeded_complex(n, batch, seed=n % 89 + 3, dtype=dt), dt)


def test_wisdom_round_trip_reproduces_the_tuned_bits(gpu_ready, tmp_path):
    """A tuned choice exported and imported again plans the same kernels: same bits."""
    import catabrafft as C

    n, batch, dt = 720, 6, np.complex64
    x = seeded_complex(n, batch, seed=11, dtype=dt)
    p1 = C.Plan(n, batch, dt, flags=C.ENCHANT)
    y1 = run_dev(lambda xd: p1.exec_torch(xd), x)
    path = tmp_path / "wisdom.txt"
    C.export_wisdom(str(path))
    lines = [l for l in path.read_text().splitlines() if l.startswith(f"{n} {batch} {n} {n} 0 2 :")]
    assert len(lines) == 1, path.read_text()
    C.import_wisdom(str(path))
    p2 = C.Plan(n, batch, dt, flags=C.ENCHANT)
    assert "(wisdom)" in p2.describe(), p2.describe()
    y2 = run_dev(lambda xd: p2.exec_torch(xd), x)
    assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))
    assert rel_l2(y1, oracle_ref(x)) <= tol(n, dt)


@pytest.mark.parametrize("n,batch,dt", [(1 << 20, 16, np.complex64), (1 << 20, 8, np.complex128),
                                        (1 << 18, 32, np.complex64)])
def test_tuned_fourstep_refined_matches_oracle_and_round_trips(gpu_ready, tmp_path, n, batch, dt):
    """MEASURE on four-step lengths: the winner of the candidate list refined over tiles,
    pipelined and (ComplexF32) line-pair passes, and the finalists timed again.  The tuned plan
    matches the oracle, and its wisdom line, exported and imported, plans the same kernels (same
    description, same bits)."""
    import catabrafft as C

    x = seeded_complex(n, batch, seed=n % 83 + batch, dtype=dt)
    p1 = C.Plan(n, batch, dt, flags=C.MEASURE)
    d1 = p1.describe()
    print(f"tuned four-step n={n} batch={batch} {np.dtype(dt)}: {d1}", flush=True)
    assert "fourstep" in d1 and ("tuned(" in d1 or "(wisdom)" in d1), d1
    _check_plan(p1, x, dt)
    y1 = run_dev(lambda xd: p1.exec_torch(xd), x)
    path = tmp_path / "wisdom.txt"
    C.export_wisdom(str(path))
    code = 0 if dt == np.complex64 else 1
    lines = [l for l in path.read_text().splitlines() if l.startswith(f"{n} {batch} {n} {n} {code} 1 :")]
    assert len(lines) == 1, path.read_text()
    C.import_wisdom(str(path))
    p2 = C.Plan(n, batch, dt, flags=C.MEASURE)
    d2 = p2.describe()
    assert "(wisdom)" in d2 and d2.split(" (wisdom)")[0] == d1.split(" tuned(")[0].split(" (wisdom)")[0], (d1, d2)
    y2 = run_dev(lambda xd: p2.exec_torch(xd), x)
    word = np.uint32 if dt == np.complex64 else np.uint64
    assert np.array_equal(y1.view(word), y2.view(word))


@pytest.mark.parametrize("n,batch,dt,fields", [
    (2520, 64, np.complex128, "16 64 0 0 0 -1 0 0 ra=8,7,5,3,3"),
    (1024, 40, np.complex64, "16 64 0 0 0 -1 0 0 wv=1"),
    (4096, 16, np.complex128, "8 64 0 0 0 -1 0 0")])
def test_unpadded_lds_lines_give_the_padded_bits(gpu_ready, tmp_path, n, batch, dt, fields):
    """The LDS layout of a line moves no arithmetic: the plan with pd=0 gives the padded plan's
    bits on the GPU."""
    import catabrafft as C

    code = 0 if dt == np.complex64 else 1
    x = seeded_complex(n, batch, seed=n % 91 + 2, dtype=dt)
    ys = {}
    for pd in (1, 0):
        w = tmp_path / f"w{pd}.txt"
        w.write_text(f"{n} {batch} {n} {n} {code} 1 : {fields} pd={pd}\n")
        C.import_wisdom(str(w))
        p = C.Plan(n, batch, dt, flags=C.MEASURE)
        assert ("nopad" in p.describe()) == (pd == 0), p.describe()
        ys[pd] = run_dev(lambda xd: p.exec_torch(xd), x)
    word = np.uint32 if dt == np.complex64 else np.uint64
    assert np.array_equal(ys[0].view(word), ys[1].view(word))
    assert rel_l2(ys[0], oracle_ref(x)) <= tol(n, dt)
