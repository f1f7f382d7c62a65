"""The exact reductions of the int8-limb path at their extremes, on the CPU.

pgw_fold.h holds the one definition of the MFMA epilogue's digit-shift fold,
the Mersenne reduction of the digit split and row mixing (red63_m31) and the
balanced digit split; tests/fold_check.cpp runs that same code on the host.
Each result is compared with exact Python integers for p = 2^31 - 1 (the
Mersenne folds), 2147483629 (the per-shift Barrett path), 65521 and 3 (the
two-digit path): every accumulator at +-(2^31 - 1) and -2^31, mixed signs,
one extreme shift at a time, and the products the digit split and row mixing
can form (below 2^62 and 2^63).
"""
import random
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PRIMES = [2**31 - 1, 2147483629, 65521, 3]
LO, HI = -2**31, 2**31 - 1


@pytest.fixture(scope="module")
def fold_check(tmp_path_factory):
    exe = tmp_path_factory.mktemp("fold") / "fold_check"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-x", "hip", "--cuda-host-only",
                        "-I", str(ROOT / "prym_green_ancillary_amd" / "csrc"),
                        str(ROOT / "tests" / "fold_check.cpp"), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]

    def run(lines):
        out = subprocess.run([str(exe)], input="\n".join(lines) + "\n", capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stderr
        return out.stdout.splitlines()
    return run


def _shift_vectors(ns, rng):
    vecs = [[HI] * ns, [LO] * ns, [-HI] * ns, [0] * ns]
    vecs += [[HI if s % 2 else LO for s in range(ns)], [LO if s % 2 else HI for s in range(ns)]]
    for s in range(ns):
        for v in (HI, LO, 1, -1):
            x = [0] * ns
            x[s] = v
            vecs.append(x)
    vecs += [[rng.choice((LO, HI, rng.randint(LO, HI))) for _ in range(ns)] for _ in range(200)]
    return vecs


@pytest.mark.parametrize("p", PRIMES)
def test_digit_shift_fold_is_exact_at_the_accumulator_bounds(fold_check, p):
    ns = 3 if p <= 65536 else 7
    vecs = _shift_vectors(ns, random.Random(p))
    got = fold_check([f"F {p} " + " ".join(map(str, x)) for x in vecs])
    want = [sum(v * 256**s for s, v in enumerate(x)) % p for x in vecs]
    assert [int(g) for g in got] == want


def test_mersenne_reduction_below_2_63(fold_check):
    p = 2**31 - 1
    rng = random.Random(7)
    vals = [0, 1, p - 1, p, p + 1, 2 * p, 2**31, 2**32 - 1, 2**32, (p - 1) ** 2, 2 * (p - 1) ** 2,
            (p - 1) ** 2 + (p - 1), 2**62 - 1, 2**62, 2**63 - 1]
    vals += [rng.randrange(2**63) for _ in range(300)]
    got = fold_check([f"R {v}" for v in vals])
    assert [int(g) for g in got] == [v % p for v in vals]


@pytest.mark.parametrize("p", PRIMES)
def test_balanced_digits_represent_every_residue(fold_check, p):
    nl = 2 if p <= 65536 else 4
    maxpos = 127 * (256**nl - 1) // 255
    rng = random.Random(p + 1)
    xs = sorted({x for x in (0, 1, 2, p - 1, p - 2, maxpos, maxpos + 1, maxpos - 1) if 0 <= x < p})
    xs += [rng.randrange(p) for _ in range(300)]
    got = fold_check([f"D {p} {x}" for x in xs])
    for x, line in zip(xs, got):
        d = [int(t) for t in line.split()]
        assert len(d) == nl and all(-128 <= e <= 127 for e in d), (x, d)
        assert sum(e * 256**j for j, e in enumerate(d)) % p == x, (x, d)
