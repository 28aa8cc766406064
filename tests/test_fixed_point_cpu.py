"""The 128-bit fixed-point conversion of the excess-mass partial sums
(csrc/fixed_point.hpp), built for the host, against exact rational
arithmetic: V = trunc(v * 2^64) toward zero for every |v| < 2^63.

The cases include the ones a conversion through e - floor(e) gets wrong:
tiny negative values, where 1 + e rounds to 1 (an out-of-range cast) or to
a 2^-53 step, and values whose last significand bits sit below 2^-64."""
import random
import subprocess
from fractions import Fraction
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "lbm_solver_3d_amd" / "csrc"
SRC = ROOT / "tests" / "native" / "fixed_point_host.cpp"

CASES = [0.0, -0.0, 3e-19, -3e-19, -1e-17, -2.0 ** -54, -2.0 ** -53, -1e-10, -0.5, 1e3,
         2.0 ** -64, -2.0 ** -64, 2.0 ** -65, -(2.0 ** -65), 5e-324, -5e-324, 2.2250738585072014e-308,
         1.0 - 2.0 ** -53, -1.0 + 2.0 ** -53, 0.1, -0.1, 1.0 / 3.0, -19.0, 12345.678901234567,
         2.0 ** 62 + 2.0 ** 10, -(2.0 ** 62) - 2.0 ** 10, 1.5e-20, -7.7e-18]


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("fixed") / "fixed_point_host"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{CSRC}", str(SRC), "-o", str(out)],
                   check=True)
    return out


def _run(exe, values):
    r = subprocess.run([str(exe), *[v.hex() for v in values]], check=True, capture_output=True, text=True)
    lines = r.stdout.split("\n")
    got = []
    for line in lines[:len(values)]:
        hi, lo = line.split()
        got.append(int(hi) * 2 ** 64 + int(lo))
    hi, lo, back = lines[len(values)].split()
    return got, int(hi) * 2 ** 64 + int(lo), float.fromhex(back)


def _exact(v):
    return int(Fraction(v) * 2 ** 64)  # int() truncates toward zero


def test_known_cases_exact(exe):
    got, _, _ = _run(exe, CASES)
    for v, q in zip(CASES, got):
        assert q == _exact(v), (v, q, _exact(v))
    # the cases a conversion through e - floor(e) mishandles
    assert got[CASES.index(-1e-17)] == -184
    assert got[CASES.index(-3e-19)] == -5
    assert got[CASES.index(-(2.0 ** -54))] == -(2 ** 10)


def test_random_values_and_sum(exe):
    rng = random.Random(7)
    vals = []
    for _ in range(400):
        e = rng.uniform(-70, 5)
        vals.append(rng.choice((-1.0, 1.0)) * rng.random() * 2.0 ** e)
    got, total, back = _run(exe, vals)
    for v, q in zip(vals, got):
        assert q == _exact(v)
        assert abs(Fraction(v) - Fraction(q, 2 ** 64)) < Fraction(1, 2 ** 64)
    # integer sums are exact and order-free; the double read-back rounds once per part
    assert total == sum(got)
    assert abs(back - float(Fraction(total, 2 ** 64))) <= 2 * abs(float(Fraction(total, 2 ** 64))) * 2.0 ** -52 + 2.0 ** -64


def test_non_finite_and_saturation(exe):
    got, _, _ = _run(exe, [float("nan"), float("inf"), -float("inf"), 2.0 ** 80, -(2.0 ** 80)])
    assert got[:3] == [0, 0, 0]
    assert got[3] == 2 ** 127 - 1 and got[4] == -(2 ** 127 - 1)
