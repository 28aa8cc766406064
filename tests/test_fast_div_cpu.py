"""Division by a run-time constant (csrc/fast_div.hpp), built for the host,
against Python's integer floor division: every numerator below 2^20 for
every divisor up to 600, and for divisors up to 2^31 - 1 the numerators
next to multiples of the divisor and next to 2^31 (the largest the kernels
may pass), where a magic number one unit off fails first."""
import random
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "lbm_solver_3d_amd" / "csrc"
SRC = ROOT / "tests" / "native" / "fast_div_host.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("fastdiv") / "fast_div_host"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{CSRC}", str(SRC), "-o", str(out)],
                   check=True)
    return out


def test_every_small_numerator_and_divisor(exe):
    r = subprocess.run([str(exe), "sweep", "600", str(1 << 20)], check=True, capture_output=True, text=True)
    assert int(r.stdout) == 0


def test_numerators_near_multiples_and_the_limit(exe):
    rng = random.Random(3)
    ds = [1, 2, 3, 5, 7, 641, 1000, 4095, 4096, 4097, 65535, 65536, 65537, (1 << 20) + 1, (1 << 24) - 1,
          (1 << 30) - 1, 1 << 30, (1 << 30) + 3, (1 << 31) - 1]
    ds += [rng.randrange(1, 1 << 31) for _ in range(200)]
    top = (1 << 31) - 1
    pairs = []
    for d in ds:
        ns = {0, 1, top, top - 1, top - d % (1 << 31), d - 1, d, d + 1}
        for _ in range(40):
            k = rng.randrange(0, top // d + 1)
            ns.update({k * d - 1, k * d, k * d + 1, k * d + d - 1})
        ns.update(rng.randrange(0, top + 1) for _ in range(40))
        pairs += [(n, d) for n in sorted(ns) if 0 <= n <= top]
    text = "".join(f"{n} {d}\n" for n, d in pairs)
    r = subprocess.run([str(exe), "pairs"], input=text, check=True, capture_output=True, text=True)
    got = [int(v) for v in r.stdout.split()]
    assert len(got) == len(pairs)
    bad = [(n, d, q) for (n, d), q in zip(pairs, got) if q != n // d]
    assert not bad, bad[:5]
