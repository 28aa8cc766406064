"""The C++ facade's host code and the oracle built for the CPU, once plainly
and once with AddressSanitizer + UndefinedBehaviorSanitizer.

* tests/native/host_harness.cpp drives the deck parser (params.cpp), the
  writers (io.cpp), the squirmer helpers (squirmer.cpp) and the analytic
  solutions (analytic.cpp) without the device library.  Besides running
  clean under the sanitizers it pins the analytic code against the
  reference's duct fixtures: the L2 error 2.105303768e-4 of
  outputs/duct_summary.txt and the exact columns of
  outputs/duct_profile_{y,z}.dat, recomputed here from the LBM field of
  outputs/duct_final.vtk (reference analytic.cpp:11-64, main.cpp:104-108).
* tests/native/oracle_harness.c runs every function of the oracle on small
  cases (duct, wall contact, x-z plane, periodic corner).
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import analytic as A

ROOT = Path(__file__).resolve().parents[1]
HOST = ROOT / "lbm_solver_3d_amd" / "host"
NATIVE = ROOT / "tests" / "native"
FIX = ROOT / "tests" / "golden" / "reference_outputs"
HOST_SRC = [HOST / f for f in ("params.cpp", "squirmer.cpp", "io.cpp", "analytic.cpp")]
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
           OMP_NUM_THREADS="2")


def _build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.fixture(scope="module", params=["plain", "sanitized"])
def host_exe(request, tmp_path_factory):
    out = tmp_path_factory.mktemp(f"host_{request.param}") / "host_harness"
    flags = ["-O2"] if request.param == "plain" else SAN
    _build(["g++", "-std=c++17", "-Wall", "-Wextra", *flags, f"-I{HOST}", f"-I{ROOT / 'include'}",
            str(NATIVE / "host_harness.cpp"), *map(str, HOST_SRC), "-o", str(out)])
    return out


@pytest.fixture(scope="module", params=["plain", "sanitized"])
def oracle_exe(request, tmp_path_factory):
    out = tmp_path_factory.mktemp(f"oracle_{request.param}") / "oracle_harness"
    flags = ["-O2"] if request.param == "plain" else SAN
    _build(["gcc", "-std=c11", "-Wall", "-Wextra", "-fopenmp", "-ffp-contract=off", *flags,
            f"-I{ROOT / 'oracle'}", str(NATIVE / "oracle_harness.c"), str(ROOT / "oracle" / "lbm_oracle.c"),
            "-lm", "-o", str(out)])
    return out


def _run(exe, *args, rc=0):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, env=ENV, timeout=300)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert r.returncode == rc, r.stderr
    return r.stdout


def test_reference_decks_parse(host_exe, tmp_path):
    decks = sorted(FIX.glob("*.in"))
    assert len(decks) >= 5
    out = _run(host_exe, "decks", tmp_path, *decks).split("\n")
    got = {Path(l.split()[0]).name: l.split()[1:] for l in out if l}
    assert got["duct.in"] == ["4", "34", "34", "0"]
    assert got["channel_puller.in"][:3] == ["64", "24", "20"] and got["channel_puller.in"][3] == "1"


def test_malformed_decks_throw(host_exe, tmp_path):
    out = _run(host_exe, "bad", tmp_path)
    assert "cannot interpret 'maybe' as a boolean" in out
    assert "unknown parameter 'unknown_key'" in out
    assert "expected 'key = value'" in out
    assert "tau must be > 0.5" in out
    assert "squirmer radius must be > 1 lattice unit" in out
    assert "cannot open input file" in out


def test_writers_formats(host_exe, tmp_path):
    _run(host_exe, "writers", tmp_path)
    d = tmp_path / "nested" / "out"
    v = A.read_vtk(d / "f.vtk")
    assert v["dims"] == (5, 4, 3)
    n = np.arange(60)
    flag = np.where(n % 7 == 0, 1, np.where(n % 5 == 0, 2, 0))
    assert np.array_equal(v["flag"], flag)
    fluid = flag == 0
    assert np.array_equal(v["density"], np.where(fluid, 1.0 + 1e-3 * n, 1.0))
    assert np.array_equal(v["velocity"][:, 0], np.where(fluid, 1e-4 * n, 0.0))
    rows = [l.split() for l in (d / "traj.dat").read_text().splitlines()]
    assert rows[0][0] == "#" and len(rows) == 4 and all(len(r) == 15 for r in rows[1:])
    assert [r[0] for r in rows[1:]] == ["50", "100", "150"]
    state = dict(l.split(" ", 1) for l in (d / "state.dat").read_text().splitlines())
    assert state["t"] == "150" and state["R"] == "4.5" and state["beta"] == "-2"


def _duct_fixture_field():
    v = A.read_vtk(FIX / "duct_final.vtk")
    return v["flag"], v["velocity"][:, 0].copy()


def test_poiseuille_error_reproduces_duct_fixture(host_exe, tmp_path):
    flag, ux = _duct_fixture_field()
    (tmp_path / "flag.bin").write_bytes(flag.astype(np.uint8).tobytes())
    (tmp_path / "ux.bin").write_bytes(ux.astype("<f8").tobytes())
    nu = (0.8 - 0.5) / 3.0
    err = float(_run(host_exe, "poiseuille", 4, 34, 34, 1, 1, repr(nu), "5e-5", tmp_path / "flag.bin",
                     tmp_path / "ux.bin", tmp_path / "duct"))
    summary = dict(l.split(" ", 1) for l in (FIX / "duct_summary.txt").read_text().splitlines())
    assert err == pytest.approx(float(summary["l2_error"]), rel=1e-8)
    for axis in ("y", "z"):
        got = np.loadtxt(tmp_path / f"duct_profile_{axis}.dat")
        ref = np.loadtxt(FIX / f"duct_profile_{axis}.dat")
        assert got.shape == ref.shape == (32, 3)
        assert np.array_equal(got[:, 0], ref[:, 0])
        assert np.allclose(got[:, 1:], ref[:, 1:], rtol=1e-9, atol=0)
        assert (tmp_path / f"duct_profile_{axis}.dat").read_text().splitlines()[0] == \
            (FIX / f"duct_profile_{axis}.dat").read_text().splitlines()[0]


def test_duct_series_closed_forms(host_exe):
    """Symmetry in y and z, zero on the walls, and the square-duct centre
    value u(0,0) = 0.2947 G a^2 / mu (White, Viscous Fluid Flow, Table 3-2:
    Q-free form of the series, 0.29469...)."""
    a, b, G, mu = 16.0, 16.0, 5e-5, 0.1
    pts = [(0, 0), (3.5, -2.5), (-3.5, 2.5), (3.5, 2.5), (a, 0), (0, b), (7.25, -b)]
    out = [float(x) for x in _run(host_exe, "duct", a, b, G, mu, *[c for p in pts for c in p]).split()]
    assert out[0] == pytest.approx(0.2946854 * G * a * a / mu, rel=1e-6)
    assert out[1] == pytest.approx(out[2], rel=1e-14) and out[1] == pytest.approx(out[3], rel=1e-14)
    for v in out[4:]:
        assert abs(v) < 1e-9 * out[0]
    # the plane-channel limit: a duct much longer in z than wide in y (the
    # 200-harmonic cut leaves a relative error of order 1e-8)
    far = float(_run(host_exe, "duct", 8.0, 400.0, G, mu, 2.0, 0.0).split()[0])
    assert far == pytest.approx(G / (2 * mu) * (8.0 ** 2 - 2.0 ** 2), rel=5e-8)


def test_oracle_cases_run_clean(oracle_exe):
    out = _run(oracle_exe)
    lines = {l.split()[0]: l.split() for l in out.splitlines() if l}
    assert set(lines) == {"duct", "wall_contact", "plane_xz", "corner", "helpers"}
    for name in ("wall_contact", "plane_xz", "corner"):
        excess = float(lines[name][lines[name].index("mass_excess") + 1])
        assert abs(excess) < 1e-8
    assert int(lines["wall_contact"][lines["wall_contact"].index("contacts") + 1]) > 0
