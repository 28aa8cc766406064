"""Pinning the oracle (oracle/lbm_oracle.c) before it is trusted as the checker.

* closed forms of the velocity set, the equilibrium and the TRT + Guo
  collision, evaluated independently in numpy (a dropped term, a wrong sign or
  a wrong opposite index fails them);
* the reference compiled from its own sources (oracle/_ref), bit for bit;
* the golden vectors in tests/golden/*.npz written from that build;
* the reference's own output fixtures (outputs/duct_*, channel_puller_traj);
* the reference's own acceptance tests (tests/test_duct_flow.py,
  tests/test_squirmer_bulk.py) re-run on the oracle.
"""
from pathlib import Path

import numpy as np
import pytest

import analytic as A

GOLD = Path(__file__).resolve().parent / "golden"
REFOUT = GOLD / "reference_outputs"
Q = 19


@pytest.fixture(scope="module")
def lat(oracle_mod):
    lib = oracle_mod._load()
    c = np.array([[lib.orc_cx(i), lib.orc_cy(i), lib.orc_cz(i)] for i in range(Q)], float)
    w = np.array([lib.orc_w(i) for i in range(Q)])
    opp = np.array([lib.orc_opp(i) for i in range(Q)])
    return c, w, opp


# --------------------------------------------------------------- closed forms
def test_velocity_set_invariants(lat):
    c, w, opp = lat
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert np.allclose(w @ c, 0.0, atol=1e-16)
    assert np.allclose(np.einsum("i,ia,ib->ab", w, c, c), np.eye(3) / 3.0, atol=1e-15)
    assert np.allclose(np.einsum("i,ia,ib,ic->abc", w, c, c, c), 0.0, atol=1e-16)
    assert np.array_equal(c[opp], -c)
    assert np.array_equal(opp[opp], np.arange(Q))
    # opposite pairs adjacent (lattice.hpp:17)
    assert all(opp[i] == (i + 1 if i % 2 else i - 1) for i in range(1, Q))
    assert len({tuple(v) for v in c}) == Q


def test_equilibrium_moments(oracle_mod, lat):
    c, w, _ = lat
    lib = oracle_mod._load()
    rng = np.random.default_rng(1)
    for _ in range(20):
        rho = 0.9 + 0.2 * rng.random()
        u = 0.05 * (rng.random(3) - 0.5)
        feq = np.array([lib.orc_equilibrium(i, rho, *u) for i in range(Q)])
        ref = w * rho * (1 + 3 * c @ u + 4.5 * (c @ u) ** 2 - 1.5 * u @ u)
        assert np.allclose(feq, ref, rtol=1e-15, atol=0)
        assert feq.sum() == pytest.approx(rho, rel=1e-14)
        assert np.allclose(feq @ c, rho * u, atol=1e-15)


def _trt_guo(f, omP, omM, F, c, w, opp):
    """TRT collision with Guo forcing from its definition, long double."""
    f = f.astype(np.longdouble)
    rho = f.sum()
    u = (f @ c + 0.5 * F) / rho
    cu, cF = c @ u, c @ F
    feq = w * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))
    S = w * (3 * (cF - u @ F) + 9 * cu * cF)
    fp, fm = 0.5 * (f + f[opp]), 0.5 * (f - f[opp])
    ep, em = 0.5 * (feq + feq[opp]), 0.5 * (feq - feq[opp])
    Sp, Sm = 0.5 * (S + S[opp]), 0.5 * (S - S[opp])
    return f - omP * (fp - ep) - omM * (fm - em) + (1 - omP / 2) * Sp + (1 - omM / 2) * Sm


@pytest.mark.parametrize("tau,lam", [(0.8, 3 / 16), (1.0, 3 / 16), (0.6, 0.25), (0.9, (0.9 - 0.5) ** 2)])
def test_collision_matches_trt_definition(oracle_mod, lat, tau, lam):
    """One step of a periodic 3x3x3 box from random populations: every pushed
    value equals the TRT+Guo post-collision value of its source node."""
    c, w, opp = lat
    F = np.array([3e-4, -2e-4, 1e-4])
    o = oracle_mod.Oracle(nx=3, ny=3, nz=3, tau=tau, **{"lambda": lam}, fx=F[0], fy=F[1], fz=F[2],
                          walls_y=0, walls_z=0)
    N = 27
    rng = np.random.default_rng(7)
    f0 = w[:, None] * (1 + 0.05 * (rng.random((Q, N)) - 0.5))
    np.ctypeslib.as_array(o.L.f, shape=(Q * N,))[:] = f0.reshape(-1)
    o.step()
    f1 = o.f()
    omP = 1.0 / tau
    omM = 1.0 / (0.5 + lam / (tau - 0.5))
    if lam == (tau - 0.5) ** 2:
        assert omM == pytest.approx(omP)  # BGK limit
    for n in range(N):
        x, y, z = n % 3, (n // 3) % 3, n // 9
        post = _trt_guo(f0[:, n], omP, omM, F, c, w, opp)
        for i in range(Q):
            m = (x + int(c[i, 0])) % 3 + 3 * ((y + int(c[i, 1])) % 3) + 9 * ((z + int(c[i, 2])) % 3)
            assert f1[i, m] == pytest.approx(float(post[i]), rel=2e-15, abs=1e-17)
    # mass conserved; momentum gains F per node (Guo)
    assert f1.sum() == pytest.approx(f0.sum(), rel=1e-14)
    assert np.allclose(f1.T.sum(0) @ c, f0.T.sum(0) @ c + N * F, rtol=0, atol=1e-14)


def test_rest_state_is_fixed_point_with_walls(oracle_mod, lat):
    _, w, _ = lat
    o = oracle_mod.Oracle(nx=5, ny=8, nz=7, walls_y=1, walls_z=1)
    o.step(25)
    f = o.f()
    fl = o.flags() == 0
    # exact in exact arithmetic; rounding of rho = sum(w) leaves ulp-level noise
    assert np.abs(f[:, fl] - w[:, None]).max() < 1e-15


def test_initial_particle_flags_brute_force(oracle_mod):
    for kv in [dict(nx=20, ny=18, nz=16, walls_y=1, walls_z=0, radius=4.3, y0=8.6, x0=19.2, z0=0.4),
               dict(nx=16, ny=16, nz=16, walls_y=0, walls_z=0, radius=3.0, x0=0.0, y0=15.5, z0=7.25)]:
        o = oracle_mod.Oracle(squirmer=1, angle=0, **kv)
        NX, NY, NZ = kv["nx"], kv["ny"], kv["nz"]
        X = np.array([kv["x0"], kv["y0"], kv["z0"]])
        z, y, x = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
        r = [x - X[0], y - X[1], z - X[2]]
        r[0] = r[0] - NX * np.round(r[0] / NX)
        if not kv["walls_y"]:
            r[1] = r[1] - NY * np.round(r[1] / NY)
        r[2] = r[2] - NZ * np.round(r[2] / NZ)
        inside = (r[0] ** 2 + r[1] ** 2 + r[2] ** 2 < kv["radius"] ** 2).reshape(-1)
        wall = np.zeros_like(inside)
        if kv["walls_y"]:
            wall = ((y == 0) | (y == NY - 1)).reshape(-1)
        expect = np.where(wall, 1, np.where(inside, 2, 0))
        assert np.array_equal(o.flags(), expect.astype(np.uint8))


def test_squirmer_known_answers(oracle_mod):
    import ctypes as C
    lib = oracle_mod._load()
    s = oracle_mod.OrcSquirmer()
    s.B1, s.beta = 0.015, -3.0
    s.e[:] = [1.0, 0.0, 0.0]
    s.U[:] = [1e-3, 2e-3, 0.0]
    s.W[:] = [0.0, 0.0, 1e-4]
    ub = (C.c_double * 3)()
    # equator r = (0, R, 0): cos th = 0, slip = -B1 e; W x r = (-1e-4*R, 0, 0)
    R = 4.0
    lib.orc_surface_velocity(C.byref(s), (C.c_double * 3)(0.0, R, 0.0), ub, 1)
    assert np.allclose(ub[:], [1e-3 - 1e-4 * R - 0.015, 2e-3, 0.0], rtol=1e-15, atol=1e-18)
    # pole r = (R, 0, 0): cos th = 1, slip = B1 (1 + beta)(r^ - e) = 0
    lib.orc_surface_velocity(C.byref(s), (C.c_double * 3)(R, 0.0, 0.0), ub, 1)
    assert np.allclose(ub[:], [1e-3, 2e-3 + 1e-4 * R, 0.0], rtol=1e-15, atol=1e-18)
    # 45 deg: r^ = (1,1,0)/sqrt2, ct = 1/sqrt2, amp = B1 (1 + beta ct)
    rh = np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    lib.orc_surface_velocity(C.byref(s), (C.c_double * 3)(*(R * rh)), ub, 1)
    ct = rh[0]
    slip = 0.015 * (1 - 3.0 * ct) * (ct * rh - np.array([1.0, 0, 0]))
    rig = np.array([1e-3, 2e-3, 0.0]) + np.cross([0, 0, 1e-4], R * rh)
    assert np.allclose(ub[:], rig + slip, rtol=1e-14, atol=1e-18)
    # without slip only the rigid-body part
    lib.orc_surface_velocity(C.byref(s), (C.c_double * 3)(*(R * rh)), ub, 0)
    assert np.allclose(ub[:], rig, rtol=1e-15)
    # soft wall force: eps FS (1 - h/hc)^2 for 0 <= h < hc, eps FS for h < 0
    assert lib.orc_soft_wall_force(2.0, 0.2, 0.5, 2.0) == 0.0
    assert lib.orc_soft_wall_force(1.0, 0.2, 0.5, 2.0) == pytest.approx(0.2 * 0.5 * 0.25)
    assert lib.orc_soft_wall_force(-1.0, 0.2, 0.5, 2.0) == pytest.approx(0.2 * 0.5)
    # Newton-Euler: U += F/M, W += T/I, X += U, e rotated by W x e and renormalised
    t = oracle_mod.OrcSquirmer()
    t.M, t.I, t.planar = 2.0, 4.0, 1
    t.e[:] = [1.0, 0.0, 0.0]
    t.Fh[:] = [0.2, 0.4, 9.0]
    t.Fc[:] = [0.0, -0.2, 0.0]
    t.Th[:] = [5.0, 5.0, 0.08]
    lib.orc_integrate(C.byref(t))
    assert np.allclose(t.U[:], [0.1, 0.1, 0.0]) and np.allclose(t.W[:], [0.0, 0.0, 0.02])
    assert np.allclose(t.X[:], [0.1, 0.1, 0.0])
    e = np.array([1.0, 0.02, 0.0]) / np.hypot(1.0, 0.02)
    assert np.allclose(t.e[:], e, rtol=1e-15)
    assert list(t.Fc[:]) == [0.0, 0.0, 0.0]


# ------------------------------------------------- the reference, bit for bit
REF_CASES = [
    dict(nx=4, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1),
    dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4, b1=0.015, beta=-3, angle=20),
    dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.03, beta=-1, angle=-40, y0=6.2),
    dict(nx=30, ny=20, nz=18, walls_y=1, walls_z=1, squirmer=1, radius=3.5, b1=0.02, beta=2, angle=10, planar=0,
         y0=9.0, z0=8.3),
]


@pytest.mark.parametrize("kv", REF_CASES)
def test_oracle_bit_identical_to_reference_build(oracle_mod, kv):
    if not oracle_mod.reference_available():
        pytest.skip("reference build (oracle/_ref) not available")
    o = oracle_mod.Oracle(**kv)
    r = oracle_mod.Reference(**kv)
    r.set_threads(1)
    for t in range(1, 121):
        o.step(); r.step()
        if kv.get("squirmer"):
            o.advance_squirmer(); r.advance_squirmer()
            if t % 10 == 0:
                o.correct_mass(); r.correct_mass()
    rf = r.fields()
    fl = rf["flag"] == 0
    assert np.array_equal(o.flags(), rf["flag"])
    assert np.array_equal(o.f()[:, fl], rf["f"][:, fl])
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(o.field(k), rf[k])
    so, sr = o.state(), r.state()
    for k in ("X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw", "wallForce"):
        assert np.array_equal(so[k], sr[k]), k
    assert so["contacts"] == sr["contacts"]


# ----------------------------------------------------------- golden vectors
def _parse_config(s):
    kv = {}
    for tok in str(s).split():
        k, v = tok.split("=")
        kv[k] = float(v) if "." in v or "e" in v else int(v)
    return kv


@pytest.mark.parametrize("name", sorted(p.stem for p in GOLD.glob("*.npz")))
def test_oracle_matches_golden(oracle_mod, name):
    g = np.load(GOLD / f"{name}.npz")
    kv = _parse_config(g["config"])
    o = oracle_mod.Oracle(**kv)
    steps, me = int(g["steps"]), int(g["mass_every"])
    hist = {k: [] for k in ("X", "U", "W", "Fh", "Th", "contacts")}
    for t in range(1, steps + 1):
        o.step()
        if me:
            o.advance_squirmer()
            if t % me == 0:
                o.correct_mass()
        s = o.state()
        for k in hist:
            hist[k].append(s[k])
    fl = g["flag"]
    assert np.array_equal(o.flags(), fl)
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(o.field(k), g[k]), k
    f = o.f()
    if "f_fluid" in g:
        assert np.array_equal(f[:, fl == 0], g["f_fluid"])
    else:
        NX, NY = o.NX, o.NY
        zc = int(g["plane_z"])
        assert np.array_equal(f[:, zc * NX * NY:(zc + 1) * NX * NY][:, fl[zc * NX * NY:(zc + 1) * NX * NY] == 0],
                              g["f_plane"][:, fl[zc * NX * NY:(zc + 1) * NX * NY] == 0])
        ff = f[:, fl == 0]
        assert np.array_equal(ff.sum(axis=1), g["f_sum"])
    for k in ("X", "U", "W", "Fh", "Th"):
        assert np.array_equal(np.array(hist[k]), g["hist_" + k]), k
    assert np.array_equal(np.array(hist["contacts"]), g["hist_contacts"])


# ------------------------------------------------ reference output fixtures
def run_to_steady(obj, steps, print_every, tol, sum_ux):
    """main.cpp:60-87 without a squirmer: stop when the relative change of
    sum(u_x) between prints drops below tol."""
    prev = 0.0
    for t in range(1, steps + 1):
        obj.step()
        if t % print_every == 0:
            sx = sum_ux()
            change = abs(sx - prev) / (abs(sx) + 1e-30)
            prev = sx
            if tol > 0 and change < tol:
                return t, True
    return steps, False


def read_kv(path):
    out = {}
    for line in Path(path).read_text().splitlines():
        p = line.split()
        if p:
            vals = [float(v) for v in p[1:]]
            out[p[0]] = vals[0] if len(vals) == 1 else np.array(vals)
    return out


@pytest.fixture(scope="module")
def duct_fixture_run(oracle_mod):
    kv = dict(nx=4, ny=34, nz=34, tau=0.8, fx=5e-5, walls_y=1, walls_z=1)  # inputs/duct.in
    o = oracle_mod.Oracle(**kv)
    done, conv = run_to_steady(o, 100000, 1000, 1e-12, lambda: o.diagnostics()["sum_ux"])
    return o, done, conv


def test_duct_summary_fixture(duct_fixture_run):
    """outputs/duct_summary.txt:1-8 (reference run of inputs/duct.in)."""
    o, done, conv = duct_fixture_run
    ref = read_kv(REFOUT / "duct_summary.txt")
    assert done == ref["steps"] and conv
    d = o.diagnostics()
    assert d["mean_density"] == pytest.approx(ref["mean_density"], abs=1e-10)
    assert d["max_speed"] == pytest.approx(ref["max_speed"], rel=1e-9)
    wf = np.array(o.L.wallForce[:])
    assert wf[0] == pytest.approx(ref["wall_force"][0], rel=1e-9)
    assert wf[0] == pytest.approx(ref["body_force_total"][0], rel=1e-6)
    err = A.poiseuille_error(o.field("ux"), o.flags(), 4, 34, 34, True, True, o.L.nu, 5e-5)
    assert err == pytest.approx(ref["l2_error"], rel=1e-8)


def test_duct_vtk_and_profiles_fixture(duct_fixture_run):
    """outputs/duct_final.vtk and duct_profile_{y,z}.dat."""
    o, _, _ = duct_fixture_run
    v = A.read_vtk(REFOUT / "duct_final.vtk")
    assert v["dims"] == (4, 34, 34)
    assert np.array_equal(v["flag"], o.flags())
    fl = o.flags() == 0
    assert np.allclose(v["density"][fl], o.field("rho")[fl], rtol=1e-12)
    assert np.allclose(v["velocity"][fl, 0], o.field("ux")[fl], rtol=1e-8, atol=1e-15)
    ux = o.field("ux").reshape(34, 34, 4)
    py = np.loadtxt(REFOUT / "duct_profile_y.dat")
    assert np.allclose(py[:, 1], ux[17, 1:33, 0], rtol=1e-8)
    pz = np.loadtxt(REFOUT / "duct_profile_z.dat")
    assert np.allclose(pz[:, 1], ux[1:33, 17, 0], rtol=1e-8)
    # exact-solution column of the fixture == this series
    a = b = 16.0
    assert np.allclose(pz[:, 2], A.duct_velocity(17 - 16.5, np.arange(1, 33) - 16.5, a, b, 5e-5, 0.1), rtol=1e-9)


def test_duct_second_order_convergence(oracle_mod):
    """tests/test_duct_flow.py:16-42 (reference acceptance): H = 8, 16, 32."""
    errs = []
    for H in (8, 16, 32):
        fx = 5e-5 * (32 / H) ** 2
        o = oracle_mod.Oracle(nx=4, ny=H + 2, nz=H + 2, tau=0.8, fx=fx, walls_y=1, walls_z=1)
        done, conv = run_to_steady(o, 100000, 1000, 1e-12, lambda: o.diagnostics()["sum_ux"])
        assert conv
        errs.append(A.poiseuille_error(o.field("ux"), o.flags(), 4, H + 2, H + 2, True, True, o.L.nu, fx))
        d = o.diagnostics()
        assert abs(d["mean_density"] - 1.0) < 1e-10
        assert abs(o.L.wallForce[0] - fx * d["count_fluid"]) / (fx * d["count_fluid"]) < 1e-6
    order = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(2)
    assert errs[-1] < 5e-4
    assert np.all(order > 1.8), order
    # tutorial.ipynb:331-334 values
    assert errs == pytest.approx([4.025e-3, 8.803e-4, 2.105e-4], rel=2e-3)


def test_plane_channel_exact(oracle_mod):
    """tests/test_duct_flow.py:52-56: TRT with Lambda = 3/16 is exact for plane Poiseuille."""
    o = oracle_mod.Oracle(nx=4, ny=34, nz=1, walls_y=1, walls_z=0, tau=0.8, fx=1e-5)
    done, conv = run_to_steady(o, 100000, 1000, 1e-12, lambda: o.diagnostics()["sum_ux"])
    assert conv
    err = A.poiseuille_error(o.field("ux"), o.flags(), 4, 34, 1, True, False, o.L.nu, 1e-5)
    assert err < 1e-6


def test_channel_puller_trajectory_fixture(oracle_mod):
    """First rows of outputs/channel_puller_traj.dat (inputs/channel_puller.in)."""
    kv = dict(nx=64, ny=24, nz=20, tau=1.0, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.015, beta=1,
              angle=30, planar=1, eps=0.2, hc=2.0, hmin=0.5)
    ref = np.loadtxt(REFOUT / "channel_puller_traj.dat")
    o = oracle_mod.Oracle(**kv)
    rows = []
    for t in range(1, 501):
        o.step(); o.advance_squirmer()
        if t % 10 == 0:
            o.correct_mass()
        if t % 50 == 0:
            s = o.state()
            rows.append([t, *s["X"], *s["e"], *s["U"], *s["W"], s["Fw"][1], s["Fw"][2]])
    rows = np.array(rows)
    r = ref[:len(rows)]
    assert np.array_equal(rows[:, 0], r[:, 0])
    assert np.allclose(rows[:, 1:7], r[:, 1:7], atol=1.5e-6)          # %.6f columns
    assert np.allclose(rows[:, 7:13], r[:, 7:13], rtol=2e-6, atol=1e-12)  # %.6e columns


def test_plane_xz_extension_is_relabelled_xy(oracle_mod):
    """The x-z orientation extension (no reference counterpart) against the
    reference's x-y behaviour with y and z relabelled: walls in y + x-y plane
    on an (nx, ny, nz) grid == walls in z + x-z plane on (nx, nz, ny)."""
    a = oracle_mod.Oracle(nx=30, ny=20, nz=14, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.02, beta=-2,
                          angle=25, planar=1, y0=8.0)
    b = oracle_mod.Oracle(nx=30, ny=14, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4, b1=0.02, beta=-2,
                          angle=25, planar=1, plane_xz=1, z0=8.0, y0=6.5)
    # Identical up to summation order for 100 steps; later a node whose
    # inside test sits within round-off of R^2, or the sweep order of the
    # remap (z-outer, so it changes under y<->z), can make the runs part.
    for t in range(1, 101):
        for o in (a, b):
            o.step(); o.advance_squirmer()
            if t % 10 == 0:
                o.correct_mass()
    sa, sb = a.state(), b.state()
    swap = [0, 2, 1]
    assert np.abs(sa["X"] - sb["X"][swap]).max() < 1e-12
    assert np.abs(sa["e"] - sb["e"][swap]).max() < 1e-13
    assert np.abs(sa["U"] - sb["U"][swap]).max() < 1e-14
    assert sa["W"][2] == pytest.approx(-sb["W"][1], rel=1e-10)   # rotation about z <-> about -y
    assert abs(sa["W"][2]) > 1e-6
    assert sa["contacts"] == sb["contacts"]


def test_exact_mass_sum_mode_agrees_on_small_lattices(oracle_mod):
    """The oracle's exact-sum mass correction (used where the reference's
    double sum rounds visibly, at millions of nodes) equals the reference's
    summation to round-off on a small lattice with remaps and corrections."""
    kv = dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4.0, b1=0.03, beta=-1.0,
              angle=-40.0, y0=6.2, hmin=0.5)
    out = []
    for exact in (False, True):
        oracle_mod.set_mass_sum_exact(exact)
        try:
            o = oracle_mod.Oracle(**kv)
            o.run(120, 10)
            out.append((o.f().copy(), o.flags().copy(), np.array(o.state()["X"])))
        finally:
            oracle_mod.set_mass_sum_exact(False)
    (fa, la, xa), (fb, lb, xb) = out
    assert np.array_equal(la, lb)
    m = la == 0
    assert np.abs(fa[:, m] - fb[:, m]).max() <= 1e-14
    assert np.abs(xa - xb).max() <= 1e-12
    # and the total mass: the exact mode restores it to round-off
    assert abs(fb[:, m].sum() / m.sum() - 1.0) < 1e-13
