"""The reference's own acceptance tests and the validation workloads (duct
64x32x32, bulk 128^3 squirmer), run on the GPU path in every storage
precision.

* duct flow: L2 error against the series solution, second-order convergence,
  wall force = body force, mass (reference tests/test_duct_flow.py);
* plane channel exact with Lambda = 3/16;
* 64x32x32 duct with tau = 1, Fx = 1e-6, 20000 steps, second order over
  cross-sections 16 / 32 / 64;
* squirmer in a periodic box: swimming speed, straight swimming, pusher >
  neutral > puller, mass, Blake near field (tests/test_squirmer_bulk.py);
* 128^3 neutral squirmer (R = 6, B1 = 1.5e-3): speed against the FP64 run
  and against Blake's 2 B1 / 3.
Reduced precisions are compared with the FP64 mode, which itself matches the
CPU oracle bit for bit (duct) or to 1e-12 (squirmer), see
test_gpu_parity_fp64.py.  Tolerances are written next to each assertion.
"""
import numpy as np
import pytest

import analytic as A

pytestmark = pytest.mark.gpu

PRECS = ["fp64", "fp32", "fp16", "bf16"]


def steady(L, lbm, steps=100000, print_every=1000, tol=1e-12):
    """main.cpp:60-87 convergence loop on the GPU path."""
    prev, t = 0.0, 0
    while t < steps:
        L.run(print_every - 1, 0, 0)
        L.step(lbm.STEP_ALL)
        t += print_every
        sx = L.diagnostics()["sum_ux"]
        change = abs(sx - prev) / (abs(sx) + 1e-30)
        prev = sx
        if tol > 0 and change < tol:
            return t, True
    return t, False


# duct L2 errors of the reference at H = 8 / 16 / 32 (tau = 0.8, fx = 5e-5 (32/H)^2),
# reproduced by the oracle (tests/test_oracle_cpu.py) and the notebook (tutorial.ipynb:331-334)
DUCT_L2 = {8: 4.025e-3, 16: 8.803e-4, 32: 2.105e-4}
# Storage rounding of a converged Stokes flow is amplified by the slowest
# relaxation time (~H^2/nu steps), so each storage format adds an error floor
# to the discretisation error.  FP64 / FP32: relative to the FP64 values.
# FP16 / BF16: the flow stagnates once the per-step change of a population
# falls below half an ulp of its stored value (wall/body force balance 0.4 %
# off for FP16, 14 % for BF16 at H = 32), so the error is a storage floor and
# is capped absolutely.  Measured: FP16 5.4e-3 / 7.5e-3 / 2.1e-2 and BF16
# 2.4e-2 / 4.0e-2 / 2.4e-1 at H = 8 / 16 / 32 (deterministic).
DUCT_TOL = {"fp64": 2e-3, "fp32": 0.25}
DUCT_CAP = {"fp16": 3e-2, "bf16": 0.35}
STEADY_TOL = {"fp64": 1e-12, "fp32": 1e-7, "fp16": 1e-5, "bf16": 1e-4}


@pytest.mark.parametrize("prec", PRECS)
def test_duct_against_exact_solution(lbm, prec):
    errs = []
    for H in (8, 16, 32):
        fx = 5e-5 * (32 / H) ** 2
        with lbm.Lattice(nx=4, ny=H + 2, nz=H + 2, tau=0.8, fx=fx, walls_y=1, walls_z=1, precision=prec) as L:
            done, conv = steady(L, lbm, steps=40000, tol=STEADY_TOL[prec])
            fd = L.fields()
            err = A.poiseuille_error(fd["ux"], fd["flag"], 4, H + 2, H + 2, True, True, L.nu, fx)
            errs.append(err)
            d = L.diagnostics()
            fw = L.wall_force()[0]
            fb = fx * d["count_fluid"]
            print(f"duct {prec} H={H}: L2 {err:.4e} (FP64 {DUCT_L2[H]:.4e}), steps {done}, "
                  f"wall/body {fw / fb:.9f}, <rho> {d['mean_density']:.12f}")
            if prec in ("fp64", "fp32"):
                assert conv
                assert abs(fw - fb) / fb < (1e-6 if prec == "fp64" else 1e-3), (H, fw, fb)  # test_duct_flow.py:45-49
                assert abs(d["mean_density"] - 1.0) < (1e-10 if prec == "fp64" else 1e-6)
    for H, err in zip((8, 16, 32), errs):
        if prec in DUCT_TOL:
            assert err == pytest.approx(DUCT_L2[H], rel=DUCT_TOL[prec]), (prec, H, err)
        else:
            assert err < DUCT_CAP[prec], (prec, H, err)
    order = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(2)
    if prec == "fp64":
        assert np.all(order > 1.8), order          # test_duct_flow.py:39-42
        assert errs[-1] < 5e-4                      # test_duct_flow.py:33-36


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_plane_channel_exact(lbm, prec):
    with lbm.Lattice(nx=4, ny=34, nz=1, walls_y=1, walls_z=0, tau=0.8, fx=1e-5, precision=prec) as L:
        done, conv = steady(L, lbm, tol=STEADY_TOL[prec])
        fd = L.fields()
        err = A.poiseuille_error(fd["ux"], fd["flag"], 4, 34, 1, True, False, L.nu, 1e-5)
    assert conv
    print(f"plane channel {prec}: L2 {err:.3e}")
    # test_duct_flow.py:52-56 (exact to round-off with Lambda = 3/16); FP32
    # storage leaves the amplified rounding floor described above
    assert err < (1e-6 if prec == "fp64" else 1e-4), err


@pytest.mark.parametrize("axis", ["y", "z"])
def test_body_force_off_x_fast_storages(lbm, axis):
    """A plane channel driven along y (walls in z) or z (walls in y): the
    FP32-collision storages pick the forcing kernels from any force
    component, and follow the FP64 run.  FP16 lags by its storage floor: the
    force adds about one half-precision ulp per step to a population here,
    so the developing flow trails FP64 by about 5 % (measured 4.7 / 5.2 %) while a
    dropped force would leave it at rest."""
    if axis == "z":
        kv = dict(nx=4, ny=34, nz=4, walls_y=1, walls_z=0, tau=0.8, fz=1e-5)
    else:
        kv = dict(nx=4, ny=4, nz=34, walls_y=0, walls_z=1, tau=0.8, fy=1e-5)
    u = {}
    for prec in ("fp64", "fp32", "fp16"):
        with lbm.Lattice(**kv, precision=prec) as L:
            L.run(3000, 0, lbm.STEP_ALL)
            fd = L.fields()
            u[prec] = fd["uz" if axis == "z" else "uy"][fd["flag"] == 0]
            other = fd["ux"][fd["flag"] == 0]
            assert np.abs(other).max() < 1e-6 * max(np.abs(u[prec]).max(), 1e-30) + 1e-12
    peak = np.abs(u["fp64"]).max()
    e32, e16 = (np.abs(u[p] - u["fp64"]).max() / peak for p in ("fp32", "fp16"))
    print(f"body force along {axis}: centre-line {peak:.4e}, FP32 {e32:.2e}, FP16 {e16:.2e} of it")
    assert peak > 5e-3  # driven: the centre-line speed after 3000 steps
    assert e32 < 1e-4
    assert e16 < 0.1


def test_duct_64x32x32_second_order_fp64(lbm):
    """64x32x32-class duct, tau = 1, Fx = 1e-6, f = feq, 20000 steps; cross-sections 16/32/64."""
    errs = {}
    for H in (16, 32, 64):
        with lbm.Lattice(nx=4, ny=H + 2, nz=H + 2, tau=1.0, fx=1e-6, walls_y=1, walls_z=1,
                         precision="fp64") as L:
            L.run(20000, 0, lbm.STEP_ALL)
            fd = L.fields()
            errs[H] = A.poiseuille_error(fd["ux"], fd["flag"], 4, H + 2, H + 2, True, True, L.nu, 1e-6)
    e = np.array([errs[16], errs[32], errs[64]])
    order = np.log(e[:-1] / e[1:]) / np.log(2)
    assert np.all(order > 1.8), (errs, order)
    assert errs[64] < 1e-4


def test_duct_64x32x32_precisions(lbm):
    """The 64x32x32 duct (tau = 1, Fx = 1e-6) itself: FP32 / FP16 / BF16 against FP64 after 20000 steps."""
    kv = dict(nx=64, ny=32, nz=32, tau=1.0, fx=1e-6, walls_y=1, walls_z=1)
    ux = {}
    for prec in PRECS:
        with lbm.Lattice(**kv, precision=prec) as L:
            L.run(20000, 0, lbm.STEP_ALL)
            fd = L.fields()
            m = fd["flag"] == 0
            ux[prec] = fd["ux"][m]
            if prec == "fp64":
                err = A.poiseuille_error(fd["ux"], fd["flag"], 64, 32, 32, True, True, L.nu, 1e-6)
                assert err < 1e-3
    ref = ux["fp64"]
    rel = {p: float(np.linalg.norm(ux[p] - ref) / np.linalg.norm(ref)) for p in PRECS}
    # measured storage error against FP64 (relative L2 of u_x over the fluid)
    assert rel["fp32"] < 1e-5, rel
    assert rel["fp16"] < 2e-2, rel
    assert rel["bf16"] < 1.5e-1, rel


# ------------------------------------------------------------------ squirmer
BULK = dict(nx=40, ny=40, nz=40, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.015, angle=0, planar=1,
            tau=1.0)
U0 = 2 * 0.015 / 3


def swim(L, steps, every=10, lbm=None):
    """Trajectory samples of (t, U, W, e) every `every` steps (driver loop, mass
    correction every 10); the last step writes the moments (like the
    reference's final VTK at step `steps`)."""
    out = []
    t = 0
    while t < steps:
        last = t + every >= steps
        L.run(every, 10, (lbm.STEP_ALL if (last and lbm) else 0))
        t += every
        s = L.squirmer()
        out.append((t, *s["U"], *s["W"], *s["e"]))
    return np.array(out)


@pytest.fixture(scope="module")
def bulk_runs(lbm):
    runs = {}
    for prec in ("fp64", "fp32", "fp16"):
        for beta in (-3.0, 0.0, 3.0):
            with lbm.Lattice(**BULK, beta=beta, precision=prec) as L:
                tr = swim(L, 2000, lbm=lbm)
                fd = L.fields()
                runs[prec, beta] = dict(tr=tr, fields=fd, state=L.squirmer(), diag=L.diagnostics())
    return runs


def late(tr, col):
    return tr[tr[:, 0] > 1000, col]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16"])
def test_bulk_swimming_speed(bulk_runs, prec):
    ratio = late(bulk_runs[prec, 0.0]["tr"], 1).mean() / U0
    assert 0.93 < ratio < 1.03, ratio                                  # test_squirmer_bulk.py:47-49
    assert ratio == pytest.approx(0.985, abs=0.003)                    # tutorial.ipynb:549-563


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16"])
def test_bulk_straight_and_ordered(bulk_runs, prec):
    u = {}
    for beta in (-3.0, 0.0, 3.0):
        tr = bulk_runs[prec, beta]["tr"]
        # test_squirmer_bulk.py:52-57.  FP16 rounding moves the body off the
        # mirror plane by ~1e-4 lattice units, after which a node on one side
        # can be covered a step before its mirror partner: each such event
        # kicks U_y by about the node's momentum over M (~U0 / 100 here) and
        # decays.  For FP16 the reference bounds apply to the means of |U_y|
        # and |W_z|; the single kicks (measured up to U0 / 80 and 5e-2 U0 / R)
        # are capped, and so is the heading they add up to (measured
        # |e_y| <= 2.3e-4 rad over 2000 steps; the bound, 2e-3 rad, is a
        # sideways drift of 0.04 nodes over the run).
        uy, wz = np.abs(late(tr, 2)), np.abs(late(tr, 6))
        if prec != "fp16":
            assert uy.max() < 1e-3 * U0
            assert wz.max() < 1e-6
            assert np.abs(tr[:, 8]).max() < 1e-6                       # e_y: no rotation
        else:
            assert uy.mean() < 1e-3 * U0 and uy.max() < 3e-2 * U0, (uy.mean(), uy.max())
            assert wz.mean() < 1e-5 and wz.max() < 5e-4, (wz.mean(), wz.max())
            assert np.abs(tr[:, 8]).max() < 2e-3
        u[beta] = late(tr, 1).mean()
        assert abs(bulk_runs[prec, beta]["diag"]["mean_density"] - 1.0) < 1e-4
    assert u[-3.0] > u[0.0] > u[3.0]                                   # test_squirmer_bulk.py:60-62
    # notebook values 1.143 / 0.985 / 0.865 (tutorial.ipynb:549-563)
    assert [u[b] / U0 for b in (-3.0, 0.0, 3.0)] == pytest.approx([1.143, 0.985, 0.865], abs=0.004)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_bulk_speed_vs_fp64(bulk_runs, prec):
    for beta in (-3.0, 0.0, 3.0):
        u = late(bulk_runs[prec, beta]["tr"], 1).mean()
        u64 = late(bulk_runs["fp64", beta]["tr"], 1).mean()
        print(f"bulk beta={beta} {prec}: <U>/<U>_fp64 - 1 = {u / u64 - 1:.3e}")
        assert abs(u / u64 - 1) < (2e-4 if prec == "fp32" else 5e-3), (beta, u / u64)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16"])
def test_bulk_near_field_blake(bulk_runs, prec):
    for beta in (-3.0, 0.0, 3.0):
        r = bulk_runs[prec, beta]
        fd = r["fields"]
        e = A.near_field_error(fd["ux"], fd["uy"], fd["flag"], 40, 40, 40, r["state"]["X"], 4.0, 0.015, beta)
        assert e < 0.30, (beta, e)                                     # test_squirmer_bulk.py:70-75
        # tutorial.ipynb:617-619: 0.21 / 0.16 / 0.22
        assert e == pytest.approx({-3.0: 0.21, 0.0: 0.16, 3.0: 0.22}[beta], abs=0.02), (beta, e)


def test_bulk_128_squirmer_speed(lbm):
    """128^3 periodic box, R = 6, B1 = 1.5e-3 (U0 = 1e-3), beta = 0, 5000 steps:
    FP32 / FP16 speed within 0.5 % of FP64; FP64 near Blake's 2 B1 / 3 up to
    the finite-box correction."""
    kv = dict(nx=128, ny=128, nz=128, walls_y=0, walls_z=0, squirmer=1, radius=6.0, b1=1.5e-3, beta=0.0, angle=0,
              planar=1, tau=1.0)
    u = {}
    for prec in ("fp64", "fp32", "fp16", "bf16"):
        with lbm.Lattice(**kv, precision=prec) as L:
            tr = swim(L, 5000, every=50, lbm=lbm)
        u[prec] = tr[tr[:, 0] > 2500, 1].mean()
    u0 = 2 * 1.5e-3 / 3
    print("bulk 128^3 speeds / U0:", {k: v / u0 for k, v in u.items()})
    assert 0.93 < u["fp64"] / u0 < 1.03, u["fp64"] / u0
    assert abs(u["fp32"] / u["fp64"] - 1) < 1e-4
    assert abs(u["fp16"] / u["fp64"] - 1) < 5e-3
    assert abs(u["bf16"] / u["fp64"] - 1) < 5e-2


def test_slit_512x256x128_storage_precisions_against_fp64():
    """The 512x256x128 pusher (the bench workload) for 4000 steps in FP32,
    FP16 and BF16 storage against the FP64 run: squirmer velocity within 1 %
    (FP32, FP16; BF16 within 5 %), measured FP16 ~1e-4."""
    import bench
    import lbm_solver_3d_amd as lbm
    rep = bench.trajectory_errors(lbm, 4000, every=500)
    print(rep)
    assert rep["fp32"]["velocity_rel_err_max"] < 1e-3, rep["fp32"]
    assert rep["fp16"]["velocity_rel_err_max"] < 1e-2, rep["fp16"]
    assert rep["bf16"]["velocity_rel_err_max"] < 5e-2, rep["bf16"]
