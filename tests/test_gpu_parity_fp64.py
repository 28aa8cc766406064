"""GPU FP64 mode against the CPU oracle (restatement of lbm.cpp / squirmer.cpp).

The FP64 mode runs the same AA in-place kernels as the fast modes, with the
reference's operation order and no fused multiply-add.  Without a squirmer
every population, density and velocity must be bit-identical to the oracle;
with a squirmer the force/torque sums are reduced in a different (fixed)
order -- per-block trees and a block-ordered sum instead of one sequential
sum -- so populations are compared at 1e-12 relative and node flags must be
identical.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DUCT = dict(nx=4, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1)


def _fluid(flag):
    return flag == 0


def _compare_exact(o, L):
    fo = o.f()
    fg = L.populations()
    fl = L.flags()
    assert np.array_equal(fl, o.flags())
    m = _fluid(fl)
    assert np.array_equal(fg[:, m], fo[:, m]), f"max |df| = {np.abs(fg[:, m] - fo[:, m]).max()}"
    fd = L.fields()
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(fd[k][m], o.field(k)[m]), k


@pytest.mark.parametrize("steps", [1, 2, 7, 150])
def test_duct_fp64_bit_identical(oracle_mod, lbm, steps):
    o = oracle_mod.Oracle(**DUCT)
    with lbm.Lattice(**DUCT, precision="fp64") as L:
        for _ in range(steps):
            o.step()
            L.step(lbm.STEP_ALL)
        _compare_exact(o, L)
        # the wall force is summed in another order (per-thread, tree, atomics):
        # components agree to round-off of the total
        wf, wo = L.wall_force(), np.array(o.L.wallForce[:])
        assert np.abs(wf - wo).max() <= 1e-11 * np.abs(wo).max()


def test_duct_fp64_non_cubic_tiles(oracle_mod, lbm):
    kv = dict(nx=70, ny=9, nz=6, tau=0.7, fx=2e-5, fy=-1e-5, fz=3e-6, walls_y=1, walls_z=0)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        for _ in range(41):
            o.step()
            L.step(lbm.STEP_ALL)
        _compare_exact(o, L)


def test_duct_fp64_padded_rows(oracle_mod, lbm):
    """Rows of 50 nodes are stored with a pitch of 64 (aligned wavefront
    segments): still bit-identical to the oracle."""
    kv = dict(nx=50, ny=9, nz=7, tau=0.7, fx=2e-5, fy=1e-5, walls_y=1, walls_z=1)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert L.info().row_pitch == 64
        for _ in range(33):
            o.step()
            L.step(lbm.STEP_ALL)
        _compare_exact(o, L)


def _run_pair(o, L, lbm, steps, mass_every=10):
    for t in range(1, steps + 1):
        o.step()
        L.step(lbm.STEP_ALL)
        o.advance_squirmer()
        L.advance_squirmer()
        if mass_every and t % mass_every == 0:
            o.correct_mass()
            L.correct_mass()


def _compare_squirmer(o, L, rtol=1e-12):
    fl = L.flags()
    assert np.array_equal(fl, o.flags())
    m = _fluid(fl)
    fo, fg = o.f()[:, m], L.populations()[:, m]
    err = np.abs(fg - fo).max() / np.abs(fo).max()
    assert err < rtol, f"population rel err {err:.3e}"
    so, sg = o.state(), L.squirmer()
    for k in ("X", "e"):
        assert np.allclose(sg[k], so[k], rtol=0, atol=1e-12), (k, sg[k], so[k])
    for k in ("U", "W"):
        scale = max(np.abs(so[k]).max(), 1e-30)
        assert np.abs(sg[k] - so[k]).max() <= 1e-9 * scale + 1e-18, (k, sg[k], so[k])
    assert sg["contacts"] == so["contacts"]


@pytest.mark.parametrize("beta,steps", [(-3.0, 151), (3.0, 200)])
def test_bulk_squirmer_fp64(oracle_mod, lbm, beta, steps):
    kv = dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4, b1=0.015, beta=beta,
              angle=20, planar=1)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert np.array_equal(L.flags(), o.flags())
        _run_pair(o, L, lbm, steps)
        _compare_squirmer(o, L)


def test_channel_squirmer_near_wall_fp64(oracle_mod, lbm):
    kv = dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.03, beta=-1,
              angle=-40, planar=1, y0=6.2, hmin=0.5)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        _run_pair(o, L, lbm, 333)
        _compare_squirmer(o, L)
        # the wall momentum exchange with the sphere close to the wall (its
        # flow on the rows next to the wall), summed in another order
        wf, wo = L.wall_force(), np.array(o.L.wallForce[:])
        assert np.abs(wf - wo).max() <= 1e-11 * np.abs(wo).max()


@pytest.mark.parametrize("kv,steps", [
    # walls in y and z, odd NX: the link, boundary and remap boxes wrap around x = 0
    (dict(nx=37, ny=23, nz=21, walls_y=1, walls_z=1, squirmer=1, radius=4.2, b1=0.03, beta=-1.5, angle=-25,
          planar=1, x0=1.3, hmin=0.5), 220),
    # fully periodic, the sphere over the corner of the box: every box wraps in x, y and z
    (dict(nx=30, ny=26, nz=28, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.02, beta=2.0, angle=30,
          planar=0, x0=0.4, y0=25.7, z0=0.6), 180),
    # rows of two 64-node chunks: the boundary box wraps around x = 0 across both chunks
    (dict(nx=128, ny=22, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4.3, b1=0.03, beta=-2.0, angle=20,
          planar=1, x0=126.6), 200),
    # the same rows, the box straddling the chunk edge at x = 64
    (dict(nx=128, ny=22, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4.3, b1=0.03, beta=1.0, angle=-15,
          planar=1, x0=64.3), 200),
    # a short last chunk (100 = 64 + 36): a wrapping box touches it and both neighbours
    (dict(nx=100, ny=22, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4.3, b1=0.03, beta=-1.0, angle=10,
          planar=1, x0=98.8), 200),
])
def test_squirmer_across_periodic_boundaries_fp64(oracle_mod, lbm, kv, steps):
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert np.array_equal(L.flags(), o.flags())
        _run_pair(o, L, lbm, steps)
        _compare_squirmer(o, L)


def test_channel_squirmer_padded_rows_fp64(oracle_mod, lbm):
    """Squirmer in a channel whose 100-node rows are stored with a pitch of
    128: links, remap and mass correction address the padded rows."""
    kv = dict(nx=100, ny=20, nz=18, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.03, beta=-2,
              angle=-30, planar=1, plane_xz=1, z0=7.5, hmin=0.5)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert L.info().row_pitch == 128
        _run_pair(o, L, lbm, 160)
        _compare_squirmer(o, L)


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_row_padding_changes_nothing(lbm, monkeypatch, prec):
    kv = dict(nx=100, ny=20, nz=18, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.03, beta=-2, angle=-30,
              planar=1, plane_xz=1, z0=7.5, hmin=0.5, precision=prec)
    out = []
    for pad in ("1", "0"):
        monkeypatch.setenv("LBM_HIP_ROW_PAD", pad)
        with lbm.Lattice(**kv) as L:
            assert L.info().row_pitch == (128 if pad == "1" else 100)
            L.run(57, 10, lbm.STEP_ALL)
            out.append((L.populations(), L.flags(), L.squirmer()["X"]))
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][2], out[1][2])


def test_plane_xz_squirmer_fp64(oracle_mod, lbm):
    kv = dict(nx=36, ny=20, nz=22, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.02, beta=-3,
              angle=-25, planar=1, plane_xz=1, z0=9.0)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        _run_pair(o, L, lbm, 250)
        _compare_squirmer(o, L)


def test_bulk_long_run_fields_fp64(oracle_mod, lbm):
    """40^3 periodic box, neutral squirmer R = 4, 2000 steps of the driver loop:
    flow field, flags and state against the oracle (the reference's own bulk
    acceptance case, tests/test_squirmer_bulk.py:20-40)."""
    import analytic as A
    oracle_mod.set_threads(max(1, min(16, __import__("os").cpu_count() or 1)))
    kv = dict(nx=40, ny=40, nz=40, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.015, beta=0.0, angle=0,
              planar=1)
    o = oracle_mod.Oracle(**kv)
    o.run(2000, 10)
    oracle_mod.set_threads(1)
    with lbm.Lattice(**kv, precision="fp64") as L:
        for t in range(1, 2001):
            L.step(lbm.STEP_ALL if t == 2000 else 0)
            L.advance_squirmer()
            if t % 10 == 0:
                L.correct_mass()
        fd = L.fields()
        sg = L.squirmer()
    assert np.array_equal(fd["flag"], o.flags())
    m = fd["flag"] == 0
    for k in ("ux", "uy", "uz"):
        ref = o.field(k)[m]
        assert np.abs(fd[k][m] - ref).max() <= 1e-9 * np.abs(ref).max(), k
    so = o.state()
    assert np.allclose(sg["X"], so["X"], rtol=0, atol=1e-9)
    e_gpu = A.near_field_error(fd["ux"], fd["uy"], fd["flag"], 40, 40, 40, sg["X"], 4.0, 0.015, 0.0)
    e_ref = A.near_field_error(o.field("ux"), o.field("uy"), o.flags(), 40, 40, 40, so["X"], 4.0, 0.015, 0.0)
    assert e_gpu == pytest.approx(e_ref, rel=1e-6)
    assert e_ref == pytest.approx(0.16, abs=0.01)   # tutorial.ipynb:617-619


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_graph_replay_equals_eager(lbm, prec):
    """lbm_hip_run (HIP-graph replay of the driver loop) == the same calls made one by one."""
    kv = dict(nx=32, ny=20, nz=18, walls_y=1, walls_z=0, squirmer=1, radius=4.0, b1=0.02, beta=-2.0, angle=20,
              precision=prec)
    with lbm.Lattice(**kv) as A_, lbm.Lattice(**kv) as B_:
        A_.run(73, 10, lbm.STEP_ALL)
        for t in range(1, 74):
            B_.step(lbm.STEP_ALL if t == 73 else (lbm.STEP_MASS if t % 10 == 0 else 0))
            B_.advance_squirmer()
            if t % 10 == 0:
                B_.correct_mass()
        assert np.array_equal(A_.flags(), B_.flags())
        m = A_.flags() == 0
        assert np.array_equal(A_.populations()[:, m], B_.populations()[:, m])
        sa, sb = A_.squirmer(), B_.squirmer()
        for k in ("X", "e", "U", "W"):
            assert np.array_equal(sa[k], sb[k]), k


def test_host_nudged_squirmer_fp64(oracle_mod, lbm):
    """The squirmer centre moved by the host between steps (by less than a
    node, as lbm_hip_set_squirmer allows): the next step still sees the old
    flags, whose links may lie outside the box of the new centre, and the
    remap reclassifies around the new one, as the reference does when its
    public member sq is changed."""
    kv = dict(nx=40, ny=24, nz=22, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.02, beta=-2.0, angle=10.0,
              planar=1, plane_xz=1, z0=11.0)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        _run_pair(o, L, lbm, 30)
        X = np.array(o.state()["X"]) + np.array([0.85, 0.0, -0.7])
        for d in range(3):
            o.L.sq.X[d] = X[d]
        L.set_squirmer(X=X)
        with pytest.raises(lbm.LbmHipError):
            L.set_squirmer(X=X + np.array([1.2, 0.0, 0.0]))
        _run_pair(o, L, lbm, 40)
        _compare_squirmer(o, L)
