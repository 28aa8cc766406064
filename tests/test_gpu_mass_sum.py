"""Mass-correction bookkeeping on the GPU.

The step kernels emit one excess-mass partial per block as a 128-bit
fixed-point number, so the total does not depend on how nodes are grouped
into tiles and kernels: every tile shape gives the same populations bit for
bit.  The remaps since the last step report their mass change (the first
sets it, later ones add), so correct_mass after a step with partials equals
correct_mass recomputing the mass from the populations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SLIT = dict(nx=96, ny=48, nz=40, walls_y=0, walls_z=1, squirmer=1, radius=6.0, b1=0.01, beta=-3.0, z0=20.0,
            angle=15.0, planar=1, plane_xz=1, eps=0.2, hc=2.0, hmin=1.0)
CORNER = dict(nx=40, ny=36, nz=32, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.05, beta=1.0, x0=2.0,
              y0=34.5, z0=1.0, angle=30.0, planar=0)


def _state(L):
    return dict(f=L.populations(), flag=L.flags(), fields=L.fields(), sq=L.squirmer())


def _assert_same(a, b):
    assert np.array_equal(a["flag"], b["flag"])
    m = a["flag"] == 0
    fa, fb = a["f"][:, m], b["f"][:, m]
    assert np.array_equal(fa, fb), f"max |df| = {np.abs(fa - fb).max()}"
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(a["fields"][k][m], b["fields"][k][m]), k
    for k in ("X", "e", "U", "W", "Fh", "Th"):
        assert np.array_equal(a["sq"][k], b["sq"][k]), k


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "bf16"])
@pytest.mark.parametrize("width,xcd", [(64, 0), (128, 1), (256, 0), (256, 2), (64, 2), (0, 1), (0, 0)])
def test_results_independent_of_tile_shape(lbm, prec, width, xcd):
    """Other tiles, other per-tile mass partials: the same state bit for bit
    after graph-replayed steps with mass corrections."""
    with lbm.Lattice(**SLIT, precision=prec) as A, lbm.Lattice(**SLIT, precision=prec) as B:
        A.set_tiling(64, 1)
        B.set_tiling(width, xcd)
        A.run(43, 10, lbm.STEP_ALL)
        B.run(43, 10, lbm.STEP_ALL)
        _assert_same(_state(A), _state(B))


def test_mass_bookkeeping_of_repeated_remaps(lbm):
    """A step with mass partials, two remaps, correct_mass (partials + link
    change + both remaps' changes) == a step without partials, two remaps,
    correct_mass summing the populations."""
    with lbm.Lattice(**CORNER, precision="fp64") as A, lbm.Lattice(**CORNER, precision="fp64") as B:
        A.run(20, 10, 0)
        B.run(20, 10, 0)
        A.step(lbm.STEP_MASS)
        B.step(0)
        for L in (A, B):
            L.advance_squirmer()
            L.advance_squirmer()
            L.correct_mass()
            L.step(lbm.STEP_ALL)
        m = A.flags() == 0
        assert np.array_equal(A.flags(), B.flags())
        fa, fb = A.populations()[:, m], B.populations()[:, m]
        # the two totals differ by rounding only; leaving out the first
        # remap's mass change would shift the populations by ~1e-7
        assert np.abs(fa - fb).max() <= 1e-13


def test_set_squirmer_rejects_larger_radius(lbm):
    kv = dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.02, planar=1)
    with lbm.Lattice(**kv, precision="fp32") as L:
        with pytest.raises(lbm.LbmHipError):
            L.set_squirmer(R=5.0)
        L.set_squirmer(R=3.5)
        L.run(4, 0, 0)
        L.synchronize()


def test_autotune_between_mass_step_and_correction(lbm):
    """Autotuning after a step with mass partials and before its
    correct_mass: the trials overwrite the partials, so the correction must
    sum the populations instead; the result equals the same calls without
    the autotune to the rounding of the two sums."""
    with lbm.Lattice(**SLIT, precision="fp64") as A, lbm.Lattice(**SLIT, precision="fp64") as B:
        for L in (A, B):
            L.run(29, 10, 0)
            L.step(lbm.STEP_MASS)
            L.advance_squirmer()
        A.autotune(trial_steps=3, mass_every=10)
        for L in (A, B):
            L.correct_mass()
            L.step(lbm.STEP_ALL)
        m = A.flags() == 0
        assert np.array_equal(A.flags(), B.flags())
        fa, fb = A.populations()[:, m], B.populations()[:, m]
        assert np.abs(fa - fb).max() <= 1e-13
        assert A.diagnostics()["mean_density"] == pytest.approx(1.0, abs=1e-12)


def test_unstable_nodes_counted(lbm):
    """A population set to NaN or to a large value makes its node (and,
    after a step, its neighbours) count as unstable; a healthy lattice
    counts none."""
    kv = dict(nx=16, ny=12, nz=10, walls_y=1, walls_z=0, precision="fp64")
    with lbm.Lattice(**kv) as L:
        L.run(3, 0, lbm.STEP_ALL)
        assert L.diagnostics()["unstable_nodes"] == 0
        f = L.populations()
        fl = L.flags()
        n_nan = int(np.flatnonzero(fl == 0)[5])
        n_big = int(np.flatnonzero(fl == 0)[-7])
        f[0, n_nan] = np.nan
        f[0, n_big] = 5.0
        L.upload_populations(f)
        L.step(lbm.STEP_ALL)
        d = L.diagnostics()
        assert d["unstable_nodes"] >= 2
        assert not np.isfinite(d["max_speed"]) or not np.isfinite(d["mean_density"])
