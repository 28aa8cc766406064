"""Checkpoint and restart (Lattice.save_checkpoint / Lattice.from_checkpoint).

The reference cannot restart a run (its VTK holds no populations,
io.cpp:50-65).  A dump of the populations, flags, squirmer state and step
count restarts one here: a run interrupted at a multiple of the mass
cadence and restarted from its dump continues as the uninterrupted run does
-- flags identical, populations and squirmer state within 1e-12 in FP64 (the
rebuilt link list adds the same force terms in another order), bit for bit
without a squirmer.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(L):
    fl = L.flags()
    out = dict(flag=fl, f=L.populations()[:, fl == 0])
    if L.has_squirmer:
        out["sq"] = L.squirmer()
    return out


@pytest.mark.parametrize("kv", [
    # a puller near a wall, pitched in the x-y plane, with wall contacts
    dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.03, beta=-1, angle=-40, planar=1,
         y0=6.2, hmin=0.5),
    # a pusher in a slit channel, pitched in the x-z plane, across x = 0
    dict(nx=64, ny=24, nz=22, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.02, beta=-3.0, angle=25.0,
         planar=1, plane_xz=1, x0=63.0, z0=11.0, hmin=1.0),
])
def test_restart_continues_the_run_fp64(lbm, tmp_path, kv):
    kv = dict(kv, precision="fp64")
    with lbm.Lattice(**kv) as A:
        A.run(240, 10, lbm.STEP_ALL)
        ref = _state(A)
    path = tmp_path / "state.npz"
    with lbm.Lattice(**kv) as B:
        B.run(120, 10, lbm.STEP_ALL)
        B.save_checkpoint(path)
    with lbm.Lattice.from_checkpoint(path) as R:
        assert R.steps_before == 120
        R.run(120, 10, lbm.STEP_ALL)
        got = _state(R)
    assert np.array_equal(got["flag"], ref["flag"])
    err = np.abs(got["f"] - ref["f"]).max() / np.abs(ref["f"]).max()
    assert err < 1e-12, err
    for k in ("X", "e"):
        assert np.allclose(got["sq"][k], ref["sq"][k], rtol=0, atol=1e-12), k
    for k in ("U", "W"):
        scale = max(np.abs(ref["sq"][k]).max(), 1e-30)
        assert np.abs(got["sq"][k] - ref["sq"][k]).max() <= 1e-9 * scale + 1e-18, k
    assert got["sq"]["contacts"] == ref["sq"]["contacts"]


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_restart_without_squirmer_is_exact(lbm, tmp_path, prec):
    kv = dict(nx=128, ny=20, nz=18, tau=0.8, fx=4e-5, walls_y=1, walls_z=1, precision=prec)
    with lbm.Lattice(**kv) as A:
        A.run(160, 10, lbm.STEP_ALL)
        ref = A.populations()
    path = tmp_path / "duct.npz"
    with lbm.Lattice(**kv) as B:
        B.run(80, 10, lbm.STEP_ALL)
        B.save_checkpoint(path)
    with lbm.Lattice.from_checkpoint(path) as R:
        R.run(80, 10, lbm.STEP_ALL)
        got = R.populations()
        fl = R.flags()
    m = fl == 0
    assert np.array_equal(got[:, m], ref[:, m])


def test_restart_rejects_inconsistent_flags(lbm, tmp_path):
    kv = dict(nx=32, ny=20, nz=20, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.02, beta=0.0, angle=0.0,
              planar=1, precision="fp64")
    path = tmp_path / "bad.npz"
    with lbm.Lattice(**kv) as B:
        B.run(20, 10, lbm.STEP_ALL)
        B.save_checkpoint(path)
    with np.load(path, allow_pickle=False) as d:
        data = {k: d[k] for k in d.files}
    data["sq_X"] = data["sq_X"] + np.array([0.0, 3.0, 0.0])  # a centre the flags do not hold
    np.savez(path, **data)
    with pytest.raises(lbm.LbmHipError):
        lbm.Lattice.from_checkpoint(path)
