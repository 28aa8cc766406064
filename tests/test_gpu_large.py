"""A lattice of more than 2^32 / 19 nodes: population offsets i * SS + n
beyond 32 bits (nodes and in-slot indices stay 32-bit)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DUCT = dict(ny=512, nz=256, tau=0.8, fx=2e-5, walls_y=1, walls_z=1)
BIG_NX = 2048  # 268M nodes; 19 N = 5.1e9


@pytest.mark.parametrize("prec", ["fp16", "fp32"])
def test_large_duct_equals_narrow_duct(lbm, prec):
    """The duct is invariant along x, so every x column of the 268M-node
    lattice must equal the column of a 64-node-wide one bit for bit."""
    assert 19 * BIG_NX * DUCT["ny"] * DUCT["nz"] > 2 ** 32
    with lbm.Lattice(nx=64, **DUCT, precision=prec) as S:
        S.run(12, 10, lbm.STEP_MACRO)
        small = S.fields(("rho", "ux"))
    ref = small["ux"].reshape(DUCT["nz"], DUCT["ny"], 64)[:, :, 0]
    assert np.abs(ref).max() > 0
    with lbm.Lattice(nx=BIG_NX, **DUCT, precision=prec) as L:
        assert L.info().n == BIG_NX * DUCT["ny"] * DUCT["nz"]
        L.run(12, 10, lbm.STEP_MACRO)
        big = L.fields(("rho", "ux"))
    for name in ("ux", "rho"):
        a = big[name].reshape(DUCT["nz"], DUCT["ny"], BIG_NX)
        r = small[name].reshape(DUCT["nz"], DUCT["ny"], 64)[:, :, :1]
        assert np.array_equal(a, np.broadcast_to(r, a.shape)), name


def test_large_channel_with_squirmer(lbm):
    """A squirmer near the top of a 268M-node channel (its slots at the
    highest offsets): mass is kept, the swimmer advances along its axis."""
    kv = dict(nx=BIG_NX, ny=512, nz=256, walls_y=0, walls_z=1, squirmer=1, radius=6.0, b1=0.01, beta=-1.0,
              z0=240.0, angle=0.0, planar=1, plane_xz=1, hmin=1.0)
    with lbm.Lattice(**kv, precision="fp16") as L:
        x0 = np.array(L.squirmer()["X"])
        L.run(40, 10, lbm.STEP_MACRO)
        s = L.squirmer()
        d = L.diagnostics()
    assert np.all(np.isfinite(s["U"])) and np.all(np.isfinite(s["X"]))
    assert s["X"][0] > x0[0]  # swims along +x
    assert abs(d["mean_density"] - 1.0) < 1e-5
