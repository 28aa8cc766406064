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


# 1024 x 1024 x 2080 = 2.18e9 nodes; every node of the squirmer's boxes has
# an index above 2^31 (z > 2048).
HUGE = dict(nx=1024, ny=1024, nz=2080, walls_y=0, walls_z=1, squirmer=1, radius=6.0, b1=0.03, beta=-1.0,
            x0=500.25, y0=600.75, z0=2066.625, angle=20.0, planar=1, plane_xz=1, hmin=1.0)
# the same place relative to the upper wall, in a 640 x 640 x 72 channel
# (x and y in the same binades as above, so those coordinates round alike)
NEAR = dict(HUGE, nx=640, ny=640, nz=72, z0=58.625)
STEPS_2_31 = 30


def test_squirmer_at_node_indices_past_2_31(lbm):
    """Without mass correction a node after k steps depends only on the
    nodes within k of it, so for 30 steps the squirmer and the flow around
    it must be those of the same squirmer near the upper wall of a small
    channel: flags of the sphere's neighbourhood identical, state equal to
    the rounding of the z coordinate (2066.6 and 58.6 have different ulps).
    The remap changes nodes in these steps; its node list holds indices
    above 2^31."""
    assert HUGE["nx"] * HUGE["ny"] * int(HUGE["z0"] - HUGE["radius"] - 3) > 2 ** 31
    print(f"[large] {HUGE['nx']}x{HUGE['ny']}x{HUGE['nz']} FP16 squirmer, {STEPS_2_31} steps", flush=True)
    out = {}
    for name, kv in (("near", NEAR), ("huge", HUGE)):
        with lbm.Lattice(**kv, precision="fp16") as L:
            f0 = None
            if name == "near":
                f0 = L.flags()
            L.run(STEPS_2_31, 0, lbm.STEP_MACRO)
            fl = L.flags().reshape(kv["nz"], kv["ny"], kv["nx"])
            X = L.squirmer()["X"]
            # the sphere's neighbourhood: 16 nodes around the centre in x and y,
            # the top 24 layers in z
            cx, cy = int(kv["x0"]), int(kv["y0"])
            box = fl[kv["nz"] - 24:, cy - 16:cy + 16, cx - 16:cx + 16].copy()
            out[name] = dict(box=box, sq=L.squirmer(), d=L.diagnostics(),
                             changed=None if f0 is None else int((L.flags() != f0).sum()))
            del fl
    a, b = out["huge"], out["near"]
    assert b["changed"] > 0, "no remap changed a node"
    assert np.array_equal(a["box"], b["box"])
    assert (a["box"] == 2).sum() > 500  # the sphere is inside the compared box
    dz = HUGE["z0"] - NEAR["z0"]
    assert np.allclose(a["sq"]["X"] - [0, 0, dz], b["sq"]["X"], rtol=0, atol=1e-9)
    for k in ("U", "W", "Fh", "Th"):
        # FP16 storage: a last-bit difference can move one stored rounding
        assert np.allclose(a["sq"][k], b["sq"][k], rtol=1e-4, atol=1e-9 * np.abs(b["sq"][k]).max()), k
    assert a["d"]["unstable_nodes"] == 0
