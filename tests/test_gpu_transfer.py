"""Population and field transfers (lbm_hip_download, lbm_hip_upload_populations)
move through a fixed staging array of 2^18 nodes per chunk: a lattice of
336,000 nodes takes two chunks, the second ragged.  Download -> upload ->
download returns the same populations (bit for bit in FP64, to the storage
precision otherwise), and the fields and flags come back whole."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX = dict(nx=80, ny=60, nz=70, walls_y=1, walls_z=0, squirmer=1, radius=6.0, b1=0.02, beta=-2.0, x0=70.0,
           y0=30.0, z0=35.0, angle=20.0, planar=1, eps=0.2, hc=2.0, hmin=1.0)
STAGE_NODES = 1 << 18


@pytest.mark.parametrize("prec,rel", [("fp64", 0.0), ("fp32", 1e-6), ("fp16", 2e-3)])
def test_round_trip_across_chunks(lbm, prec, rel):
    N = BOX["nx"] * BOX["ny"] * BOX["nz"]
    assert STAGE_NODES < N < 2 * STAGE_NODES
    with lbm.Lattice(**BOX, precision=prec) as L:
        L.run(40, 10)
        f1 = L.populations()
        fl = L.flags()
        fd = L.fields()
        fluid = fl == 0
        assert np.isfinite(f1[:, fluid]).all()
        # the chunk seam and the last node carry ordinary values
        for n in (STAGE_NODES - 1, STAGE_NODES, N - 1):
            if fluid[n]:
                assert 0.0 < f1[:, n].sum() < 2.0
        # the fields of the last step (float storage in the reduced modes:
        # converted chunk by chunk), whole and in range
        for k in ("rho", "ux", "uy", "uz"):
            assert np.isfinite(fd[k][fluid]).all(), k
        assert np.abs(fd["rho"][fluid] - 1.0).max() < 0.05
        assert np.abs(fd["ux"][fluid]).max() > 0.0
        assert np.array_equal(fd["flag"], fl)
        L.upload_populations(f1)
        f2 = L.populations()
        d = np.abs(f2[:, fluid] - f1[:, fluid]).max()
        if prec == "fp64":
            assert np.array_equal(f2[:, fluid], f1[:, fluid])
        else:
            assert d <= rel * np.abs(f1[:, fluid]).max(), d
        # the lattice steps on from the uploaded state
        L.run(10, 10)
        f3 = L.populations()
        assert np.isfinite(f3[:, fluid]).all()
        assert abs(L.diagnostics()["mean_density"] - 1.0) < 1e-6
