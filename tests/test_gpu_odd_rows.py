"""Odd steps in row blocks (aa_odd_row_kernel).

On lattices whose stored rows are whole wavefronts (the row pitch a multiple
of 64, at most 1024: NX a multiple of 64, or rows padded to one) the odd step
of FP16 storage runs whole rows per block (other storages on request): the ten stores with c_x != 0 are
moved one lane by a wavefront shift and the value that crosses a wavefront
edge goes through LDS, so every store is line-aligned.  Only the lane that
issues a store changes, never its address or value:

- FP64 with the row kernel forced on matches the oracle exactly as the tile
  kernel does (bit-identical without a squirmer, 1e-12 with one), for rows of
  one wavefront (the shift wraps within the wavefront), two and sixteen,
  padded rows (the row's last node inside its last wavefront) and several
  rows per block, and equals the tile kernel bit for bit through mass
  corrections;
- FP16 and BF16 are bit-identical with the row kernel and with the odd step
  in tiles, through eager steps, mass corrections, remaps and graph replay;
- a batch runs its odd steps in row blocks on request, every member then
  equal to the same lattice alone.
The environment variable LBM_HIP_ODD_ROWS is read when a lattice is created
(0: odd steps in tiles; 1: in row blocks).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture
def odd_rows(monkeypatch):
    def set_mode(v):
        monkeypatch.setenv("LBM_HIP_ODD_ROWS", str(v))
    return set_mode


def _exact(o, L):
    fl = L.flags()
    assert np.array_equal(fl, o.flags())
    m = fl == 0
    fo, fg = o.f()[:, m], L.populations()[:, m]
    assert np.array_equal(fg, fo), f"max |df| = {np.abs(fg - fo).max()}"
    fd = L.fields()
    for k in ("rho", "ux", "uy", "uz"):
        assert np.array_equal(fd[k][m], o.field(k)[m]), k


@pytest.mark.parametrize("kv", [
    dict(nx=64, ny=9, nz=7, tau=0.7, fx=2e-5, fy=1e-5, walls_y=1, walls_z=1),     # one wavefront per row
    dict(nx=128, ny=6, nz=9, tau=0.9, fx=1e-5, fz=-2e-5, walls_y=0, walls_z=1),   # two
    dict(nx=1024, ny=5, nz=6, tau=0.6, fx=3e-5, walls_y=1, walls_z=0),            # sixteen, the largest block
    dict(nx=50, ny=7, nz=9, tau=0.8, fx=2e-5, fz=1e-5, walls_y=1, walls_z=1),     # padded to 64, 4 rows per block
    dict(nx=100, ny=9, nz=5, tau=0.7, fx=-1e-5, fy=2e-5, walls_y=0, walls_z=1),   # padded to 128, 2 rows per block
    dict(nx=200, ny=5, nz=7, tau=0.9, fx=2e-5, walls_y=1, walls_z=0),             # padded to 256
])
def test_fp64_rows_bit_identical_to_oracle(oracle_mod, lbm, odd_rows, kv):
    odd_rows(1)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert L.info().odd_row_waves == (kv["nx"] + 63) // 64
        for _ in range(43):
            o.step()
            L.step(lbm.STEP_ALL)
        _exact(o, L)


@pytest.mark.parametrize("kv", [
    dict(nx=64, ny=9, nz=7, tau=0.7, fx=2e-5, fy=1e-5, walls_y=1, walls_z=1),
    dict(nx=1024, ny=5, nz=6, tau=0.6, fx=3e-5, walls_y=1, walls_z=0),
])
def test_fp64_rows_equal_tiles_with_mass_correction(lbm, odd_rows, kv):
    """Mass corrections every 10 steps sum the row kernel's per-row
    fixed-point partials: the same state as from the tiles' partials."""
    out = []
    for mode in (0, 1):
        odd_rows(mode)
        with lbm.Lattice(**kv, precision="fp64") as L:
            for t in range(1, 44):
                L.step(lbm.STEP_ALL)
                if t % 10 == 0:
                    L.correct_mass()
            out.append(L.populations())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("nx,x0", [(64, 63.2), (100, 98.8)])
def test_fp64_rows_squirmer_matches_oracle(oracle_mod, lbm, odd_rows, nx, x0):
    """A squirmer across x = 0 on one-wavefront rows and on padded rows: the
    remap, link pass and mass correction after the row kernel's odd steps,
    against the oracle."""
    odd_rows(1)
    kv = dict(nx=nx, ny=22, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4.3, b1=0.03, beta=-2.0, angle=20,
              planar=1, x0=x0)
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        assert L.info().odd_row_waves == (nx + 63) // 64
        for t in range(1, 201):
            o.step(); L.step(lbm.STEP_ALL)
            o.advance_squirmer(); L.advance_squirmer()
            if t % 10 == 0:
                o.correct_mass(); L.correct_mass()
        fl = L.flags()
        assert np.array_equal(fl, o.flags())
        m = fl == 0
        fo, fg = o.f()[:, m], L.populations()[:, m]
        assert np.abs(fg - fo).max() / np.abs(fo).max() < 1e-12
        so, sg = o.state(), L.squirmer()
        for k in ("X", "e"):
            assert np.allclose(sg[k], so[k], rtol=0, atol=1e-12), k
        assert sg["contacts"] == so["contacts"]


def _state(L):
    fl = L.flags()
    out = dict(flag=fl, f=L.populations()[:, fl == 0])
    if L.has_squirmer:
        out["sq"] = L.squirmer()
    return out


@pytest.mark.parametrize("prec,nx", [("fp16", 192), ("bf16", 192), ("fp16", 132)])
def test_rows_equal_tiles_bit_for_bit(lbm, odd_rows, prec, nx):
    kv = dict(nx=nx, ny=48, nz=40, walls_y=0, walls_z=1, squirmer=1, radius=6.0, b1=2e-3, beta=-3.0,
              angle=15.0, planar=1, plane_xz=1, z0=20.0, x0=nx - 1.5, hmin=1.0)
    states = []
    for mode in (0, 1):
        odd_rows(mode)
        with lbm.Lattice(**kv, precision=prec) as L:
            assert L.info().odd_row_waves == (3 if mode else 0)
            for t in range(1, 24):  # eager, the reference's call pattern
                L.step(lbm.STEP_ALL)
                L.advance_squirmer()
                if t % 10 == 0:
                    L.correct_mass()
            L.run(137, 10, lbm.STEP_ALL)  # graph replay
            states.append(_state(L))
    a, b = states
    assert np.array_equal(a["flag"], b["flag"])
    assert np.array_equal(a["f"], b["f"]), float(np.abs(a["f"] - b["f"]).max())
    for k in ("X", "e", "U", "W", "Fh", "Th"):
        assert np.array_equal(a["sq"][k], b["sq"][k]), k


@pytest.mark.parametrize("kv", [
    # a duct with a body force along x and y: the forcing variants of the FP16 row kernel
    dict(nx=256, ny=20, nz=18, tau=0.8, fx=4e-5, fy=-1e-5, walls_y=1, walls_z=1),
    # walls in y, squirmer in the x-y plane near the lower wall, rows padded to 192
    dict(nx=150, ny=40, nz=24, walls_y=1, walls_z=0, squirmer=1, radius=5.0, b1=2e-3, beta=2.0, angle=-30.0,
         planar=1, y0=8.0, hmin=0.5),
])
def test_rows_equal_tiles_forced_and_walls_y(lbm, odd_rows, kv):
    states = []
    for mode in (0, 1):
        odd_rows(mode)
        with lbm.Lattice(**kv, precision="fp16") as L:
            assert L.info().odd_row_waves == ((kv["nx"] + 63) // 64 if mode else 0)
            L.run(101, 10, lbm.STEP_ALL)
            states.append(_state(L))
    a, b = states
    assert np.array_equal(a["flag"], b["flag"])
    assert np.array_equal(a["f"], b["f"]), float(np.abs(a["f"] - b["f"]).max())
    if "sq" in a:
        for k in ("X", "e", "U", "W"):
            assert np.array_equal(a["sq"][k], b["sq"][k]), k


def test_row_kernel_selection(lbm, odd_rows, monkeypatch):
    monkeypatch.delenv("LBM_HIP_ODD_ROWS", raising=False)
    for kv, prec, waves in [
        (dict(nx=512, ny=8, nz=8), "fp16", 8),
        (dict(nx=512, ny=8, nz=8), "bf16", 0),    # 2 % slower in rows: tiles
        (dict(nx=512, ny=8, nz=8), "fp32", 0),    # no faster in rows: tiles
        (dict(nx=512, ny=8, nz=8), "fp64", 0),
        (dict(nx=100, ny=8, nz=8), "fp16", 2),    # rows padded to 128 nodes
        (dict(nx=40, ny=8, nz=8), "fp16", 0),     # rows not padded (64 would add more than half): not whole wavefronts
        (dict(nx=2048, ny=4, nz=4), "fp16", 0),   # more than 1024 threads per row
    ]:
        with lbm.Lattice(**kv, precision=prec) as L:
            info = L.info()
            assert info.odd_row_waves == waves, (kv, prec)
            if waves:
                assert info.occupancy_odd == 0  # the row kernel runs uncapped by default


BATCH_MEMBERS = [
    # slit channels of the ensemble's kind: rows of 2, 3 (padded) and 6 wavefronts
    dict(nx=128, ny=64, nz=32, walls_y=0, walls_z=1, squirmer=1, radius=5.0, b1=1e-3, beta=-2.0, angle=20.0,
         z0=12.0, planar=1, plane_xz=1, hmin=1.0, x0=126.5),
    dict(nx=164, ny=82, nz=41, walls_y=0, walls_z=1, squirmer=1, radius=6.0, b1=1e-3, beta=3.0, angle=40.0,
         z0=25.0, planar=1, plane_xz=1, hmin=1.0),
    dict(nx=384, ny=20, nz=24, walls_y=0, walls_z=1, squirmer=1, radius=4.0, b1=1e-3, beta=0.5, angle=5.0,
         z0=8.0, planar=1, plane_xz=1, hmin=1.0),
    # a duct with a body force and no squirmer, one wavefront per row
    dict(nx=64, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1),
]


@pytest.mark.parametrize("prec", ["fp16", "fp64"])
def test_batch_rows_equal_members_alone(lbm, odd_rows, prec):
    """A batch whose members' rows are whole wavefronts runs its odd steps in
    row blocks on request (the same launch for all members; by default its
    odd steps run in tiles): every member equals the same lattice run alone
    bit for bit."""
    with lbm.Batch(BATCH_MEMBERS, precision=prec) as B:
        assert B.info()["odd_row_blocks"] == 0
    odd_rows(1)
    with lbm.Batch(BATCH_MEMBERS, precision=prec) as B:
        info = B.info()
        assert info["odd_row_blocks"] > 0 and info["odd_row_block_waves"] == 8
        B.run(127, 10, lbm.STEP_ALL)
        got = [_state(B.member(i)) for i in range(len(BATCH_MEMBERS))]
    for kv, a in zip(BATCH_MEMBERS, got):
        with lbm.Lattice(**kv, precision=prec) as L:
            L.run(127, 10, lbm.STEP_ALL)
            b = _state(L)
        assert np.array_equal(a["flag"], b["flag"])
        assert np.array_equal(a["f"], b["f"]), (kv["nx"], float(np.abs(a["f"] - b["f"]).max()))
        if "sq" in a:
            for k in ("X", "e", "U", "W"):
                assert np.array_equal(a["sq"][k], b["sq"][k]), (kv["nx"], k)
