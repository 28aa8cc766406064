"""The device's PARTICLE link list against the links of the flags.

LBM::step (lbm.cpp:179-199) visits every link (node n, direction i) with n
FLUID and n + c_i PARTICLE, found from the flags at every step.  Here the
links live in a list on the device: built from the flags, extended by the
remap with the links its flag changes create, pruned by the link pass, and
compacted when removed entries pile up.  These tests check, after remaps
that moved the sphere, that the live links are exactly the flags' links,
each once: in FP64 against the oracle's flags (with which the GPU's agree
bit for bit) and in FP16 against the GPU's own flags, through the graph-
replayed driver loop, the eager path, the host reading the force between
step and remap, two remaps without a step between them, and runs long enough
for the list to be compacted several times."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# lattice.hpp:20-23 (velocity order of the reference)
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]


def links_of_flags(flag, nx, ny, nz):
    """Brute force over the whole lattice: every (n, i) with n FLUID (0) and
    n + c_i PARTICLE (2), periodic wrap (a walled axis never wraps a FLUID
    node's neighbour)."""
    f = flag.reshape(nz, ny, nx)
    out = set()
    for i in range(1, 19):
        nb = np.roll(f, shift=(-CZ[i], -CY[i], -CX[i]), axis=(0, 1, 2))  # nb[z, y, x] = f[z + cz, y + cy, x + cx]
        for n in np.flatnonzero(((f == 0) & (nb == 2)).reshape(-1)):
            out.add((int(n), i))
    return out


def check_links(L, flag, kv):
    node, d = L.particle_links()
    got = list(zip(node.tolist(), d.tolist()))
    assert len(got) == len(set(got)), "a link is listed twice"
    want = links_of_flags(flag, kv["nx"], kv["ny"], kv["nz"])
    assert set(got) == want, (len(set(got) - want), len(want - set(got)))
    return len(got)


CHANNEL = dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.03, beta=-1, angle=-40,
               planar=1, y0=6.2, hmin=0.5)
XZ = dict(nx=48, ny=20, nz=18, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.03, beta=-2, angle=-30,
          planar=1, plane_xz=1, z0=7.5, hmin=0.5)
CORNER = dict(nx=30, ny=26, nz=28, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.02, beta=2.0, angle=30,
              planar=0, x0=0.4, y0=25.7, z0=0.6)


@pytest.mark.parametrize("kv", [CHANNEL, XZ, CORNER], ids=["walls_y", "plane_xz", "periodic_corner"])
def test_link_list_equals_oracle_links_fp64(oracle_mod, lbm, kv):
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        check_links(L, o.flags(), kv)  # as built at creation
        for t in range(1, 201):
            o.step(); L.step(lbm.STEP_ALL)
            o.advance_squirmer(); L.advance_squirmer()
            if t % 10 == 0:
                o.correct_mass(); L.correct_mass()
            if t % 50 == 0:
                flo = o.flags()
                assert np.array_equal(L.flags(), flo)
                check_links(L, flo, kv)
        # the sphere moved: the remaps changed nodes, so the list was extended and pruned
        assert not np.array_equal(o.flags(), oracle_mod.Oracle(**kv).flags())


def test_link_list_through_graph_and_compaction_fp16(lbm):
    """3000 graph-replayed steps of a fast pusher in FP16: the list is
    extended at every remap that changes nodes and compacted many times."""
    kv = dict(nx=64, ny=40, nz=40, walls_y=0, walls_z=1, squirmer=1, radius=5.5, b1=0.04, beta=-3, angle=20,
              planar=1, plane_xz=1, z0=14.0, hmin=1.0)
    with lbm.Lattice(**kv, precision="fp16") as L:
        f0 = L.flags()
        n0 = check_links(L, f0, kv)
        for _ in range(6):
            L.run(500, 10, lbm.STEP_MACRO)
            n = check_links(L, L.flags(), kv)
            assert abs(n - n0) < 0.25 * n0
        assert not np.array_equal(L.flags(), f0)


def test_link_list_eager_force_reads_and_double_remap_fp64(oracle_mod, lbm):
    """The reference's call pattern with the force read between step and
    remap (the link pass then runs on its own and the remap after it), and
    two remaps in a row without a step (the list is rebuilt from the flags
    before the next link pass): parity and exact links throughout."""
    kv = CHANNEL
    o = oracle_mod.Oracle(**kv)
    with lbm.Lattice(**kv, precision="fp64") as L:
        for t in range(1, 121):
            o.step(); L.step(lbm.STEP_ALL)
            sg = L.squirmer()  # runs the pending link pass
            so = o.state()
            assert np.abs(np.array(sg["Fh"]) - np.array(so["Fh"])).max() <= 1e-9 * max(
                np.abs(np.array(so["Fh"])).max(), 1e-30) + 1e-18
            o.advance_squirmer(); L.advance_squirmer()
            if t % 40 == 0:  # a second remap before the next step
                o.advance_squirmer(); L.advance_squirmer()
            if t % 10 == 0:
                o.correct_mass(); L.correct_mass()
        flo = o.flags()
        assert np.array_equal(L.flags(), flo)
        check_links(L, flo, kv)
        m = flo == 0
        fo, fg = o.f()[:, m], L.populations()[:, m]
        assert np.abs(fg - fo).max() / np.abs(fo).max() < 1e-12
