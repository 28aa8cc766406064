"""The ensemble (lbm_hip_batch_*) and the tile-shape autotuner.

Every member of a batch must advance exactly as the same lattice run alone
through lbm_hip_run (same kernels, bodies and reduction partitions; only the
launches are shared), and autotuning must leave the simulation state as it
found it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = [
    # slit channel with a pusher near the upper wall, pitched in x-z
    dict(nx=40, ny=24, nz=20, walls_y=0, walls_z=1, squirmer=1, radius=4.0, b1=1.5e-2, beta=-3.0, angle=20.0,
         z0=12.0, planar=1, plane_xz=1, hmin=1.0),
    # periodic box with a neutral squirmer
    dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=1.5e-2, beta=0.0, angle=0.0),
    # duct without a squirmer, body force
    dict(nx=4, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1),
    # wide channel whose tiles are 256 nodes long
    dict(nx=300, ny=20, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=5.0, b1=1.5e-2, beta=1.0, angle=10.0),
]


def _state(L):
    fl = L.flags()
    out = dict(flag=fl, f=L.populations()[:, fl == 0])
    if L.has_squirmer:
        out["sq"] = L.squirmer()
    return out


def _same(a, b, rtol=0.0):
    assert np.array_equal(a["flag"], b["flag"])
    if rtol == 0.0:
        assert np.array_equal(a["f"], b["f"]), float(np.abs(a["f"] - b["f"]).max())
    else:
        assert np.allclose(a["f"], b["f"], rtol=rtol, atol=rtol)
    if "sq" in a:
        for k in ("X", "e", "U", "W", "Fh", "Th"):
            if rtol == 0.0:
                assert np.array_equal(a["sq"][k], b["sq"][k]), k
            else:
                assert np.allclose(a["sq"][k], b["sq"][k], rtol=rtol, atol=1e-15), k
        assert a["sq"]["contacts"] == b["sq"]["contacts"]


@pytest.mark.parametrize("prec,width,order", [("fp64", 128, 1), ("fp16", 128, 1), ("fp16", 128, 2),
                                              ("fp32", 128, 0), ("fp16", 0, 1), ("fp64", 0, 0), ("bf16", 64, 2)])
def test_batch_members_equal_separate_runs(lbm, prec, width, order):
    """Tile order 2 permutes the ensemble's tile map on the host (whole tile
    rows per XCD); width 0 is row runs; the members' own runs use the
    default tiles."""
    steps = 37  # eager head, graph-replayed cycles, eager tail with moments
    with lbm.Batch(MEMBERS, precision=prec) as B:
        B.set_tiling(width, order)
        info = B.info()
        assert info["count"] == len(MEMBERS)
        assert info["total_nodes"] == sum(m["nx"] * m["ny"] * m["nz"] for m in MEMBERS)
        B.run(steps, 10, lbm.STEP_ALL)
        got = [_state(B.member(i)) for i in range(len(MEMBERS))]
        assert B.info()["steps_done"] == steps
    for i, kv in enumerate(MEMBERS):
        with lbm.Lattice(**kv, precision=prec) as L:
            L.run(steps, 10, lbm.STEP_ALL)
            _same(got[i], _state(L))


def test_batch_member_fields_after_macro_step(lbm):
    with lbm.Batch(MEMBERS[:2], precision="fp32") as B:
        B.run(12, 10, lbm.STEP_MACRO)
        for i in range(2):
            L = B.member(i)
            d = L.diagnostics()
            assert abs(d["mean_density"] - 1.0) < 1e-5
            assert d["count_fluid"] == int((L.flags() == 0).sum())


def test_batch_rejects_member_advanced_alone(lbm):
    with lbm.Batch(MEMBERS[:2], precision="fp32") as B:
        B.run(3, 10, 0)
        B.member(1).step(0)
        with pytest.raises(lbm.LbmHipError):
            B.run(2, 10, 0)


def test_batch_autotune_keeps_state_and_picks_a_candidate(lbm):
    with lbm.Batch(MEMBERS, precision="fp64") as B:
        B.run(6, 10, 0)
        before = [_state(B.member(i)) for i in range(len(MEMBERS))]
        t = B.autotune(trial_steps=4, mass_every=10)
        assert t["tile_width"] in (0, 64, 128, 256) and t["xcd_remap"] in (0, 1, 2)
        assert len(t["candidates"]) == lbm.TUNE_CANDIDATES
        assert t["ms_per_step"] == min(c["ms_per_step"] for c in t["candidates"])
        info = B.info()
        assert (info["tile_width"], info["xcd_remap"]) == (t["tile_width"], t["xcd_remap"])
        assert info["steps_done"] == 6
        after = [_state(B.member(i)) for i in range(len(MEMBERS))]
        for a, b in zip(before, after):
            _same(a, b)
        B.run(20, 10, 0)
        tuned = [_state(B.member(i)) for i in range(len(MEMBERS))]
    with lbm.Batch(MEMBERS, precision="fp64") as B:
        B.run(26, 10, 0)
        plain = [_state(B.member(i)) for i in range(len(MEMBERS))]
    # the mass partials are summed exactly: any tile shape gives the same state
    for a, b in zip(tuned, plain):
        _same(a, b)


@pytest.mark.parametrize("width,order", [(64, 0), (128, 2), (256, 2), (64, 2), (0, 1), (0, 0)])
def test_tile_shapes_agree_without_squirmer(lbm, width, order):
    kv = dict(nx=70, ny=18, nz=9, tau=0.8, fx=2e-5, walls_y=1, walls_z=1, precision="fp64")
    with lbm.Lattice(**kv) as A, lbm.Lattice(**kv) as B:
        B.set_tiling(width, order)
        A.run(25, 0, lbm.STEP_ALL)
        B.run(25, 0, lbm.STEP_ALL)
        assert np.array_equal(A.populations(), B.populations())


def test_single_autotune_keeps_state(lbm):
    kv = dict(MEMBERS[0], precision="fp16")
    with lbm.Lattice(**kv) as L:
        L.run(5, 10, 0)
        a = _state(L)
        t = L.autotune(trial_steps=4, mass_every=10)
        assert t["tile_width"] in (0, 64, 128, 256)
        _same(a, _state(L))


# Members for the oracle comparison: walls in y, walls in z (x-z plane),
# fully periodic, and one without a squirmer.
ORACLE_MEMBERS = [
    dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4.0, b1=0.03, beta=-1.0, angle=-30.0, y0=8.0,
         planar=1, hmin=0.5),
    dict(nx=36, ny=20, nz=22, walls_y=0, walls_z=1, squirmer=1, radius=4.5, b1=0.02, beta=-3.0, angle=-25.0,
         planar=1, plane_xz=1, z0=9.0),
    dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4.0, b1=0.015, beta=-3.0, angle=20.0, planar=1),
    dict(nx=4, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1),
]


def test_batch_members_match_oracle_fp64(oracle_mod, lbm):
    """Every member of an FP64 batch against the oracle run alone (the
    reference's driver loop, main.cpp:52-87 over lbm.cpp:128-339): 160
    steps with a mass correction every 10, graph-replayed.  Populations to
    1e-12 relative (the force sums run in another order) and flags exactly;
    the member without a squirmer bit for bit."""
    steps = 160
    with lbm.Batch(ORACLE_MEMBERS, precision="fp64") as B:
        B.run(steps, 10, lbm.STEP_ALL)
        got = []
        for i in range(len(ORACLE_MEMBERS)):
            L = B.member(i)
            fl = L.flags()
            got.append(dict(flag=fl, f=L.populations(), sq=L.squirmer() if L.has_squirmer else None,
                            rho=L.fields(("rho",))["rho"]))
    for kv, g in zip(ORACLE_MEMBERS, got):
        o = oracle_mod.Oracle(**kv)
        o.run(steps, 10)
        assert np.array_equal(g["flag"], o.flags())
        m = g["flag"] == 0
        fo, fg = o.f()[:, m], g["f"][:, m]
        if not kv.get("squirmer"):
            assert np.array_equal(fg, fo)
            assert np.array_equal(g["rho"][m], o.field("rho")[m])
            continue
        err = np.abs(fg - fo).max() / np.abs(fo).max()
        assert err < 1e-12, f"population rel err {err:.3e}"
        so = o.state()
        for k in ("X", "e"):
            assert np.allclose(g["sq"][k], so[k], rtol=0, atol=1e-12), k
        for k in ("U", "W"):
            assert np.abs(g["sq"][k] - so[k]).max() <= 1e-9 * max(np.abs(so[k]).max(), 1e-30) + 1e-18, k
        assert g["sq"]["contacts"] == so["contacts"]
        # the remaps moved the sphere off its initial nodes
        assert not np.array_equal(g["flag"], oracle_mod.Oracle(**kv).flags())


def test_c4_ensemble_full_size(lbm):
    """The 64 slit-channel runs of seed 0 (200.2M nodes, FP16 store) in one
    batch, 40 graph-replayed steps and one with moments: every member stays
    finite at unit mean density, and the smallest, the largest and one R = 8
    member equal the same lattice run alone, bit for bit."""
    members = lbm.ensemble_members(64, 0)
    H = [m["nz"] for m in members]
    pick = sorted({int(np.argmin(H)), int(np.argmax(H)),
                   next(i for i, m in enumerate(members) if m["radius"] == 8.0)})
    steps = 41
    with lbm.Batch(members, precision="fp16") as B:
        info = B.info()
        assert info["count"] == 64 and info["total_nodes"] == sum(m["nx"] * m["ny"] * m["nz"] for m in members)
        B.run(steps, 10, lbm.STEP_MACRO)
        for i in range(64):
            d = B.member(i).diagnostics()
            assert d["unstable_nodes"] == 0, i
            assert np.isfinite(d["max_speed"]) and d["max_speed"] > 0, i
            assert abs(d["mean_density"] - 1.0) < 1e-5, (i, d["mean_density"])
            s = B.member(i).squirmer()
            assert np.all(np.isfinite(s["X"])) and np.all(np.isfinite(s["U"])), i
        got = {i: _state(B.member(i)) for i in pick}
    for i in pick:
        with lbm.Lattice(**members[i], precision="fp16") as L:
            L.run(steps, 10, lbm.STEP_MACRO)
            _same(got[i], _state(L))


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_occupancy_caps_leave_results_unchanged(lbm, prec):
    """The occupancy cap only reserves LDS per block: every cap gives the
    same state bit for bit, through eager steps and graph replay."""
    kv = dict(MEMBERS[0], precision=prec)
    defaults = {"fp64": (3, 3), "fp16": (0, 6)}[prec]
    states = []
    free = None  # blocks per CU the registers allow (no cap)
    for caps in ((0, 0), None, (2, 2), (8, 3)):
        with lbm.Lattice(**kv) as L:
            if caps is None:
                caps = defaults
            else:
                L.set_occupancy(*caps)
            info = L.info()
            assert (info.occupancy_even, info.occupancy_odd) == caps
            if free is None:
                free = (info.resident_even, info.resident_odd)
                assert min(free) >= 1
            # the LDS reservation holds each launch to its cap (HIP's occupancy calculator)
            for cap, res, f in zip(caps, (info.resident_even, info.resident_odd), free):
                assert res == (min(cap, f) if cap else f), (caps, res, free)
            L.run(33, 10, lbm.STEP_ALL)
            states.append(_state(L))
    for s in states[1:]:
        _same(states[0], s)


def test_occupancy_cap_rejects_out_of_range(lbm):
    with lbm.Lattice(nx=8, ny=8, nz=8, precision="fp32") as L:
        for caps in ((1, 0), (0, 9), (-1, 4)):
            with pytest.raises(lbm.LbmHipError):
                L.set_occupancy(*caps)
        info = L.info()
        assert (info.occupancy_even, info.occupancy_odd) == (4, 4)


def test_batch_occupancy_cap(lbm):
    with lbm.Batch(MEMBERS, precision="fp16") as A, lbm.Batch(MEMBERS, precision="fp16") as B:
        B.set_occupancy(3, 2)
        A.run(25, 10, 0)
        B.run(25, 10, 0)
        for i in range(len(MEMBERS)):
            a, b = A.member(i).info(), B.member(i).info()
            assert (a.occupancy_even, a.occupancy_odd) == (6, 6)  # the batch default for FP16
            assert (b.occupancy_even, b.occupancy_odd) == (3, 2)
            _same(_state(A.member(i)), _state(B.member(i)))
