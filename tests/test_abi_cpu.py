"""C ABI checks that need no GPU: the library loads, exports every entry
point include/lbm_hip.h declares, validates parameters with the reference's
rules (params.cpp:120-140) and refuses to run without a GPU."""
import ctypes as C

import pytest


def test_library_exports_header_symbols(lbm):
    lib = lbm.load_library()
    syms = lbm.exported_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in lbm_hip.h but not exported"


def test_default_params_match_reference(lbm):
    p = lbm.make_params()
    # struct Params defaults, params.hpp:16-56
    assert (p.nx, p.ny, p.nz) == (64, 34, 32)
    assert p.tau == 1.0 and p.lambda_ == 3.0 / 16.0
    assert (p.walls_y, p.walls_z, p.squirmer) == (1, 0, 0)
    assert (p.radius, p.b1, p.beta, p.rho_p) == (4.0, 0.015, 0.0, 1.0)
    assert (p.x0, p.y0, p.z0, p.angle, p.planar) == (-1, -1, -1, 30.0, 1)
    assert (p.eps, p.hc, p.hmin) == (0.2, 2.0, 0.5)


@pytest.mark.parametrize("kv,msg", [
    (dict(nx=0), "grid dimensions must be positive"),
    (dict(ny=3, walls_y=1), "ny must be >= 4 with walls in y"),
    (dict(nz=3, walls_z=1), "nz must be >= 4 with walls in z"),
    (dict(tau=0.5), "tau must be > 0.5"),
    (dict(lambda_=0.0), "lambda must be > 0"),
    (dict(squirmer=1, radius=1.0), "squirmer radius must be > 1 lattice unit"),
    (dict(squirmer=1, angle=45.0), "initial angle must satisfy"),
    (dict(squirmer=1, rho_p=0.0), "rho_p must be > 0"),
    (dict(squirmer=1, hc=0.0), "eps >= 0, hc > 0, hmin >= 0 required"),
    (dict(squirmer=1, y0=4.0), "overlaps (or is too close to) a y-wall"),
    (dict(squirmer=1, nx=9), "domain too small for the squirmer"),
    (dict(squirmer=1, b1=0.09), "swimming speed U0"),
    (dict(nx=2048, ny=2048, nz=1024), "grid too large for 32-bit node indexing"),
])
def test_validate_rejects_like_reference(lbm, kv, msg):
    lib = lbm.load_library()
    p = lbm.make_params()
    for k, v in kv.items():
        setattr(p, k, v)
    rc = lib.lbm_hip_validate(C.byref(p))
    assert rc == -1
    assert msg in lib.lbm_hip_last_error().decode()


def test_validate_accepts_reference_decks(lbm):
    lib = lbm.load_library()
    for kv in [dict(nx=4, ny=34, nz=34, tau=0.8, fx=5e-5, walls_y=1, walls_z=1),
               dict(nx=48, ny=48, nz=48, walls_y=0, squirmer=1, angle=0),
               dict(nx=64, ny=34, nz=32, squirmer=1, angle=30),
               dict(nx=2048, ny=512, nz=256, walls_y=1, walls_z=1)]:  # 19 N > 2^32: 64-bit slot offsets
        p = lbm.make_params(**kv)
        assert lib.lbm_hip_validate(C.byref(p)) == 0, lib.lbm_hip_last_error()


def test_unknown_key_rejected(lbm):
    with pytest.raises(ValueError, match="unknown parameter"):
        lbm.make_params(not_a_key=1)


def test_ensemble_members_are_deterministic_and_valid(lbm):
    """The batched-ensemble workload (64 slit-channel runs from seed 0):
    reproducible draws inside the stated ranges, every member valid."""
    a, b = lbm.ensemble_members(64, 0), lbm.ensemble_members(64, 0)
    assert a == b and len(a) == 64
    assert a != lbm.ensemble_members(64, 1)
    for m in a:
        H, R = m["nz"], m["radius"]
        assert 32 <= H <= 96 and (m["nx"], m["ny"]) == (4 * H, 2 * H)
        assert 4 <= R <= 8 and R == int(R)
        assert -5 <= m["beta"] <= 5 and 0 <= m["angle"] < 45
        assert 2 * R <= m["z0"] <= H - 2 * R
        lbm.validate(**m)


def test_copy_bandwidth_rejects_bad_arguments(lbm):
    lib = lbm.load_library()
    out = C.c_double()
    assert lib.lbm_hip_copy_bandwidth(0, 100, 1, C.byref(out)) == -1
    assert lib.lbm_hip_copy_bandwidth(0, 1 << 20, 0, C.byref(out)) == -1
    assert lib.lbm_hip_copy_bandwidth(0, 1 << 20, 1, None) == -1
