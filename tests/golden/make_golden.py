#!/usr/bin/env python3
"""Write the golden vectors under tests/golden/ from the reference solver.

The reference's own sources (src/utils/lbm.cpp, squirmer.cpp, params.cpp) are
compiled by oracle/Makefile into oracle/_ref/libref_lbm.so (FMA contraction
off, so the result is independent of the host's FMA units) and driven here,
single-threaded, through oracle/ref_harness.cpp.  The fixtures are data only:
inputs (the key=value configuration) and outputs (populations, moments,
flags, squirmer state histories).  Run from the repository root:

    make -C oracle && python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent

CASES = {
    # plain method: duct with body force, walls in y and z (lbm.cpp step only)
    "duct_4x10x12": dict(kv=dict(nx=4, ny=10, nz=12, tau=0.8, fx=5e-5, walls_y=1, walls_z=1), steps=200,
                         squirmer=False, full_f=True),
    # non-square grid, all three force components, walls in y only
    "chan_f3_13x9x6": dict(kv=dict(nx=13, ny=9, nz=6, tau=0.7, fx=2e-5, fy=-1e-5, fz=3e-6, walls_y=1, walls_z=0),
                           steps=97, squirmer=False, full_f=True),
    # squirmer in a periodic box (moving bounce-back, remap, mass correction)
    "bulk_pusher_24": dict(kv=dict(nx=24, ny=24, nz=24, walls_y=0, walls_z=0, squirmer=1, radius=4, b1=0.015,
                                   beta=-3, angle=20, planar=1), steps=120, squirmer=True, full_f=False),
    # squirmer driven into a y-wall: soft potential and hard-gap contacts
    "wall_contact_40x22x16": dict(kv=dict(nx=40, ny=22, nz=16, walls_y=1, walls_z=0, squirmer=1, radius=4, b1=0.03,
                                          beta=-1, angle=-40, planar=1, y0=6.2), steps=160, squirmer=True,
                                  full_f=False),
    # non-planar motion with walls in y and z
    "duct_puller_3d": dict(kv=dict(nx=30, ny=20, nz=18, walls_y=1, walls_z=1, squirmer=1, radius=3.5, b1=0.02,
                                   beta=2, angle=10, planar=0, y0=9.0, z0=8.3), steps=140, squirmer=True,
                           full_f=False),
}

STATE_KEYS = ["X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw"]


def run_case(name: str, spec: dict) -> dict:
    kv, steps = spec["kv"], spec["steps"]
    r = O.Reference(**kv)
    r.set_threads(1)
    hist = {k: [] for k in STATE_KEYS}
    hist["contacts"], hist["wallForce"] = [], []
    for t in range(1, steps + 1):
        r.step()
        if spec["squirmer"]:
            r.advance_squirmer()
            if t % 10 == 0:
                r.correct_mass()
        s = r.state()
        for k in STATE_KEYS + ["wallForce"]:
            hist[k].append(s[k])
        hist["contacts"].append(s["contacts"])
    fd = r.fields()
    flag = fd["flag"]
    out = dict(
        config=np.array(" ".join(f"{k}={v}" for k, v in kv.items())),
        steps=np.array(steps), mass_every=np.array(10 if spec["squirmer"] else 0),
        flag=flag, rho=fd["rho"], ux=fd["ux"], uy=fd["uy"], uz=fd["uz"],
        diag=np.array([r.diagnostics()[k] for k in ("count_fluid", "mean_density", "max_speed", "sum_ux")]),
    )
    f = fd["f"]
    fluid = flag == 0
    if spec["full_f"]:
        out["f_fluid"] = f[:, fluid]
    else:
        # populations of one z-plane through the sphere and moments of all
        NX, NY = r.NX, r.NY
        zc = int(round(r.state()["X"][2])) % r.NZ
        sl = slice(zc * NX * NY, (zc + 1) * NX * NY)
        out["f_plane"] = f[:, sl]
        out["plane_z"] = np.array(zc)
        ff = f[:, fluid]
        out["f_sum"] = ff.sum(axis=1)
        out["f_sumsq"] = (ff * ff).sum(axis=1)
    for k, v in hist.items():
        out["hist_" + k] = np.array(v)
    return out


def main():
    if not O.reference_available():
        sys.exit("oracle/_ref/libref_lbm.so is not built (needs the reference tree)")
    for name, spec in CASES.items():
        out = run_case(name, spec)
        path = OUT / f"{name}.npz"
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({path.stat().st_size / 1024:.1f} KiB)")


if __name__ == "__main__":
    main()
