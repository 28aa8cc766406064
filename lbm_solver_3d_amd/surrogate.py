"""Learned velocity surrogate with the lattice-Boltzmann solver as fallback.

A squirmer in a slit channel (walls in z, motion in the x-z plane) moves, in
the Stokes regime, with a velocity fixed by its configuration.  The
surrogate predicts that velocity from five numbers -- the distance from the
centre to the nearest wall over R, sin and cos of the pitch, beta and R/H --
with an ensemble of four MLPs (include/lbm_mlp.h, trained and evaluated by
the library's HIP kernels) and integrates trajectories with it.  Where the
ensemble disagrees by more than 2 % the trajectory is run by the full
lattice-Boltzmann solver instead (lbm_hip_batch_*).  Training data come from
batched lattice-Boltzmann runs of the same channels.

Wall frame: the configuration is mirrored in z so that the nearest wall lies
below (z' = z if the lower wall z = 0.5 is nearer, z' = -z otherwise;
velocities flip their z component, angular velocities their x and y
components).  Targets are U / U0 and Omega R / U0 in that frame, with
U0 = 2 B1 / 3 (squirmer.hpp:42).
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import STEP_MACRO, Batch, MLP

MASS_EVERY = 10


# ---------------------------------------------------------------------------
# configuration <-> features / targets
# ---------------------------------------------------------------------------
def wall_frame(X, nz):
    """Distance from the centre to the nearest wall and its side (+1: the
    lower wall z = 0.5, -1: the upper wall z = nz - 1.5, halfway bounce-back)."""
    lo, hi = X[2] - 0.5, (nz - 1.5) - X[2]
    return (lo, 1.0) if lo <= hi else (hi, -1.0)


def features(X, e, R, beta, nz) -> np.ndarray:
    d, s = wall_frame(X, nz)
    ez = s * e[2]
    th = math.atan2(ez, e[0])
    return np.array([d / R, math.sin(th), math.cos(th), beta, R / (nz - 2.0)], np.float64)


def targets(X, U, W, R, B1, nz) -> np.ndarray:
    _, s = wall_frame(X, nz)
    U0 = 2.0 * B1 / 3.0
    return np.array([U[0] / U0, U[1] / U0, s * U[2] / U0, s * W[0] * R / U0, s * W[1] * R / U0, W[2] * R / U0])


def to_lab(y, X, R, B1, nz):
    """Wall-frame prediction -> lab-frame U, Omega (lattice units)."""
    _, s = wall_frame(X, nz)
    U0 = 2.0 * B1 / 3.0
    U = np.array([y[0], y[1], s * y[2]]) * U0
    W = np.array([s * y[3], s * y[4], y[5]]) * U0 / R
    return U, W


def free_space(x: np.ndarray) -> np.ndarray:
    """The method without the learned part: a squirmer in unbounded fluid
    swims with U0 along its axis and does not rotate (Blake 1971), whatever
    the walls."""
    y = np.zeros((x.shape[0], 6))
    y[:, 0] = x[:, 2]  # cos(pitch)
    y[:, 2] = x[:, 1]  # sin(pitch)
    return y


# ---------------------------------------------------------------------------
# lattice-Boltzmann runs
# ---------------------------------------------------------------------------
def channel_configs(count: int, seed: int, H=(16, 32), R=(3, 6), beta=(-3.0, 3.0), pitch=(-40.0, 40.0),
                    B1=1e-2) -> list[dict]:
    """Slit channels 4H x 2H x H (walls in z, x and y periodic) with one
    squirmer each, drawn with numpy's default_rng(seed): H, R integers;
    beta, pitch (x-z plane, toward +z when positive) and the centre height
    uniform, the height keeping the gap to either wall at least 1.5 (R at
    most (H - 5) / 2).  B1 = 1e-2 (U0 = 6.7e-3, Reynolds number 2 R U0 / nu
    0.24 - 0.48): the swimmer crosses a lattice spacing in ~150 steps, so
    window means average the staircase of its lattice surface out."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        h = int(rng.integers(H[0], H[1] + 1))
        r = int(rng.integers(R[0], min(R[1], (h - 5) // 2) + 1))  # room for a gap of 1.5 on both sides
        b = float(rng.uniform(*beta))
        p = float(rng.uniform(*pitch))
        lo, hi = 0.5 + r + 1.5, (h - 1.5) - r - 1.5
        z0 = float(rng.uniform(lo, hi))
        out.append(dict(nx=4 * h, ny=2 * h, nz=h, tau=1.0, walls_y=0, walls_z=1, squirmer=1, radius=float(r), b1=B1,
                        beta=b, angle=p, z0=z0, planar=1, plane_xz=1, eps=0.2, hc=2.0, hmin=1.0))
    return out


def lbm_samples(configs: list[dict], sample_at: list[int], precision: str = "fp16"):
    """Run the configurations as one batch and record, for every window
    between consecutive step counts of `sample_at`, the features at the
    window's midpoint and the window-mean velocities: U = dX / dt (exact,
    since X advances by U every step) and Omega_y = -d(pitch) / dt.  The
    instantaneous velocity of a squirmer this small carries the staircase
    of its lattice surface (remap impulses); the window mean does not.
    Windows holding a hard-gap contact (lbm.cpp:245-253) are left out: the
    surrogate's integrator applies the gap itself."""
    xs, ys = [], []
    done, prev = 0, None
    with Batch(configs, precision=precision) as B:
        for t in sample_at:
            B.run(t - done, MASS_EVERY, 0)
            cur = [B.member(i).squirmer() for i in range(len(configs))]
            if prev is not None:
                dt = float(t - done)
                for kv, a, b in zip(configs, prev, cur):
                    if b["contacts"] != a["contacts"]:
                        continue
                    Xm = 0.5 * (a["X"] + b["X"])
                    em = a["e"] + b["e"]
                    em /= np.linalg.norm(em)
                    th_a = math.atan2(a["e"][2], a["e"][0])
                    th_b = math.atan2(b["e"][2], b["e"][0])
                    dth = math.remainder(th_b - th_a, 2.0 * math.pi)
                    U = (b["X"] - a["X"]) / dt
                    W = np.array([0.0, -dth / dt, 0.0])
                    xs.append(features(Xm, em, b["R"], b["beta"], kv["nz"]))
                    ys.append(targets(Xm, U, W, b["R"], b["B1"], kv["nz"]))
            prev, done = cur, t
    return np.array(xs), np.array(ys)


def lbm_trajectories(configs: list[dict], steps: int, every: int, precision: str, warmup: int = 0):
    """Positions and orientations of every squirmer every `every` steps after
    `warmup` steps: arrays (runs, samples, 3)."""
    Xs, es = [], []
    with Batch(configs, precision=precision) as B:
        if warmup:
            B.run(warmup, MASS_EVERY, 0)
        for k in range(steps // every + 1):
            if k:
                B.run(every, MASS_EVERY, 0)
            row = [B.member(i).squirmer() for i in range(len(configs))]
            Xs.append([r["X"] for r in row])
            es.append([r["e"] for r in row])
    return np.array(Xs).transpose(1, 0, 2), np.array(es).transpose(1, 0, 2)


# ---------------------------------------------------------------------------
# the surrogate
# ---------------------------------------------------------------------------
class Surrogate:
    """Normalisation around an ensemble of MLPs."""

    def __init__(self, nets: int = 4, mode: str = "bf16", seed: int = 0):
        self.mlp = MLP(nets=nets, mode=mode, seed=seed)
        self.xm = self.xs = self.ym = self.ys = None

    def close(self):
        self.mlp.close()

    def fit(self, x: np.ndarray, y: np.ndarray, epochs: int = 400, batch: int = 64, lr: float = 2e-3):
        self.xm, self.xs = x.mean(0), x.std(0) + 1e-6
        self.ym, self.ys = y.mean(0), y.std(0) + 1e-6
        return self.mlp.train((x - self.xm) / self.xs, (y - self.ym) / self.ys, epochs, batch, lr)

    def predict(self, x: np.ndarray):
        """Ensemble mean (targets' units) and relative spread per input."""
        xn = (np.atleast_2d(x) - self.xm) / self.xs
        yk = self.mlp.forward(xn) * self.ys + self.ym          # (nets, n, 6)
        mean = yk.mean(0)
        sd = yk.std(0, ddof=1) if yk.shape[0] > 1 else np.zeros_like(mean)
        spread = sd.max(1) / np.maximum(np.abs(mean).max(1), 1e-12)
        return mean, spread

    def trajectories(self, configs: list[dict], X0, e0, steps: int, every: int, spread_limit: float = 0.02):
        """Integrate every configuration from (X0, e0) for `steps` steps, the
        velocity re-evaluated every `every` steps (explicit Euler on X and on
        e, e renormalised, the hard gap applied as in lbm.cpp:245-253).
        Returns positions, orientations (runs, samples, 3) and, per run,
        whether the ensemble spread ever exceeded `spread_limit`."""
        n = len(configs)
        X, e = np.array(X0, float).copy(), np.array(e0, float).copy()
        Xs, es = [X.copy()], [e.copy()]
        uncertain = np.zeros(n, bool)
        for _ in range(steps // every):
            x = np.array([features(X[i], e[i], c["radius"], c["beta"], c["nz"]) for i, c in enumerate(configs)])
            y, sp = self.predict(x)
            uncertain |= sp > spread_limit
            for i, c in enumerate(configs):
                U, W = to_lab(y[i], X[i], c["radius"], c["b1"], c["nz"])
                U[1] = 0.0; W[0] = W[2] = 0.0        # planar motion in x-z (squirmer.cpp:58, 65)
                X[i] = X[i] + U * every
                e[i] = e[i] + np.cross(W, e[i]) * every
                e[i][1] = 0.0
                e[i] /= np.linalg.norm(e[i])
                lo = 0.5 + c["radius"] + c["hmin"]
                hi = (c["nz"] - 1.5) - c["radius"] - c["hmin"]
                X[i][2] = min(max(X[i][2], lo), hi)
            Xs.append(X.copy())
            es.append(e.copy())
        return np.array(Xs).transpose(1, 0, 2), np.array(es).transpose(1, 0, 2), uncertain


def study(n_train_runs=256, n_test_runs=64, n_traj=1000, n_ref=16, seed=0,
          sample_at=tuple(range(3000, 5001, 250)), epochs=400, traj_warmup=3000, traj_steps=4000, every=100,
          fallback_precision="fp16", log=print):
    """The surrogate end to end: training data from batched LBM runs, 4-net
    training on the GPU, held-out error against the free-space model, 1000
    trajectories driven by the surrogate with LBM fallback where the
    ensemble disagrees, and the trajectory error against full-LBM FP64 runs
    of a subset.  Returns a report (dict)."""
    rep = {}
    t0 = time.perf_counter()
    train_cfg = channel_configs(n_train_runs, seed)
    test_cfg = channel_configs(n_test_runs, seed + 1)
    xtr, ytr = lbm_samples(train_cfg, list(sample_at))
    xte, yte = lbm_samples(test_cfg, list(sample_at))
    rep["data"] = dict(train_samples=len(xtr), test_samples=len(xte), seconds=time.perf_counter() - t0)
    log(f"[surrogate] data: {len(xtr)} train / {len(xte)} held-out samples in {rep['data']['seconds']:.1f} s")

    S = Surrogate(nets=4, mode="bf16", seed=seed)
    t1 = time.perf_counter()
    losses = S.fit(xtr, ytr, epochs=epochs)
    log(f"[surrogate] trained 4 nets: normalised loss {[round(float(v), 5) for v in losses]} "
        f"in {time.perf_counter() - t1:.1f} s")
    rep["train"] = dict(loss=[float(v) for v in losses], seconds=time.perf_counter() - t1)
    pred, spread = S.predict(xte)
    rms = lambda a: float(np.sqrt(np.mean(a ** 2)))
    rep["held_out"] = dict(rms_mlp=rms(pred - yte), rms_free_space=rms(free_space(xte) - yte),
                           rms_mlp_speed=rms(pred[:, [0, 2]] - yte[:, [0, 2]]),
                           rms_free_space_speed=rms(free_space(xte)[:, [0, 2]] - yte[:, [0, 2]]),
                           median_spread=float(np.median(spread)))
    log(f"[surrogate] held-out RMS error (targets in U0, U0/R): MLP {rep['held_out']['rms_mlp']:.4f}, "
        f"free-space model {rep['held_out']['rms_free_space']:.4f}")

    # trajectories: warm the flow up with the LBM, then continue with the surrogate
    traj_cfg = channel_configs(n_traj, seed + 2)
    t2 = time.perf_counter()
    X0, e0 = [], []
    for a in range(0, n_traj, 256):
        Xw, ew = lbm_trajectories(traj_cfg[a:a + 256], 0, 1, fallback_precision, warmup=traj_warmup)
        X0 += list(Xw[:, 0]); e0 += list(ew[:, 0])
    t_warm = time.perf_counter() - t2
    t3 = time.perf_counter()
    Xs, es, unc = S.trajectories(traj_cfg, X0, e0, traj_steps, every)
    t_sur = time.perf_counter() - t3
    fb = np.flatnonzero(unc)
    t4 = time.perf_counter()
    for a in range(0, len(fb), 256):
        idx = fb[a:a + 256]
        Xl, el = lbm_trajectories([traj_cfg[i] for i in idx], traj_steps, every, fallback_precision, warmup=traj_warmup)
        Xs[idx], es[idx] = Xl, el
    t_fb = time.perf_counter() - t4
    rep["trajectories"] = dict(count=n_traj, steps=traj_steps, fallback=int(len(fb)), surrogate_seconds=t_sur,
                               fallback_seconds=t_fb, warmup_seconds=t_warm)
    log(f"[surrogate] {n_traj} trajectories x {traj_steps} steps: surrogate {t_sur:.2f} s, "
        f"{len(fb)} fell back to the LBM ({t_fb:.1f} s)")

    # error against full-LBM FP64 runs of the first n_ref trajectories
    ref = list(range(n_ref))
    Xr, er = lbm_trajectories([traj_cfg[i] for i in ref], traj_steps, every, "fp64", warmup=traj_warmup)
    Xf = np.array([X0[i] for i in ref])
    ef = np.array([e0[i] for i in ref])
    # free-space model from the same starting states
    Xb = Xf.copy()
    eb = ef.copy()
    for _ in range(traj_steps // every):
        for j, i in enumerate(ref):
            U0 = 2.0 * traj_cfg[i]["b1"] / 3.0
            Xb[j] = Xb[j] + U0 * eb[j] * every
    travel = np.linalg.norm(Xr[:, -1] - Xr[:, 0], axis=1)
    err = np.linalg.norm(Xs[ref][:, -1] - Xr[:, -1], axis=1)
    err_b = np.linalg.norm(Xb - Xr[:, -1], axis=1)
    ang = np.degrees(np.arccos(np.clip(np.sum(es[ref][:, -1] * er[:, -1], 1), -1, 1)))
    rep["vs_fp64"] = dict(runs=n_ref, mean_travel=float(travel.mean()),
                          mean_position_error=float(err.mean()), max_position_error=float(err.max()),
                          mean_position_error_free_space=float(err_b.mean()),
                          mean_heading_error_deg=float(ang.mean()),
                          fallback_among_them=int(np.sum(unc[ref])))
    log(f"[surrogate] vs FP64 LBM ({n_ref} runs, mean travel {travel.mean():.3f}): position error "
        f"{err.mean():.4f} (free-space model {err_b.mean():.4f}), heading error {ang.mean():.3f} deg")
    S.close()
    return rep
