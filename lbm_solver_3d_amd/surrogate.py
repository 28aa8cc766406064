"""Learned velocity surrogate with the lattice-Boltzmann solver as fallback.

A squirmer in a slit channel (walls in z, motion in the x-z plane) moves, in
the Stokes regime, with a velocity fixed by its configuration.  The
surrogate predicts that velocity from five numbers -- the distance from the
centre to the nearest wall over R, sin and cos of the pitch, beta and R/H --
with an ensemble of four MLPs (include/lbm_mlp.h, trained and evaluated by
the library's HIP kernels) and integrates trajectories with it window by
window.  A window that starts where the ensemble disagrees by more than 2 %
is run by the full lattice-Boltzmann solver instead (lbm_hip_batch_*), from
the trajectory's state at that point.  Training data come from batched
lattice-Boltzmann runs of the same channels.

The channels are the slit channels of the 512 x 256 x 128 pusher benchmark
at smaller sizes: 4H x 2H x H, walls in z, x and y periodic, B1 = 1e-3
(U0 = 6.7e-4), tau = 1, the hard gap 1 and the soft wall potential of that
benchmark.

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
def channel_configs(count: int, seed: int, H=(24, 64), R=(4, 8), beta=(-5.0, 5.0), pitch=(-44.0, 44.0),
                    B1=1e-3) -> list[dict]:
    """Slit channels 4H x 2H x H (walls in z, x and y periodic) with one
    squirmer each, drawn with numpy's default_rng(seed): H, R integers;
    beta, pitch (x-z plane, toward +z when positive) and the centre height
    uniform, the height keeping the gap to either wall at least 1.5.
    B1 = 1e-3 (U0 = 6.7e-4, Reynolds number 2 R U0 / nu 0.03 - 0.06)."""
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


class LBMFollowers:
    """The LBM fallback: a batch of lattices, each following one trajectory
    through the windows where the ensemble is uncertain.  A lattice is built
    from the trajectory's channel with the squirmer at the trajectory's
    current position X and orientation e; during `warmup` steps the body is
    put back at (X, e) every `hold_every` steps (its velocities keep
    relaxing, so the flow settles around the body where it is).  After that
    every `advance` runs free LBM steps and moves the followed trajectories by
    their lattices' displacements, window after window, without another
    warm-up.  A trajectory that the ensemble is sure of again is released:
    its lattice no longer follows it (it keeps running with the batch until
    every member is released)."""

    def __init__(self, configs: list[dict], index, X, e, warmup: int, hold_every: int = 100,
                 precision: str = "fp16"):
        self.index = [int(i) for i in index]
        self.alive = [True] * len(self.index)
        kv, self.ref = [], []
        for c, x in zip(configs, X):
            k = dict(c)
            x0 = np.array([x[0] % c["nx"], x[1] % c["ny"], x[2]])
            k.update(x0=float(x0[0]), y0=float(x0[1]), z0=float(x0[2]), angle=0.0)
            kv.append(k)
            self.ref.append(x0)
        self.B = Batch(kv, precision=precision)
        held = [dict(X=list(x0), e=list(ee)) for x0, ee in zip(self.ref, e)]
        for j, h in enumerate(held):
            self.B.member(j).set_squirmer(**h)
        done = 0
        while done < warmup:
            n = min(hold_every, warmup - done)
            self.B.run(n, MASS_EVERY, 0)
            done += n
            for j, h in enumerate(held):
                self.B.member(j).set_squirmer(**h)

    def release(self, i: int):
        self.alive[self.index.index(i)] = False

    def active(self) -> bool:
        return any(self.alive)

    def advance(self, steps: int, X: np.ndarray, e: np.ndarray):
        """`steps` LBM steps; X[i], e[i] of every followed trajectory i move
        with its lattice (X in the trajectory's unwrapped frame)."""
        self.B.run(steps, MASS_EVERY, 0)
        for j, i in enumerate(self.index):
            if self.alive[j]:
                q = self.B.member(j).squirmer()
                X[i] = X[i] + (q["X"] - self.ref[j])
                e[i] = q["e"]
                self.ref[j] = np.array(q["X"], float)

    def close(self):
        self.B.close()


def _features_v(X, e, R, beta, nz):
    """features() for arrays of configurations."""
    lo, hi = X[:, 2] - 0.5, (nz - 1.5) - X[:, 2]
    sgn = np.where(lo <= hi, 1.0, -1.0)
    d = np.minimum(lo, hi)
    th = np.arctan2(sgn * e[:, 2], e[:, 0])
    return np.stack([d / R, np.sin(th), np.cos(th), beta, R / (nz - 2.0)], 1), sgn


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

    def fit(self, x: np.ndarray, y: np.ndarray, epochs: int = 600, batch: int = 256, lr: float = 5e-4):
        """Normalise inputs and targets and train every net.  The defaults
        were chosen on held-out windows of the slit-channel study: minibatches
        of 64 at lr 2e-3 fit the noise of the window means (train loss lower,
        held-out error higher) and the nets then disagree on 48 % of unseen
        states at the 2 % spread limit, against 3.5 % with these."""
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

    def advance(self, configs: list[dict], X, e, steps: int, every: int):
        """Integrate from (X, e) for `steps` steps with the velocity
        re-evaluated every `every` steps (explicit Euler on X and e, e
        renormalised, planar motion in x-z, the hard gap of lbm.cpp:245-253).
        Returns the new (X, e)."""
        R = np.array([c["radius"] for c in configs])
        beta = np.array([c["beta"] for c in configs])
        nz = np.array([c["nz"] for c in configs], float)
        U0 = np.array([2.0 * c["b1"] / 3.0 for c in configs])
        hmin = np.array([c["hmin"] for c in configs])
        X, e = np.array(X, float).copy(), np.array(e, float).copy()
        for _ in range(steps // every):
            x, sgn = _features_v(X, e, R, beta, nz)
            y, _ = self.predict(x)
            U = np.stack([y[:, 0], np.zeros_like(sgn), sgn * y[:, 2]], 1) * U0[:, None]
            Wy = sgn * y[:, 4] * U0 / R                  # only the rotation about y acts in the x-z plane
            X = X + U * every
            e = e + np.stack([Wy * e[:, 2], np.zeros_like(Wy), -Wy * e[:, 0]], 1) * every
            e[:, 1] = 0.0
            e /= np.linalg.norm(e, axis=1)[:, None]
            X[:, 2] = np.clip(X[:, 2], 0.5 + R + hmin, (nz - 1.5) - R - hmin)
        return X, e

    def trajectories(self, configs: list[dict], X0, e0, steps: int, every: int, window: int,
                     spread_limit: float = 0.02, fallback_warmup: int = 2000, fallback_precision: str = "fp16",
                     batch: int = 128, log=None):
        """Integrate every configuration from (X0, e0) for `steps` steps in
        windows of `window` steps.  A window whose starting configuration has
        an ensemble spread above `spread_limit` is run by the LBM instead of
        the surrogate: by the lattice already following that trajectory if
        the previous window was uncertain too, otherwise by a new one
        (LBMFollowers, `batch` trajectories per batch, warmed up from the
        trajectory's state).  Returns positions and orientations at the
        window ends (runs, windows + 1, 3), the number of windows run by the
        LBM per run, the seconds spent in the surrogate and in the fallback,
        and counts of the fallback's work (lattices started, member-windows
        run).  `log`, if given, receives one line per window."""
        n = len(configs)
        X, e = np.array(X0, float).copy(), np.array(e0, float).copy()
        R = np.array([c["radius"] for c in configs])
        beta = np.array([c["beta"] for c in configs])
        nz = np.array([c["nz"] for c in configs], float)
        Xs, es = [X.copy()], [e.copy()]
        fallback = np.zeros(n, int)
        t_sur = t_fb = 0.0
        groups, follower = [], {}
        started = member_windows = 0
        try:
            for _ in range(steps // window):
                t0 = time.perf_counter()
                _, sp = self.predict(_features_v(X, e, R, beta, nz)[0])
                unc = sp > spread_limit
                for i in [i for i in follower if not unc[i]]:
                    follower.pop(i).release(i)
                ok = np.flatnonzero(~unc)
                if len(ok):
                    X[ok], e[ok] = self.advance([configs[i] for i in ok], X[ok], e[ok], window, every)
                t1 = time.perf_counter()
                for g in [g for g in groups if not g.active()]:
                    g.close()
                    groups.remove(g)
                new = [int(i) for i in np.flatnonzero(unc) if int(i) not in follower]
                for a in range(0, len(new), batch):
                    idx = new[a:a + batch]
                    g = LBMFollowers([configs[i] for i in idx], idx, X[idx], e[idx], fallback_warmup,
                                     precision=fallback_precision)
                    groups.append(g)
                    follower.update({i: g for i in idx})
                    started += len(idx)
                for g in groups:
                    g.advance(window, X, e)
                    member_windows += len(g.index)
                fallback[unc] += 1
                t_sur += t1 - t0
                t_fb += time.perf_counter() - t1
                if log:
                    log(f"[surrogate] window {len(Xs)}: {len(ok)} by the surrogate, {int(unc.sum())} by the LBM "
                        f"({len(new)} new lattices; {t_sur:.2f} s / {t_fb:.1f} s so far)")
                Xs.append(X.copy())
                es.append(e.copy())
        finally:
            for g in groups:
                g.close()
        info = dict(lattices_started=started, member_windows=member_windows)
        return (np.array(Xs).transpose(1, 0, 2), np.array(es).transpose(1, 0, 2), fallback, t_sur, t_fb, info)


def study(n_train_runs=256, n_test_runs=64, n_traj=1000, n_ref=16, seed=0, H=(24, 64), warmup=6000, window=3000,
          windows=10, epochs=600, traj_warmup=3000, traj_steps=8000, every=100, traj_window=1000,
          fallback_warmup=2000, fallback_precision="fp16", log=print):
    """The surrogate end to end: training data from batched LBM runs, 4-net
    training on the GPU, held-out error against the free-space model, 1000
    trajectories driven by the surrogate with LBM fallback over the windows
    where the ensemble disagrees, and the trajectory error against full-LBM
    FP64 runs of a subset.  Returns a report (dict)."""
    rep = {}
    t0 = time.perf_counter()
    sample_at = [warmup + k * window for k in range(windows + 1)]

    def samples(cfg):
        xs, ys = [], []
        for a in range(0, len(cfg), 64):
            x, y = lbm_samples(cfg[a:a + 64], sample_at)
            xs.append(x)
            ys.append(y)
            log(f"[surrogate] data: runs {a}..{min(a + 64, len(cfg)) - 1} of {len(cfg)} done "
                f"({time.perf_counter() - t0:.1f} s)")
        return np.concatenate(xs), np.concatenate(ys)

    train_cfg = channel_configs(n_train_runs, seed, H=H)
    test_cfg = channel_configs(n_test_runs, seed + 1, H=H)
    xtr, ytr = samples(train_cfg)
    xte, yte = samples(test_cfg)
    rep["data"] = dict(train_runs=n_train_runs, train_samples=len(xtr), test_samples=len(xte), H=list(H),
                       warmup=warmup, window=window, windows=windows, seconds=time.perf_counter() - t0)
    log(f"[surrogate] data: {len(xtr)} train / {len(xte)} held-out samples in {rep['data']['seconds']:.1f} s")

    S = Surrogate(nets=4, mode="bf16", seed=seed)
    t1 = time.perf_counter()
    losses = S.fit(xtr, ytr, epochs=epochs)
    log(f"[surrogate] trained 4 nets: normalised loss {[round(float(v), 5) for v in losses]} "
        f"in {time.perf_counter() - t1:.1f} s")
    rep["train"] = dict(loss=[float(v) for v in losses], seconds=time.perf_counter() - t1, epochs=epochs)
    pred, spread = S.predict(xte)
    rms = lambda a: float(np.sqrt(np.mean(a ** 2)))
    rep["held_out"] = dict(rms_mlp=rms(pred - yte), rms_free_space=rms(free_space(xte) - yte),
                           rms_mlp_speed=rms(pred[:, [0, 2]] - yte[:, [0, 2]]),
                           rms_free_space_speed=rms(free_space(xte)[:, [0, 2]] - yte[:, [0, 2]]),
                           median_spread=float(np.median(spread)))
    log(f"[surrogate] held-out RMS error (targets in U0, U0/R): MLP {rep['held_out']['rms_mlp']:.4f}, "
        f"free-space model {rep['held_out']['rms_free_space']:.4f}")

    # trajectories: the flow develops with the LBM, then the surrogate takes over
    traj_cfg = channel_configs(n_traj, seed + 2, H=H)
    t2 = time.perf_counter()
    X0, e0 = [], []
    for a in range(0, n_traj, 128):
        Xw, ew = lbm_trajectories(traj_cfg[a:a + 128], 0, 1, fallback_precision, warmup=traj_warmup)
        X0 += list(Xw[:, 0])
        e0 += list(ew[:, 0])
        log(f"[surrogate] trajectory warm-up: {min(a + 128, n_traj)} of {n_traj} ({time.perf_counter() - t2:.1f} s)")
    X0, e0 = np.array(X0), np.array(e0)
    t_warm = time.perf_counter() - t2
    Xs, es, fb, t_sur, t_fb, info = S.trajectories(traj_cfg, X0, e0, traj_steps, every, traj_window,
                                                   fallback_warmup=fallback_warmup,
                                                   fallback_precision=fallback_precision, log=log)
    nwin = traj_steps // traj_window
    rep["trajectories"] = dict(count=n_traj, steps=traj_steps, window=traj_window, windows=nwin,
                               fallback_windows=int(fb.sum()), runs_with_fallback=int(np.sum(fb > 0)),
                               surrogate_seconds=t_sur, fallback_seconds=t_fb, warmup_seconds=t_warm, **info)
    log(f"[surrogate] {n_traj} trajectories x {traj_steps} steps: surrogate {t_sur:.2f} s, "
        f"{int(fb.sum())} of {n_traj * nwin} windows by the LBM ({t_fb:.1f} s)")

    # error against full-LBM FP64 runs of the first n_ref trajectories
    ref = list(range(n_ref))
    Xr, er = lbm_trajectories([traj_cfg[i] for i in ref], traj_steps, traj_window, "fp64", warmup=traj_warmup)
    # the model without the learned part from the same starting states
    Xb = X0[ref].copy()
    for j, i in enumerate(ref):
        U0 = 2.0 * traj_cfg[i]["b1"] / 3.0
        Xb[j] = Xb[j] + U0 * e0[i] * traj_steps
    travel = np.linalg.norm(Xr[:, -1] - Xr[:, 0], axis=1)
    err = np.linalg.norm(Xs[ref][:, -1] - Xr[:, -1], axis=1)
    err_b = np.linalg.norm(Xb - Xr[:, -1], axis=1)
    start = np.linalg.norm(X0[ref] - Xr[:, 0], axis=1)
    ang = np.degrees(np.arccos(np.clip(np.sum(es[ref][:, -1] * er[:, -1], 1), -1, 1)))
    rep["vs_fp64"] = dict(runs=n_ref, mean_travel=float(travel.mean()), mean_start_offset=float(start.mean()),
                          mean_position_error=float(err.mean()), max_position_error=float(err.max()),
                          mean_position_error_free_space=float(err_b.mean()),
                          mean_heading_error_deg=float(ang.mean()),
                          fallback_windows_among_them=int(fb[ref].sum()))
    log(f"[surrogate] vs FP64 LBM ({n_ref} runs, mean travel {travel.mean():.3f}): position error "
        f"{err.mean():.4f} (free-space model {err_b.mean():.4f}), heading error {ang.mean():.3f} deg")
    S.close()
    return rep
