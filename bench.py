#!/usr/bin/env python3
"""Benchmark: MLUPS and % of HBM peak of the D3Q19 TRT step with squirmer
coupling on the 512x256x128 slit channel (pusher, R = 12) in FP16, FP32 and
FP64 storage, on one MI355X.

A "step" is one full time step of the driver loop (reference main.cpp:60-64):
the fused collide-stream kernels, the device-side squirmer update + remap,
and the mass correction every 10 steps.  Populations, flags and the
squirmer state are resident in HBM when the timed region starts.

Prints one JSON line.  `value` is the FP16-store whole-step MLUPS; the FP32
and FP64 figures and the FP16/FP32 ratio are reported beside it, and the
roofline of the collide-stream update (the bulk kernel plus, on even steps,
the boundary kernel, timed with a HIP event pair on the step's stream).
The CPU baseline is the reference's own step (compiled from its sources
into oracle/_ref) timed on this host's cores on a bounded sample of the
same workload (y and z relabelled, as the reference can express it), or
the oracle restatement when that build is absent.

Other workloads (--workload): the 64-run ensemble, the learned surrogate,
the storage-precision trajectories, the surrogate's training kernels, the
streaming-copy probe and the host transfers.
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published peak
# 512x256x128 slit channel: walls in z, x and y periodic, pusher squirmer
# R = 12, beta = -3, B1 = 1e-3, centre height z = 64, pitched 15 deg toward
# the upper wall (x-z plane), soft wall potential with hard gap 1, F = 0.
WORKLOAD = dict(nx=512, ny=256, nz=128, tau=1.0, walls_y=0, walls_z=1, squirmer=1, radius=12.0, b1=1e-3,
                beta=-3.0, z0=64.0, angle=15.0, planar=1, plane_xz=1, eps=0.2, hc=2.0, hmin=1.0)
MASS_EVERY = 10


def measure(lbm, precision: str, steps: int, warmup: int, tiling=None) -> dict:
    with lbm.Lattice(**WORKLOAD, precision=precision) as L:
        info = L.info()
        if tiling:  # a fixed tile shape (profiling runs)
            L.set_tiling(*tiling)
            tuning = dict(tile_width=tiling[0], xcd_remap=tiling[1], candidates=[])
        else:       # tile shape chosen by timing the candidates (state left unchanged)
            tuning = L.autotune(trial_steps=20, mass_every=MASS_EVERY)
        if warmup:
            L.time_run(warmup, MASS_EVERY)
        L.synchronize()
        t0 = time.perf_counter()
        total_ms, kern_ms = L.time_run(steps, MASS_EVERY)
        L.synchronize()
        wall = time.perf_counter() - t0
        sq = L.squirmer()
    N = int(info.n)
    b_lup = info.bytes_per_lup
    kern_avg_s = kern_ms / steps * 1e-3
    achieved = b_lup * N / kern_avg_s / 1e9
    return dict(
        precision=precision, N=N, steps=steps, total_ms=total_ms, wall_s=wall,
        ms_per_step=total_ms / steps, mlups=N * steps / (total_ms * 1e3),
        kernel_ms=kern_ms / steps, kernel_mlups=N / kern_avg_s / 1e6,
        bytes_per_lup=b_lup, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
        U=list(sq["U"]), X=list(sq["X"]),
        tile_width=tuning["tile_width"] if tuning else None, xcd_remap=tuning["xcd_remap"] if tuning else None,
        tuning=tuning, odd_row_waves=int(info.odd_row_waves),
    )


def trajectory_errors(lbm, steps: int, every: int = 500, precisions=("fp64", "fp32", "fp16", "bf16"),
                      log=None) -> dict:
    """The 512x256x128 pusher run `steps` steps in each storage precision,
    the squirmer state recorded every `every` steps; errors against the FP64
    run: velocity |U - U64| / |U64| (mean and max over the samples after the
    first 2000 steps), position |X - X64| relative to the FP64 distance
    travelled, heading angle."""
    runs = {}
    for prec in precisions:
        t0 = time.perf_counter()
        rows = []
        with lbm.Lattice(**WORKLOAD, precision=prec) as L:
            done = 0
            while done < steps:
                L.run(every, MASS_EVERY, 0)
                done += every
                s = L.squirmer()
                rows.append((done, s["X"], s["U"], s["e"], s["W"]))
        runs[prec] = rows
        if log:
            log(f"[bench] trajectory {prec}: {steps} steps in {time.perf_counter() - t0:.1f} s")
    ref = runs["fp64"]
    X0 = ref[0][1] - ref[0][2] * every  # start (approximately: X before the first window)
    out = {"steps": steps, "every": every,
           "fp64_final": {"X": list(map(float, ref[-1][1])), "U": list(map(float, ref[-1][2]))},
           "fp64_travel": float(np.linalg.norm(ref[-1][1] - X0))}
    for prec, rows in runs.items():
        if prec == "fp64":
            continue
        ve, pe, he = [], [], []
        for (t, X, U, e, _), (_, X64, U64, e64, _) in zip(rows, ref):
            if t < 2000:
                continue
            ve.append(np.linalg.norm(U - U64) / np.linalg.norm(U64))
            pe.append(np.linalg.norm(X - X64))
            he.append(np.degrees(np.arccos(np.clip(np.dot(e, e64), -1.0, 1.0))))
        out[prec] = {"velocity_rel_err_mean": float(np.mean(ve)), "velocity_rel_err_max": float(np.max(ve)),
                     "position_err_final": float(pe[-1]), "position_err_rel_travel": float(pe[-1] / out["fp64_travel"]),
                     "heading_err_deg_final": float(he[-1])}
    return out


def measure_ensemble(lbm, steps: int, warmup: int, count: int = 64, sequential: bool = True, tiling=None) -> dict:
    """The 64-run slit-channel ensemble (seed 0, FP16 store): every phase of a
    step is one launch for all runs, the step cycle graph-replayed, tile shape
    autotuned (or fixed by `tiling`).  Beside it, the same runs advanced one
    after another, each on its own (graph-replayed, autotuned) lattice."""
    members = lbm.ensemble_members(count, 0)
    with lbm.Batch(members, precision="fp16") as B:
        info = B.info()
        if tiling:
            B.set_tiling(*tiling)
            tuning = dict(tile_width=tiling[0], xcd_remap=tiling[1], candidates=[])
        else:
            tuning = B.autotune(trial_steps=10, mass_every=MASS_EVERY)
        if warmup:
            B.time_run(warmup, MASS_EVERY)
        t0 = time.perf_counter()
        ms = B.time_run(steps, MASS_EVERY)
        wall = time.perf_counter() - t0
        info = B.info()
    nodes = int(info["total_nodes"])
    # the collide-stream traffic of one step (19 populations read and written,
    # 2 bytes each) against a streaming copy of as many bytes: the copy's rate
    # falls as the bytes per launch grow
    step_bytes = 2 * 19 * 2 * nodes
    copy_gbs = lbm.copy_bandwidth(step_bytes)
    seq_ms = 0.0
    for kv in (members if sequential else []):
        with lbm.Lattice(**kv, precision="fp16") as L:
            L.autotune(trial_steps=10, mass_every=MASS_EVERY)
            L.time_run(warmup or 10, MASS_EVERY)
            tot, _ = L.time_run(steps, MASS_EVERY)
            seq_ms += tot
    return dict(runs=count, total_nodes=nodes, steps=steps, total_ms=ms, wall_s=wall, ms_per_step=ms / steps,
                mlups=nodes * steps / (ms * 1e3), sequential_ms_per_step=seq_ms / steps if sequential else None,
                sequential_mlups=nodes * steps / (seq_ms * 1e3) if sequential else None, tuning=tuning,
                device_gb=info["device_bytes"] / 1e9, step_bytes=step_bytes, copy_gbs=copy_gbs,
                step_gbs=step_bytes / (ms / steps) / 1e6)


def sweep_ensemble_tiling(lbm, steps: int, count: int = 64) -> dict:
    """Every tile shape the batch tuner tries, `steps` graph-replayed steps
    each after two untimed mass-correction cycles, in the tuner's order (the
    candidates come from one short tuner pass that runs first): per-shape
    step times, and under rocprofv3 per dispatch counters that follow the
    same order."""
    members = lbm.ensemble_members(count, 0)
    rows = []
    with lbm.Batch(members, precision="fp16") as B:
        cands = B.autotune(trial_steps=2, mass_every=MASS_EVERY)["candidates"]
        for c in cands:
            B.set_tiling(c["tile_width"], c["xcd_remap"])
            B.time_run(2 * MASS_EVERY, MASS_EVERY)  # untimed: the new shape's graph capture and first cycles
            ms = B.time_run(steps, MASS_EVERY)
            rows.append(dict(tile_width=c["tile_width"], xcd_remap=c["xcd_remap"], ms_per_step=round(ms / steps, 4)))
            print(f"[sweep] tile width {c['tile_width']}, XCD order {c['xcd_remap']}: {ms / steps:.3f} ms per step",
                  file=sys.stderr, flush=True)
        nodes = int(B.info()["total_nodes"])
    return dict(workload="ensemble tiling sweep", runs=count, total_nodes=nodes, steps=steps, rows=rows)


def measure_mlp(lbm, samples: int = 4096, epochs: int = 20, batch: int = 256) -> dict:
    """Training throughput of the surrogate's networks alone: 4 nets
    5-128-128-128-6 (BF16 operands, FP32 accumulation) on synthetic data."""
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, size=(samples, 5)).astype(np.float32)
    t = np.sin(x @ rng.normal(size=(5, 6))).astype(np.float32)
    with lbm.MLP(nets=4, mode="bf16", seed=1) as m:
        m.train(x, t, epochs=1, batch=batch, lr=2e-3)  # warm-up (the first launch loads the code objects)
        t0 = time.perf_counter()
        loss = m.train(x, t, epochs=epochs, batch=batch, lr=2e-3)
        dt = time.perf_counter() - t0
        m.forward(x)
    return {"metric": "surrogate MLP training throughput, 4 nets 5-128-128-128-6, BF16 MFMA",
            "value": round(4 * samples * epochs / dt, 1), "unit": "sample-updates/s (all 4 nets)", "n_gpus": 1,
            "steps": epochs * (samples // batch), "warmup": samples // batch, "ms_per_step": round(dt * 1e3 / (epochs * (samples // batch)), 4),
            "higher_is_better": True, "dtype": "bf16", "data": "synthetic",
            "config": {"workload": f"{samples} samples, {epochs} epochs, minibatch {batch}"},
            "final_loss": [round(float(v), 6) for v in loss]}


def measure_transfer(lbm, reps: int = 3) -> dict:
    """Host transfers on the 512x256x128 channel (lbm_hip_download /
    lbm_hip_upload_populations through the library's chunked staging):
    the 19 N populations as doubles each way and the four moment fields,
    against the step time, in FP16 and FP64 storage; with a download after
    every step the rate is the PCIe-inclusive one."""
    rows = {}
    for prec in ("fp16", "fp64"):
        with lbm.Lattice(**WORKLOAD, precision=prec) as L:
            L.run(20, MASS_EVERY, lbm.STEP_ALL)
            f = L.populations()
            t0 = time.perf_counter()
            for _ in range(reps):
                f = L.populations()
            td = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                L.upload_populations(f)
            tu = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                L.fields(("rho", "ux", "uy", "uz"))
            tf = (time.perf_counter() - t0) / reps
            L.run(2, MASS_EVERY, lbm.STEP_ALL)
            L.synchronize()
            t0 = time.perf_counter()
            L.run(200, MASS_EVERY)
            L.synchronize()
            ts = (time.perf_counter() - t0) / 200
            nb = 19 * L.N * 8
            rows[prec] = dict(download_s=round(td, 4), download_GBps=round(nb / td / 1e9, 2), upload_s=round(tu, 4),
                              upload_GBps=round(nb / tu / 1e9, 2), fields_s=round(tf, 4), step_ms=round(ts * 1e3, 4),
                              mlups_with_download_every_step=round(L.N / (ts + td) / 1e6, 1))
            print(f"[transfer] {prec}: {json.dumps(rows[prec])}", file=sys.stderr, flush=True)
    return {"workload": "host transfers, 512x256x128 channel", "reps": reps, "rows": rows}


def measure_copy(lbm, reps: int = 10) -> dict:
    """Streaming-copy bandwidth (lbm_hip_copy_bandwidth_parts) against the
    footprint: the bytes one collide-stream launch moves for the 512x256x128
    channel in FP16 (1.27 GB) up to those of the 64-run ensemble (15.2 GB),
    in one pair of buffers and in 64 pairs copied by one launch (the
    ensemble's members are separate allocations)."""
    cases = [(1.275e9, 1), (2.55e9, 1), (5.1e9, 1), (10.2e9, 1), (15.22e9, 1), (15.22e9, 64), (1.275e9, 64)]
    rows = []
    for nbytes, parts in cases:
        g = lbm.copy_bandwidth(int(nbytes), reps=reps, parts=parts)
        rows.append({"bytes": int(nbytes), "parts": parts, "GBps": round(g, 1)})
        print(f"[copy] {nbytes / 1e9:.2f} GB in {parts} part(s): {g:.0f} GB/s", file=sys.stderr, flush=True)
    return {"workload": "copy probe", "reps": reps, "rows": rows}


# The same slit channel as the reference can express it (its orientation
# angle and planar motion are in x-y only, params.hpp:48-50): y and z
# relabelled, walls in y, so the reference runs this very workload.
REFERENCE_DECK = dict(nx=512, ny=128, nz=256, tau=1.0, walls_y=1, walls_z=0, squirmer=1, radius=12.0, b1=1e-3,
                      beta=-3.0, y0=64.0, angle=15.0, planar=1, eps=0.2, hc=2.0, hmin=1.0)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> tuple[int, int]:
    """(threads used, cores this process may run on).  The threads are
    OMP_NUM_THREADS when the environment sets it (a host shared by several
    GPUs gives each one a share of its cores and sets it to that share),
    else every core of the affinity mask."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (max(1, min(int(env), aff)) if env.isdigit() else aff), aff


def cpu_baseline(budget_s: float = 15.0) -> dict | None:
    """The reference's own step (compiled from its sources into oracle/_ref,
    with its Makefile's optimisation) on this host's cores, on a bounded
    sample of the workload: the y<->z-relabelled deck it can express; the
    oracle restatement of the same workload where that build is absent."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        print(f"cpu baseline unavailable: {e}", file=sys.stderr)
        return None
    threads, aff = cpu_threads()
    O.set_threads(threads)
    use_ref = (O.HERE / "_ref" / "libref_lbm_fast.so").exists()
    if use_ref:
        obj = O.Reference(variant="_fast", **REFERENCE_DECK)
        obj.set_threads(threads)
        kind, deck = "reference", "512x128x256 walls in y (the workload with y and z relabelled)"
    else:
        obj = O.Oracle(**WORKLOAD)
        kind, deck = "port", "512x256x128 walls in z"

    def one(n):
        t = time.perf_counter()
        obj.run(n, MASS_EVERY)
        return time.perf_counter() - t

    t1 = one(2)
    per = t1 / 2
    n = max(2, min(400, int(budget_s / max(per, 1e-6))))
    dt = one(n)
    N = 512 * 256 * 128
    return dict(value=N * n / dt / 1e6, unit="MLUPS", cores=threads, kind=kind, cpu_model=cpu_model(),
                affinity_cores=aff,
                sample=f"{n} full steps (step + squirmer update + mass correction every {MASS_EVERY}) of the pusher "
                       f"channel, {deck}, after 2 untimed steps, FP64, {threads} OpenMP threads")


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_aa_step_512x256x128.txt"


def measured_traffic(precision: str, tile_width: int, xcd_remap: int):
    """HBM bytes of the collide-stream update of one step (bulk kernel of
    both parities and the even-step boundary kernel, averaged over the two
    parities) from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/ for one tile shape, scaled by the counters' reading of a copy
    of known size; None when the shape in use was not profiled."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            p = line.split()
            if len(p) == 5 and p[0] == "traffic_bytes_per_step" and p[1] == precision and \
                    (int(p[2]), int(p[3])) == (tile_width, xcd_remap):
                return float(p[4])
    except (OSError, ValueError):
        pass
    return None


ENSEMBLE_TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_ensemble_64runs_fp16.txt"


def measured_ensemble_traffic(tile_width: int, xcd_remap: int):
    """HBM bytes per launch of the batched collide-stream kernel for one tile
    shape (mean of the even and odd launches, which alternate), from the
    counter runs recorded in profiles/; None for a shape not recorded there."""
    try:
        for line in ENSEMBLE_TRAFFIC_FILE.read_text().splitlines():
            p = line.split()
            if len(p) == 4 and p[0] == "traffic_bytes_per_launch" and (int(p[1]), int(p[2])) == (tile_width, xcd_remap):
                return float(p[3])
    except (OSError, ValueError):
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--precisions", default="fp16,fp32,fp64")
    ap.add_argument("--workload", default="channel",
                    choices=["channel", "ensemble", "surrogate", "trajectory", "mlp", "copy", "transfer"],
                    help="channel: the 512x256x128 slit channel (headline); ensemble: 64 slit-channel runs batched; "
                         "surrogate: the learned velocity surrogate with LBM fallback, 1000 trajectories; "
                         "trajectory: the 512x256x128 pusher in FP32 / FP16 / BF16 storage against FP64 "
                         "(--steps steps); mlp: training of the surrogate's 4-net ensemble alone on 4096 "
                         "synthetic samples, 20 epochs of 256-sample minibatches; copy: the streaming-copy "
                         "probe over footprints from 1.27 GB to 15.2 GB, in one allocation pair and in 64; "
                         "transfer: host download / upload rates on the 512x256x128 channel")
    ap.add_argument("--tiling", default=None,
                    help="WIDTH,XCD: use this collide-stream tile shape instead of autotuning (e.g. 256,0; "
                         "width 0: row runs); channel and ensemble workloads.  'sweep' (ensemble): time "
                         "--steps steps of every shape the tuner tries, in its order")
    ap.add_argument("--no-sequential", action="store_true",
                    help="ensemble: skip timing the runs one after another")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (torch) to calibrate rocprofv3 FETCH/WRITE counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU x-slab decomposition is a design "
                                   "only, so --gpus must be 1"}))
        sys.exit(2)

    import lbm_solver_3d_amd as lbm
    lbm.load_library()
    if args.workload == "ensemble":
        print(f"[bench] ensemble: warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
        if args.tiling == "sweep":
            print(json.dumps(sweep_ensemble_tiling(lbm, args.steps)))
            return
        tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else None
        e = measure_ensemble(lbm, args.steps, args.warmup, sequential=not args.no_sequential, tiling=tiling)
        print(json.dumps({
            "metric": "MLUPS, batched ensemble of 64 slit-channel squirmer runs (seed 0), FP16 store",
            "value": round(e["mlups"], 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(e["ms_per_step"], 5), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "64 runs from seed 0: H in [32,96], grid 4Hx2HxH, walls in z, R in [4,8], "
                                   "beta in [-5,5], pitch in [0,45) deg, z0 in [2R,H-2R], B1=1e-3, FP16 store",
                       "total_nodes": e["total_nodes"], "parallelism": "single GPU"},
            "sequential": ({"mlups": round(e["sequential_mlups"], 1),
                            "ms_per_step": round(e["sequential_ms_per_step"], 5)} if e["sequential_mlups"] else None),
            "batch_over_sequential": (round(e["mlups"] / e["sequential_mlups"], 3) if e["sequential_mlups"] else None),
            "tuning": e["tuning"], "device_gb": round(e["device_gb"], 2),
            # whole step (every phase) at the collide-stream bytes: a lower bound of the kernel's rate
            "roofline": {"bound": "hbm", "achieved": round(e["step_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(e["step_gbs"] / HBM_PEAK_GBS, 4),
                         "traffic": measured_ensemble_traffic(e["tuning"]["tile_width"], e["tuning"]["xcd_remap"]),
                         "timing": "whole step", "bytes_per_step": e["step_bytes"],
                         "copy_gbs": round(e["copy_gbs"], 1),
                         "frac_of_copy": round(e["step_gbs"] / e["copy_gbs"], 4)}}))
        return
    if args.workload == "trajectory":
        log = lambda m: print(m, file=sys.stderr, flush=True)
        rep = trajectory_errors(lbm, args.steps, log=log)
        print(json.dumps({"metric": "squirmer velocity / trajectory error of FP32, FP16, BF16 storage against FP64, "
                                    "512x256x128 pusher", "value": rep["fp16"]["velocity_rel_err_mean"],
                          "unit": "relative velocity error (FP16 store, mean)", "n_gpus": 1, "steps": args.steps,
                          "warmup": 0, "higher_is_better": False, "data": "synthetic", "report": rep}))
        return
    if args.workload == "mlp":
        print(json.dumps(measure_mlp(lbm)))
        return
    if args.workload == "copy":
        print(json.dumps(measure_copy(lbm)))
        return
    if args.workload == "transfer":
        print(json.dumps(measure_transfer(lbm)))
        return
    if args.workload == "surrogate":
        from lbm_solver_3d_amd import surrogate as S
        log = lambda m: print(m, file=sys.stderr, flush=True)
        rep = S.study(log=log)
        t = rep["trajectories"]
        traj_s = t["count"] / (t["surrogate_seconds"] + t["fallback_seconds"])
        print(json.dumps({
            "metric": "trajectories/s, learned velocity surrogate with LBM fallback (1000 slit-channel swimmers)",
            "value": round(traj_s, 2), "unit": "trajectories/s", "n_gpus": 1, "steps": t["steps"], "warmup": 0,
            "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
            "data": "synthetic", "config": {"workload": "MLP 5-128-128-128-6 x4, trained on batched LBM runs of "
                                                        "slit channels 4Hx2HxH, walls in z, H in [24,64], R in [4,8], "
                                                        "beta in [-5,5], B1=1e-3; 1000 trajectories of 8000 steps in "
                                                        "windows of 1000, a window with ensemble spread > 2% run by "
                                                        "the LBM",
                                            "parallelism": "single GPU"},
            "report": rep}))
        return
    res = {}
    for prec in args.precisions.split(","):
        print(f"[bench] {prec}: warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
        tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else None
        res[prec] = measure(lbm, prec, args.steps, args.warmup, tiling)
        print(f"[bench] {prec}: {res[prec]['mlups']:.0f} MLUPS, kernel {res[prec]['kernel_ms']:.4f} ms "
              f"({res[prec]['frac'] * 100:.1f}% of HBM peak)", file=sys.stderr, flush=True)
    head = res.get("fp16") or next(iter(res.values()))
    # measured ceiling: a streaming copy (lbm_hip_copy_bandwidth) moving the
    # algorithmic bytes of one collide-stream launch
    copy_gbs = lbm.copy_bandwidth(int(head["bytes_per_lup"] * head["N"]))
    print(f"[bench] streaming copy of the same byte count: {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    # the copy ceiling falls as the bytes per launch grow, so every precision
    # is held against a copy of its own launch's bytes
    copy_of = {}
    for k, v in res.items():
        nb = int(v["bytes_per_lup"] * v["N"])
        copy_of[k] = copy_gbs if v is head else lbm.copy_bandwidth(nb)
        print(f"[bench] {k}: copy of {nb / 1e9:.2f} GB at {copy_of[k]:.0f} GB/s, kernel at "
              f"{v['achieved_gbs'] / copy_of[k] * 100:.1f}% of it", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
    out = {
        "metric": "MLUPS + % HBM peak, D3Q19 TRT 512x256x128; FP16-store vs FP32 at equal error",
        "value": round(head["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "slit channel 512x256x128, walls in z, pusher squirmer R=12 beta=-3 B1=1e-3 at z=64 pitched "
                        "15 deg toward the wall, TRT Lambda=3/16 tau=1, F=0; full time step incl. coupling and "
                        "mass correction every 10 steps",
            "storage": head["precision"],
            "grid": [512, 256, 128],
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(head["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(head["frac"], 4),
            "traffic": measured_traffic(head["precision"], head["tile_width"], head["xcd_remap"]),
            "traffic_source": f"{TRAFFIC_FILE.relative_to(ROOT)} for tiles {head['tile_width']},{head['xcd_remap']}",
            "kernel": ("aa_step_kernel + aa_boundary_kernel (even steps), "
                       + ("aa_odd_row_kernel" if head.get("odd_row_waves") else "aa_step_kernel")
                       + " (odd steps): the collide-stream update of a step"),
            "copy_gbs": round(copy_gbs, 1),
            "frac_of_copy": round(head["achieved_gbs"] / copy_gbs, 4),
            "bytes_per_lup": head["bytes_per_lup"],
            "kernel_ms": round(head["kernel_ms"], 5),
            "tile_width": head["tile_width"],
            "xcd_remap": head["xcd_remap"],
        },
        "cpu_baseline": cpu,
    }
    for k in ("fp32", "fp64"):
        if k in res and res[k] is not head:
            out[k] = {"value": round(res[k]["mlups"], 1), "ms_per_step": round(res[k]["ms_per_step"], 5),
                      "kernel_ms": round(res[k]["kernel_ms"], 5), "bytes_per_lup": res[k]["bytes_per_lup"],
                      "roofline_frac": round(res[k]["frac"], 4), "achieved_gbs": round(res[k]["achieved_gbs"], 1),
                      "copy_gbs": round(copy_of[k], 1), "frac_of_copy": round(res[k]["achieved_gbs"] / copy_of[k], 4),
                      "traffic": measured_traffic(k, res[k]["tile_width"], res[k]["xcd_remap"])}
    if "fp32" in res and "fp16" in res:
        out["fp16_over_fp32"] = round(res["fp16"]["mlups"] / res["fp32"]["mlups"], 3)
    for k, v in res.items():
        out.setdefault("per_precision", {})[k] = {kk: (round(vv, 6) if isinstance(vv, float) else vv)
                                                  for kk, vv in v.items() if kk not in ("U", "X", "tuning")}
        out["per_precision"][k]["copy_gbs"] = round(copy_of[k], 1)
        if v.get("tuning"):
            out["per_precision"][k]["tuning_ms_per_step"] = {
                f"{c['tile_width'] or 'run'}/order{c['xcd_remap']}": round(c["ms_per_step"], 5)
                for c in v["tuning"]["candidates"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
