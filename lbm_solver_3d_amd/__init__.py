"""lbm_solver_3d_amd -- MI355X-native D3Q19 TRT lattice Boltzmann step with
squirmer coupling.

The compute path is liblbm_hip.so (hand-written HIP for gfx950) behind the C
ABI declared in include/lbm_hip.h.  This module is a thin ctypes binding of
that ABI, mirroring the reference's class LBM (src/utils/lbm.hpp:19-84):

    >>> L = Lattice(nx=64, ny=34, nz=34, tau=0.8, fx=5e-5, walls_y=1, walls_z=1)
    >>> L.step(); L.advance_squirmer(); L.correct_mass()

There is no CPU fallback: constructing a Lattice without the built library
or without a GPU raises.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
HEADER = ROOT / "include" / "lbm_hip.h"
LIB_PATH = PKG / "liblbm_hip.so"
Q = 19

FP64, FP32, FP16, BF16 = 0, 1, 2, 3
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp16": FP16, "bf16": BF16}
STEP_MACRO, STEP_WALL_FORCE, STEP_MASS = 0x1, 0x2, 0x4
STEP_ALL = STEP_MACRO | STEP_WALL_FORCE | STEP_MASS

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EDEVICE", -4: "ESTATE", -5: "EOVERFLOW"}


class LbmHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    """struct lbm_hip_params (include/lbm_hip.h) == reference struct Params fields."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("tau", C.c_double), ("lambda_", C.c_double),
        ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double),
        ("walls_y", C.c_int), ("walls_z", C.c_int),
        ("squirmer", C.c_int),
        ("radius", C.c_double), ("b1", C.c_double), ("beta", C.c_double), ("rho_p", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("angle", C.c_double),
        ("planar", C.c_int),
        ("eps", C.c_double), ("hc", C.c_double), ("hmin", C.c_double),
        ("plane_xz", C.c_int),
        ("precision", C.c_int),
        ("device", C.c_int),
    ]


class SquirmerState(C.Structure):
    _fields_ = [
        ("R", C.c_double), ("B1", C.c_double), ("beta", C.c_double), ("M", C.c_double), ("I", C.c_double),
        ("FS", C.c_double),
        ("X", C.c_double * 3), ("e", C.c_double * 3), ("U", C.c_double * 3), ("W", C.c_double * 3),
        ("Fh", C.c_double * 3), ("Th", C.c_double * 3), ("Fc", C.c_double * 3), ("Tc", C.c_double * 3),
        ("Fw", C.c_double * 3),
        ("eps", C.c_double), ("hc", C.c_double), ("hmin", C.c_double),
        ("contacts", C.c_int64),
        ("planar", C.c_int32), ("plane_xz", C.c_int32),
    ]


class Diag(C.Structure):
    _fields_ = [("count_fluid", C.c_int64), ("mean_density", C.c_double), ("max_speed", C.c_double),
                ("sum_ux", C.c_double), ("unstable_nodes", C.c_int64)]


class Info(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("precision", C.c_int),
        ("n", C.c_uint64), ("steps_done", C.c_int64),
        ("tau", C.c_double), ("om_plus", C.c_double), ("om_minus", C.c_double), ("nu", C.c_double),
        ("bytes_per_lup", C.c_double), ("device_bytes", C.c_uint64),
        ("has_squirmer", C.c_int), ("walls_y", C.c_int), ("walls_z", C.c_int),
        ("count_fluid", C.c_int64),
        ("row_pitch", C.c_int), ("tile_width", C.c_int), ("xcd_remap", C.c_int),
        ("occupancy_even", C.c_int), ("occupancy_odd", C.c_int),
        ("resident_even", C.c_int), ("resident_odd", C.c_int), ("odd_row_waves", C.c_int),
    ]


# Reference parameter keys and aliases (params.cpp:40-70) -> Params fields.
KEY_ALIASES = {
    "nx": "nx", "ny": "ny", "nz": "nz", "tau": "tau", "lambda": "lambda_",
    "fx": "fx", "fy": "fy", "fz": "fz",
    "walls_y": "walls_y", "wallsy": "walls_y", "walls_z": "walls_z", "wallsz": "walls_z",
    "squirmer": "squirmer", "sq": "squirmer", "radius": "radius", "r": "radius",
    "b1": "b1", "beta": "beta", "rho_p": "rho_p", "rhop": "rho_p",
    "x0": "x0", "y0": "y0", "z0": "z0", "angle": "angle", "planar": "planar",
    "eps": "eps", "hc": "hc", "hmin": "hmin",
    "plane_xz": "plane_xz", "precision": "precision", "device": "device",
}

_lib = None


def exported_symbols() -> list[str]:
    """Entry points declared in the C headers under include/."""
    import re
    names = set()
    for h in sorted(HEADER.parent.glob("*.h")):
        names |= set(re.findall(r"\b(lbm_(?:hip|mlp)_\w+)\s*\(", h.read_text()))
    return sorted(names)


TUNE_CANDIDATES = 11


class Tuning(C.Structure):
    _fields_ = [("tile_width", C.c_int), ("xcd_remap", C.c_int), ("ms_per_step", C.c_double),
                ("cand_tile_width", C.c_int * TUNE_CANDIDATES), ("cand_xcd_remap", C.c_int * TUNE_CANDIDATES),
                ("cand_ms_per_step", C.c_double * TUNE_CANDIDATES)]

    def as_dict(self) -> dict:
        return dict(tile_width=self.tile_width, xcd_remap=self.xcd_remap, ms_per_step=self.ms_per_step,
                    candidates=[dict(tile_width=self.cand_tile_width[k], xcd_remap=self.cand_xcd_remap[k],
                                     ms_per_step=self.cand_ms_per_step[k]) for k in range(TUNE_CANDIDATES)])


class BatchInfo(C.Structure):
    _fields_ = [("count", C.c_int), ("precision", C.c_int), ("steps_done", C.c_longlong),
                ("total_nodes", C.c_longlong), ("total_tiles", C.c_int), ("boundary_blocks", C.c_int),
                ("link_blocks", C.c_int), ("tile_width", C.c_int), ("xcd_remap", C.c_int),
                ("device_bytes", C.c_size_t), ("odd_row_blocks", C.c_int), ("odd_row_block_waves", C.c_int)]


def load_library(path: Path | None = None):
    """Load liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    P, ctx = C.POINTER(Params), C.c_void_p
    dp = C.POINTER(C.c_double)
    lib.lbm_hip_default_params.argtypes = [P]
    lib.lbm_hip_default_params.restype = None
    lib.lbm_hip_validate.argtypes = [P]
    lib.lbm_hip_create.argtypes = [P, C.POINTER(C.c_void_p)]
    lib.lbm_hip_destroy.argtypes = [ctx]
    lib.lbm_hip_destroy.restype = None
    lib.lbm_hip_step.argtypes = [ctx, C.c_int]
    lib.lbm_hip_advance_squirmer.argtypes = [ctx]
    lib.lbm_hip_correct_mass.argtypes = [ctx]
    lib.lbm_hip_run.argtypes = [ctx, C.c_int, C.c_int, C.c_int]
    lib.lbm_hip_get_squirmer.argtypes = [ctx, C.POINTER(SquirmerState)]
    lib.lbm_hip_set_squirmer.argtypes = [ctx, C.POINTER(SquirmerState)]
    lib.lbm_hip_get_wall_force.argtypes = [ctx, dp]
    lib.lbm_hip_diagnostics.argtypes = [ctx, C.POINTER(Diag)]
    lib.lbm_hip_download.argtypes = [ctx, dp, dp, dp, dp, dp, C.POINTER(C.c_uint8)]
    lib.lbm_hip_upload_populations.argtypes = [ctx, dp]
    lib.lbm_hip_synchronize.argtypes = [ctx]
    lib.lbm_hip_get_info.argtypes = [ctx, C.POINTER(Info)]
    lib.lbm_hip_particle_links.argtypes = [ctx, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), C.c_int64,
                                           C.POINTER(C.c_int64)]
    lib.lbm_hip_time_run.argtypes = [ctx, C.c_int, C.c_int, dp, dp]
    lib.lbm_hip_autotune.argtypes = [ctx, C.c_int, C.c_int, C.POINTER(Tuning)]
    lib.lbm_hip_set_tiling.argtypes = [ctx, C.c_int, C.c_int]
    lib.lbm_hip_set_occupancy.argtypes = [ctx, C.c_int, C.c_int]
    bt = C.c_void_p
    lib.lbm_hip_batch_create.argtypes = [P, C.c_int, C.POINTER(C.c_void_p)]
    lib.lbm_hip_batch_destroy.argtypes = [bt]
    lib.lbm_hip_batch_destroy.restype = None
    lib.lbm_hip_batch_run.argtypes = [bt, C.c_int, C.c_int, C.c_int]
    lib.lbm_hip_batch_synchronize.argtypes = [bt]
    lib.lbm_hip_batch_member.argtypes = [bt, C.c_int, C.POINTER(C.c_void_p)]
    lib.lbm_hip_batch_info_get.argtypes = [bt, C.POINTER(BatchInfo)]
    lib.lbm_hip_batch_autotune.argtypes = [bt, C.c_int, C.c_int, C.POINTER(Tuning)]
    lib.lbm_hip_batch_set_tiling.argtypes = [bt, C.c_int, C.c_int]
    lib.lbm_hip_batch_set_occupancy.argtypes = [bt, C.c_int, C.c_int]
    lib.lbm_hip_batch_time_run.argtypes = [bt, C.c_int, C.c_int, dp]
    lib.lbm_hip_copy_bandwidth.argtypes = [C.c_int, C.c_size_t, C.c_int, dp]
    lib.lbm_hip_copy_bandwidth_parts.argtypes = [C.c_int, C.c_size_t, C.c_int, C.c_int, dp]
    fp = C.POINTER(C.c_float)
    mh = C.c_void_p
    lib.lbm_mlp_create.argtypes = [C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.POINTER(C.c_void_p)]
    lib.lbm_mlp_destroy.argtypes = [mh]
    lib.lbm_mlp_destroy.restype = None
    lib.lbm_mlp_get_params.argtypes = [mh, C.c_int, fp]
    lib.lbm_mlp_set_params.argtypes = [mh, C.c_int, fp]
    lib.lbm_mlp_forward.argtypes = [mh, fp, C.c_int, fp]
    lib.lbm_mlp_predict.argtypes = [mh, fp, C.c_int, fp, fp]
    lib.lbm_mlp_predict_noalloc.argtypes = [mh, fp, C.c_int, fp, fp]
    lib.lbm_mlp_loss_grad.argtypes = [mh, C.c_int, fp, fp, C.c_int, fp, fp]
    lib.lbm_mlp_train.argtypes = [mh, fp, fp, C.c_int, C.c_int, C.c_int, C.c_float, fp]
    lib.lbm_hip_last_error.restype = C.c_char_p
    lib.lbm_hip_last_error.argtypes = []
    if path is None:
        _lib = lib
    return lib


def _check(lib, rc: int):
    if rc != 0:
        raise LbmHipError(rc, lib.lbm_hip_last_error().decode())


def copy_bandwidth(nbytes: int, reps: int = 20, device: int = 0, parts: int = 1) -> float:
    """Device streaming-copy bandwidth in GB/s (bytes read + written) moving
    nbytes in total per copy, over `parts` pairs of separate buffers copied
    in one launch (lbm_hip_copy_bandwidth_parts)."""
    lib = load_library()
    out = C.c_double()
    _check(lib, lib.lbm_hip_copy_bandwidth_parts(int(device), int(nbytes), int(parts), int(reps), C.byref(out)))
    return out.value


def make_params(**kv) -> Params:
    lib = load_library()
    p = Params()
    lib.lbm_hip_default_params(C.byref(p))
    for k, v in kv.items():
        k = k.lower()
        if k not in KEY_ALIASES:
            raise ValueError(f"unknown parameter '{k}'")
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v.lower()]
        setattr(p, KEY_ALIASES[k], v)
    return p


def validate(**kv) -> None:
    """Params::validate (params.cpp:120-140); raises LbmHipError(EINVAL)."""
    lib = load_library()
    p = make_params(**kv)
    _check(lib, lib.lbm_hip_validate(C.byref(p)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class Lattice:
    """One lattice on one GPU: the reference's class LBM over the C ABI."""

    def __init__(self, _handle=None, **kv):
        self.lib = load_library()
        self.owned = _handle is None
        self.kv = dict(kv)  # creation parameters (a checkpoint records them)
        self.steps_before = 0  # steps of the run a checkpoint restarted from
        if self.owned:
            self.params = make_params(**kv)
            h = C.c_void_p()
            _check(self.lib, self.lib.lbm_hip_create(C.byref(self.params), C.byref(h)))
            self.h = h
        else:  # a member of a Batch: a view, freed with the batch
            self.h = _handle
        info = self.info()
        self.NX, self.NY, self.NZ, self.N = info.nx, info.ny, info.nz, int(info.n)
        self.nu, self.omPlus, self.omMinus = info.nu, info.om_plus, info.om_minus
        self.has_squirmer = bool(info.has_squirmer)

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.lib.lbm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- reference entry points ----
    def step(self, flags: int = STEP_ALL):
        _check(self.lib, self.lib.lbm_hip_step(self.h, flags))

    def advance_squirmer(self):
        _check(self.lib, self.lib.lbm_hip_advance_squirmer(self.h))

    def correct_mass(self):
        _check(self.lib, self.lib.lbm_hip_correct_mass(self.h))

    def run(self, steps: int, mass_every: int = 10, last_flags: int = STEP_ALL):
        _check(self.lib, self.lib.lbm_hip_run(self.h, steps, mass_every, last_flags))

    def synchronize(self):
        _check(self.lib, self.lib.lbm_hip_synchronize(self.h))

    # ---- data ----
    def info(self) -> Info:
        o = Info()
        _check(self.lib, self.lib.lbm_hip_get_info(self.h, C.byref(o)))
        return o

    def squirmer(self) -> dict:
        s = SquirmerState()
        _check(self.lib, self.lib.lbm_hip_get_squirmer(self.h, C.byref(s)))
        out = {k: np.array(getattr(s, k)[:]) for k in ("X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw")}
        for k in ("R", "B1", "beta", "M", "I", "FS", "eps", "hc", "hmin"):
            out[k] = getattr(s, k)
        out["contacts"] = int(s.contacts)
        return out

    def set_squirmer(self, **changes):
        s = SquirmerState()
        _check(self.lib, self.lib.lbm_hip_get_squirmer(self.h, C.byref(s)))
        for k, v in changes.items():
            if isinstance(getattr(s, k), C.Array):
                getattr(s, k)[:] = list(v)
            else:
                setattr(s, k, v)
        _check(self.lib, self.lib.lbm_hip_set_squirmer(self.h, C.byref(s)))

    def wall_force(self) -> np.ndarray:
        o = np.zeros(3)
        _check(self.lib, self.lib.lbm_hip_get_wall_force(self.h, _dp(o)))
        return o

    def diagnostics(self) -> dict:
        d = Diag()
        _check(self.lib, self.lib.lbm_hip_diagnostics(self.h, C.byref(d)))
        return dict(count_fluid=int(d.count_fluid), mean_density=d.mean_density, max_speed=d.max_speed,
                    sum_ux=d.sum_ux, unstable_nodes=int(d.unstable_nodes))

    def populations(self) -> np.ndarray:
        f = np.empty(Q * self.N)
        _check(self.lib, self.lib.lbm_hip_download(self.h, _dp(f), None, None, None, None, None))
        return f.reshape(Q, self.N)

    def fields(self, names=("rho", "ux", "uy", "uz", "flag")) -> dict:
        """Moments of the last step that wrote them (as doubles) and the flags;
        only the named ones are transferred."""
        N = self.N
        out = {k: (np.empty(N, np.uint8) if k == "flag" else np.empty(N)) for k in names}
        g = lambda k: _dp(out[k]) if k in out else None
        fl = out["flag"].ctypes.data_as(C.POINTER(C.c_uint8)) if "flag" in out else None
        _check(self.lib, self.lib.lbm_hip_download(self.h, None, g("rho"), g("ux"), g("uy"), g("uz"), fl))
        return out

    def flags(self) -> np.ndarray:
        flag = np.empty(self.N, np.uint8)
        _check(self.lib, self.lib.lbm_hip_download(self.h, None, None, None, None, None,
                                                   flag.ctypes.data_as(C.POINTER(C.c_uint8))))
        return flag

    def particle_links(self) -> tuple[np.ndarray, np.ndarray]:
        """The moving-boundary links (node, direction) of the current flags, in
        the order of the device's link list."""
        cnt = C.c_int64(0)
        _check(self.lib, self.lib.lbm_hip_particle_links(self.h, None, None, 0, C.byref(cnt)))
        node = np.empty(cnt.value, np.uint32)
        d = np.empty(cnt.value, np.uint8)
        _check(self.lib, self.lib.lbm_hip_particle_links(self.h, node.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                         d.ctypes.data_as(C.POINTER(C.c_uint8)), cnt.value,
                                                         C.byref(cnt)))
        return node, d

    def upload_populations(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if f.size != Q * self.N:
            raise ValueError("populations must have 19*N entries")
        _check(self.lib, self.lib.lbm_hip_upload_populations(self.h, _dp(f)))

    _SQ_ARRAYS = ("X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw")

    def save_checkpoint(self, path) -> None:
        """Dump of the state for a restart: the populations in the reference
        layout (a pending mass-correction shift included), the flags, the
        squirmer state, the steps taken and the creation parameters, as an
        .npz of plain arrays (nothing pickled).  Lattice.from_checkpoint
        restarts from it.  The reference has no checkpoint: its VTK holds no
        populations (io.cpp:50-65)."""
        if not self.owned:
            raise ValueError("a batch member cannot be checkpointed on its own")
        arrays = dict(f=self.populations(), flag=self.flags(), steps=np.int64(self.steps_before + self.info().steps_done),
                      params=np.array(json.dumps(self.kv)))
        if self.has_squirmer:
            sq = self.squirmer()
            for k in self._SQ_ARRAYS:
                arrays["sq_" + k] = sq[k]
            arrays["sq_contacts"] = np.int64(sq["contacts"])
        np.savez(path, **arrays)

    @classmethod
    def from_checkpoint(cls, path, **overrides) -> "Lattice":
        """A lattice restarted from save_checkpoint's dump.  It is created from
        the saved parameters (`overrides` may change e.g. the device) with the
        squirmer at its saved centre, so that its initial flags are the
        classification the saved flags hold (checked); then the squirmer state
        and the populations are uploaded.  `steps_before` holds the saved run's
        steps.  The link list is rebuilt from the flags, so the first force
        sums after a restart add the same links in another order: an FP64 run
        continues within round-off of the uninterrupted one."""
        with np.load(path, allow_pickle=False) as d:
            data = {k: d[k] for k in d.files}
        kv = json.loads(str(data["params"]))
        kv.update(overrides)
        if "sq_X" in data:
            X = data["sq_X"]
            kv.update(x0=float(X[0]), y0=float(X[1]), z0=float(X[2]))
        L = cls(**kv)
        try:
            if L.N != data["flag"].size or not np.array_equal(L.flags(), data["flag"]):
                raise LbmHipError(-4, "checkpoint flags differ from the classification at the saved centre")
            if "sq_X" in data:
                L.set_squirmer(**{k: data["sq_" + k] for k in cls._SQ_ARRAYS}, contacts=int(data["sq_contacts"]))
            L.upload_populations(data["f"])
        except BaseException:
            L.close()
            raise
        L.steps_before = int(data["steps"])
        return L

    def time_run(self, steps: int, mass_every: int = 10) -> tuple[float, float]:
        tot, k = C.c_double(), C.c_double()
        _check(self.lib, self.lib.lbm_hip_time_run(self.h, steps, mass_every, C.byref(tot), C.byref(k)))
        return tot.value, k.value

    def autotune(self, trial_steps: int = 20, mass_every: int = 10) -> dict:
        """Pick the collide-stream tile shape by timing the candidates from
        the current state; the state is left as it was."""
        t = Tuning()
        _check(self.lib, self.lib.lbm_hip_autotune(self.h, trial_steps, mass_every, C.byref(t)))
        return t.as_dict()

    def set_tiling(self, tile_width: int, xcd_remap: int = 1):
        _check(self.lib, self.lib.lbm_hip_set_tiling(self.h, tile_width, xcd_remap))

    def set_occupancy(self, even_blocks: int, odd_blocks: int):
        """Cap on collide-stream blocks per CU for even / odd steps (0: none)."""
        _check(self.lib, self.lib.lbm_hip_set_occupancy(self.h, even_blocks, odd_blocks))


class Batch:
    """An ensemble of lattices advanced together (one launch per phase for
    all members, graph-replayed time step)."""

    def __init__(self, members: list[dict], precision: str = "fp16"):
        self.lib = load_library()
        arr = (Params * len(members))()
        for k, kv in enumerate(members):
            kv = dict(kv)
            kv.setdefault("precision", precision)
            arr[k] = make_params(**kv)
        h = C.c_void_p()
        _check(self.lib, self.lib.lbm_hip_batch_create(arr, len(members), C.byref(h)))
        self.h = h
        self.count = len(members)
        self._members = [None] * self.count

    def close(self):
        if getattr(self, "h", None):
            for m in self._members:
                if m is not None:
                    m.h = None
            self.lib.lbm_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, steps: int, mass_every: int = 10, last_flags: int = 0):
        _check(self.lib, self.lib.lbm_hip_batch_run(self.h, steps, mass_every, last_flags))

    def synchronize(self):
        _check(self.lib, self.lib.lbm_hip_batch_synchronize(self.h))

    def member(self, i: int) -> Lattice:
        """Member i as a Lattice, after the batch's queued work (one stream
        synchronisation; the member's own getters report its errors)."""
        hm = C.c_void_p()
        _check(self.lib, self.lib.lbm_hip_batch_member(self.h, i, C.byref(hm)))
        if self._members[i] is None:
            self._members[i] = Lattice(_handle=hm)
        return self._members[i]

    def info(self) -> dict:
        o = BatchInfo()
        _check(self.lib, self.lib.lbm_hip_batch_info_get(self.h, C.byref(o)))
        return {k: getattr(o, k) for k, _ in BatchInfo._fields_}

    def autotune(self, trial_steps: int = 20, mass_every: int = 10) -> dict:
        t = Tuning()
        _check(self.lib, self.lib.lbm_hip_batch_autotune(self.h, trial_steps, mass_every, C.byref(t)))
        return t.as_dict()

    def set_tiling(self, tile_width: int, xcd_remap: int = 1):
        _check(self.lib, self.lib.lbm_hip_batch_set_tiling(self.h, tile_width, xcd_remap))

    def set_occupancy(self, even_blocks: int, odd_blocks: int):
        _check(self.lib, self.lib.lbm_hip_batch_set_occupancy(self.h, even_blocks, odd_blocks))

    def time_run(self, steps: int, mass_every: int = 10) -> float:
        tot = C.c_double()
        _check(self.lib, self.lib.lbm_hip_batch_time_run(self.h, steps, mass_every, C.byref(tot)))
        return tot.value


def ensemble_members(count: int = 64, seed: int = 0) -> list[dict]:
    """The batched-ensemble workload: `count` slit-channel squirmer runs drawn
    with numpy's default_rng(seed), in order per run: channel height H
    (integer in [32, 96]), radius R (integer in [4, 8]), beta (uniform in
    [-5, 5]), pitch toward the wall (uniform in [0, 45) deg), centre height
    (uniform in [2R, H - 2R]).  Grid 4H x 2H x H, walls in z, x and y
    periodic, B1 = 1e-3, tau = 1, hard gap 1, pitch in the x-z plane."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        H = int(rng.integers(32, 97))
        R = int(rng.integers(4, 9))
        beta = float(rng.uniform(-5.0, 5.0))
        pitch = float(rng.uniform(0.0, 45.0))
        z0 = float(rng.uniform(2 * R, H - 2 * R))
        out.append(dict(nx=4 * H, ny=2 * H, nz=H, tau=1.0, walls_y=0, walls_z=1, squirmer=1, radius=float(R),
                        b1=1e-3, beta=beta, angle=pitch, z0=z0, planar=1, plane_xz=1, eps=0.2, hc=2.0, hmin=1.0))
    return out


# ---------------------------------------------------------------------------
# Learned velocity surrogate (include/lbm_mlp.h)
# ---------------------------------------------------------------------------
MLP_IN, MLP_OUT, MLP_PARAMS = 5, 6, 34566
MLP_MODES = {"fp32": 0, "bf16": 1}


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


class MLP:
    """An ensemble of 5 -> 128 -> 128 -> 128 -> 6 tanh nets trained and
    evaluated by the library's HIP kernels."""

    def __init__(self, nets: int = 4, mode: str = "bf16", seed: int = 0, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib, self.lib.lbm_mlp_create(nets, MLP_MODES[mode], seed, device, C.byref(h)))
        self.h, self.nets, self.mode = h, nets, mode

    def close(self):
        if getattr(self, "h", None):
            self.lib.lbm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _f32(a, cols):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.ndim != 2 or a.shape[1] != cols:
            raise ValueError(f"expected an (n, {cols}) array")
        return a

    def params(self, k: int) -> np.ndarray:
        p = np.empty(MLP_PARAMS, np.float32)
        _check(self.lib, self.lib.lbm_mlp_get_params(self.h, k, _fp(p)))
        return p

    def set_params(self, k: int, p):
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        if p.size != MLP_PARAMS:
            raise ValueError("wrong parameter count")
        _check(self.lib, self.lib.lbm_mlp_set_params(self.h, k, _fp(p)))

    def forward(self, x) -> np.ndarray:
        x = self._f32(x, MLP_IN)
        y = np.empty((self.nets, x.shape[0], MLP_OUT), np.float32)
        _check(self.lib, self.lib.lbm_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)))
        return y

    def predict(self, x) -> tuple[np.ndarray, np.ndarray]:
        x = self._f32(x, MLP_IN)
        mean = np.empty((x.shape[0], MLP_OUT), np.float32)
        spread = np.empty(x.shape[0], np.float32)
        _check(self.lib, self.lib.lbm_mlp_predict(self.h, _fp(x), x.shape[0], _fp(mean), _fp(spread)))
        return mean, spread

    def loss_grad(self, k: int, x, t) -> tuple[float, np.ndarray]:
        x, t = self._f32(x, MLP_IN), self._f32(t, MLP_OUT)
        loss = C.c_float()
        g = np.empty(MLP_PARAMS, np.float32)
        _check(self.lib, self.lib.lbm_mlp_loss_grad(self.h, k, _fp(x), _fp(t), x.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, x, t, epochs: int, batch: int = 256, lr: float = 1e-3) -> np.ndarray:
        x, t = self._f32(x, MLP_IN), self._f32(t, MLP_OUT)
        losses = np.empty(self.nets, np.float32)
        _check(self.lib, self.lib.lbm_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], epochs, batch, lr,
                                                _fp(losses)))
        return losses
