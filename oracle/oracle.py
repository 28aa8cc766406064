"""ctypes bindings for the CPU checkers.  TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the bench's CPU-baseline leg import this
module; the product library never does.

* ``Oracle``     -- the C restatement of the reference path (liboracle.so,
                    lbm_oracle.c), FP64, push scheme, two arrays.
* ``Reference``  -- the reference's own sources compiled here into
                    ``_ref/libref_lbm.so`` (absent where the reference tree
                    was not available at build time).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
Q = 19


class OrcParams(C.Structure):
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
    ]


class OrcSquirmer(C.Structure):
    _fields_ = [
        ("R", C.c_double), ("B1", C.c_double), ("beta", C.c_double), ("M", C.c_double), ("I", C.c_double),
        ("planar", C.c_int), ("plane_xz", C.c_int),
        ("X", C.c_double * 3), ("e", C.c_double * 3), ("U", C.c_double * 3), ("W", C.c_double * 3),
        ("Fh", C.c_double * 3), ("Th", C.c_double * 3), ("Fc", C.c_double * 3), ("Tc", C.c_double * 3),
        ("Fw", C.c_double * 3),
        ("eps", C.c_double), ("hc", C.c_double), ("hmin", C.c_double), ("FS", C.c_double),
        ("contacts", C.c_long),
    ]


class OrcLBM(C.Structure):
    _fields_ = [
        ("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int), ("N", C.c_size_t),
        ("tau", C.c_double), ("omPlus", C.c_double), ("omMinus", C.c_double), ("nu", C.c_double),
        ("F", C.c_double * 3),
        ("wallsY", C.c_int), ("wallsZ", C.c_int),
        ("f", C.POINTER(C.c_double)), ("fnew", C.POINTER(C.c_double)),
        ("rho", C.POINTER(C.c_double)), ("ux", C.POINTER(C.c_double)),
        ("uy", C.POINTER(C.c_double)), ("uz", C.POINTER(C.c_double)),
        ("flag", C.POINTER(C.c_uint8)),
        ("wallForce", C.c_double * 3),
        ("hasSquirmer", C.c_int),
        ("sq", OrcSquirmer),
    ]


# Reference key names (params.cpp:40-70) -> OrcParams fields.
_KEYS = {
    "nx": "nx", "ny": "ny", "nz": "nz", "tau": "tau", "lambda": "lambda_",
    "fx": "fx", "fy": "fy", "fz": "fz", "walls_y": "walls_y", "walls_z": "walls_z",
    "squirmer": "squirmer", "radius": "radius", "b1": "b1", "beta": "beta", "rho_p": "rho_p",
    "x0": "x0", "y0": "y0", "z0": "z0", "angle": "angle", "planar": "planar",
    "eps": "eps", "hc": "hc", "hmin": "hmin", "plane_xz": "plane_xz",
}

_lib = None


def _load():
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} not built (run make -C oracle)")
        lib = C.CDLL(str(path))
        lib.orc_create.restype = C.POINTER(OrcLBM)
        lib.orc_create.argtypes = [C.POINTER(OrcParams)]
        lib.orc_default_params.argtypes = [C.POINTER(OrcParams)]
        for fn in ("orc_destroy", "orc_step", "orc_advance_squirmer", "orc_correct_mass"):
            getattr(lib, fn).argtypes = [C.POINTER(OrcLBM)]
        lib.orc_run.argtypes = [C.POINTER(OrcLBM), C.c_int, C.c_int]
        lib.orc_count_fluid.restype = C.c_long
        for fn in ("orc_mean_density", "orc_max_speed", "orc_sum_ux"):
            getattr(lib, fn).restype = C.c_double
        for fn in ("orc_count_fluid", "orc_mean_density", "orc_max_speed", "orc_sum_ux"):
            getattr(lib, fn).argtypes = [C.POINTER(OrcLBM)]
        lib.orc_equilibrium.restype = C.c_double
        lib.orc_equilibrium.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        lib.orc_soft_wall_force.restype = C.c_double
        lib.orc_soft_wall_force.argtypes = [C.c_double] * 4
        lib.orc_surface_velocity.argtypes = [C.POINTER(OrcSquirmer), C.POINTER(C.c_double),
                                             C.POINTER(C.c_double), C.c_int]
        lib.orc_integrate.argtypes = [C.POINTER(OrcSquirmer)]
        lib.orc_offset_from_squirmer.argtypes = [C.POINTER(OrcLBM), C.c_double, C.c_double, C.c_double,
                                                 C.POINTER(C.c_double)]
        lib.orc_w.restype = C.c_double
        lib.orc_num_threads.restype = C.c_int
        _lib = lib
    return _lib


def params(**kv) -> OrcParams:
    lib = _load()
    p = OrcParams()
    lib.orc_default_params(C.byref(p))
    for k, v in kv.items():
        setattr(p, _KEYS[k.lower()], v)
    return p


def set_mass_sum_exact(exact: bool) -> None:
    """Mass-correction summation of the oracle: the reference's double sum
    (default) or the exact fixed-point sum (orc_set_mass_sum_exact)."""
    _load().orc_set_mass_sum_exact(C.c_int(1 if exact else 0))


def set_threads(n: int) -> None:
    """OpenMP threads for the oracle (and the reference build)."""
    os.environ["OMP_NUM_THREADS"] = str(n)
    try:
        C.CDLL(None).omp_set_num_threads  # noqa: B018
    except AttributeError:
        pass
    _load()
    libgomp = C.CDLL("libgomp.so.1")
    libgomp.omp_set_num_threads(C.c_int(n))


class Oracle:
    """The C restatement of class LBM (reference lbm.hpp:19-84)."""

    def __init__(self, **kv):
        self.lib = _load()
        self.p = params(**kv)
        self.h = self.lib.orc_create(C.byref(self.p))
        if not self.h:
            raise MemoryError("orc_create failed")
        L = self.h.contents
        self.NX, self.NY, self.NZ, self.N = L.NX, L.NY, L.NZ, L.N

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    # ---- entry points ----
    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.orc_step(self.h)

    def advance_squirmer(self):
        self.lib.orc_advance_squirmer(self.h)

    def correct_mass(self):
        self.lib.orc_correct_mass(self.h)

    def run(self, steps: int, mass_every: int = 10):
        self.lib.orc_run(self.h, steps, mass_every)

    # ---- data ----
    @property
    def L(self) -> OrcLBM:
        return self.h.contents

    def f(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.L.f, shape=(Q * self.N,)).reshape(Q, self.N).copy()

    def field(self, name: str) -> np.ndarray:
        ptr = getattr(self.L, name)
        return np.ctypeslib.as_array(ptr, shape=(self.N,)).copy()

    def flags(self) -> np.ndarray:
        return np.ctypeslib.as_array(self.L.flag, shape=(self.N,)).copy()

    def state(self) -> dict:
        s = self.L.sq
        out = {k: np.array(getattr(s, k)[:]) for k in ("X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw")}
        out["contacts"] = s.contacts
        out["wallForce"] = np.array(self.L.wallForce[:])
        for k in ("R", "B1", "beta", "M", "I", "FS"):
            out[k] = getattr(s, k)
        return out

    def diagnostics(self) -> dict:
        return dict(count_fluid=self.lib.orc_count_fluid(self.h), mean_density=self.lib.orc_mean_density(self.h),
                    max_speed=self.lib.orc_max_speed(self.h), sum_ux=self.lib.orc_sum_ux(self.h))


def reference_available() -> bool:
    return (HERE / "_ref" / "libref_lbm.so").exists()


_ref_libs: dict = {}


def _load_ref(variant: str = ""):
    name = f"libref_lbm{variant}.so"
    if name not in _ref_libs:
        path = HERE / "_ref" / name
        if not path.exists():
            raise RuntimeError(f"{path} not built (needs the reference tree at build time)")
        lib = C.CDLL(str(path))
        lib.ref_create.restype = C.c_void_p
        lib.ref_create.argtypes = [C.c_char_p]
        lib.ref_last_error.restype = C.c_char_p
        for fn in ("ref_destroy", "ref_step", "ref_advance", "ref_correct_mass"):
            getattr(lib, fn).argtypes = [C.c_void_p]
        lib.ref_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.ref_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        dp = C.POINTER(C.c_double)
        lib.ref_get_fields.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, C.POINTER(C.c_uint8)]
        lib.ref_get_state.argtypes = [C.c_void_p, dp]
        lib.ref_diagnostics.argtypes = [C.c_void_p, dp]
        lib.ref_set_threads.argtypes = [C.c_int]
        _ref_libs[name] = lib
    return _ref_libs[name]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class Reference:
    """The reference class LBM compiled from its own sources (lbm.cpp etc.)."""

    def __init__(self, variant: str = "", **kv):
        self.lib = _load_ref(variant)
        spec = " ".join(f"{k}={v}" for k, v in kv.items())
        self.h = self.lib.ref_create(spec.encode())
        if not self.h:
            raise ValueError(self.lib.ref_last_error().decode())
        d = (C.c_int * 4)()
        self.lib.ref_dims(self.h, d)
        self.NX, self.NY, self.NZ, self.has_squirmer = d[0], d[1], d[2], d[3]
        self.N = self.NX * self.NY * self.NZ

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_destroy(self.h)
            self.h = None

    def step(self, n: int = 1):
        for _ in range(n):
            self.lib.ref_step(self.h)

    def advance_squirmer(self):
        self.lib.ref_advance(self.h)

    def correct_mass(self):
        self.lib.ref_correct_mass(self.h)

    def run(self, steps: int, mass_every: int = 10):
        self.lib.ref_run(self.h, steps, mass_every)

    def fields(self) -> dict:
        N = self.N
        f = np.empty(Q * N)
        rho, ux, uy, uz = (np.empty(N) for _ in range(4))
        flag = np.empty(N, np.uint8)
        self.lib.ref_get_fields(self.h, _dp(f), _dp(rho), _dp(ux), _dp(uy), _dp(uz),
                                flag.ctypes.data_as(C.POINTER(C.c_uint8)))
        return dict(f=f.reshape(Q, N), rho=rho, ux=ux, uy=uy, uz=uz, flag=flag)

    def state(self) -> dict:
        o = np.empty(38)
        self.lib.ref_get_state(self.h, _dp(o))
        names = ["X", "e", "U", "W", "Fh", "Th", "Fc", "Tc", "Fw", "wallForce"]
        out = {k: o[3 * i:3 * i + 3].copy() for i, k in enumerate(names)}
        out["contacts"] = int(o[30])
        out["omPlus"], out["omMinus"], out["nu"] = o[31], o[32], o[33]
        out["M"], out["I"], out["FS"], out["U0"] = o[34], o[35], o[36], o[37]
        return out

    def diagnostics(self) -> dict:
        o = np.empty(4)
        self.lib.ref_diagnostics(self.h, _dp(o))
        return dict(count_fluid=int(o[0]), mean_density=o[1], max_speed=o[2], sum_ux=o[3])

    def set_threads(self, n: int):
        self.lib.ref_set_threads(n)
