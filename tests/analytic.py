"""Exact solutions and validation metrics used by the tests.

Restated from the mathematics the reference uses:
* rectangular-duct Poiseuille series (reference src/utils/analytic.cpp:11-21,
  README.md:421-424; F. M. White, "Viscous Fluid Flow"), 200 odd terms;
* plane Poiseuille profile (analytic.cpp:23-25);
* relative L2 error of u_x on the cross-section x = 0 (analytic.cpp:27-46);
* Blake's squirmer flow field (postprocessing/lbm_post.py:153-170, Blake 1971)
  and the near-field error metric of lbm_post.py:350-361;
* lateral extrema of a slit-channel trajectory (lbm_post.py:224-238).
"""
from __future__ import annotations

import numpy as np


def duct_velocity(y, z, a, b, G, mu, nterms=200):
    y, z = np.asarray(y, float), np.asarray(z, float)
    s = np.zeros(np.broadcast(y, z).shape)
    for n in range(1, 2 * nterms, 2):
        k = n * np.pi / (2.0 * a)
        A, B = k * np.abs(z), k * b
        ratio = np.exp(A - B) * (1.0 + np.exp(-2.0 * A)) / (1.0 + np.exp(-2.0 * B))
        sign = 1.0 if ((n - 1) // 2) % 2 == 0 else -1.0
        s = s + sign * (1.0 - ratio) * np.cos(k * y) / float(n) ** 3
    return 16.0 * a * a * G / (mu * np.pi ** 3) * s


def plane_velocity(y, a, G, mu):
    return G / (2.0 * mu) * (a * a - np.asarray(y, float) ** 2)


def poiseuille_error(ux, flag, NX, NY, NZ, walls_y, walls_z, nu, G):
    """Relative L2 error of u_x at x = 0 (analytic.cpp:27-46); ux/flag flat arrays."""
    if not walls_y and not walls_z:
        return -1.0
    yc, zc = 0.5 * (NY - 1), 0.5 * (NZ - 1)
    a, b = 0.5 * (NY - 2), 0.5 * (NZ - 2)
    y, z = np.meshgrid(np.arange(NY), np.arange(NZ), indexing="xy")  # [z, y]
    if walls_y and walls_z:
        ue = duct_velocity(y - yc, z - zc, a, b, G, nu)
    elif walls_y:
        ue = plane_velocity(y - yc, a, G, nu)
    else:
        ue = plane_velocity(z - zc, b, G, nu)
    u = ux.reshape(NZ, NY, NX)[:, :, 0]
    fl = flag.reshape(NZ, NY, NX)[:, :, 0] == 0
    num = np.sum((u[fl] - ue[fl]) ** 2)
    den = np.sum(ue[fl] ** 2)
    return float(np.sqrt(num / den)) if den > 0 else -1.0


def blake_velocity(x, y, z, R, B1, beta, e=(1.0, 0.0, 0.0)):
    """Lab-frame Stokes flow around a squirmer at the origin; NaN inside."""
    x, y, z = np.broadcast_arrays(*(np.asarray(v, float) for v in (x, y, z)))
    e = np.asarray(e, float)
    r = np.maximum(np.sqrt(x ** 2 + y ** 2 + z ** 2), 1e-12)
    rh = np.stack([x / r, y / r, z / r])
    ct = np.tensordot(e, rh, axes=1)
    E = e.reshape((3,) + (1,) * x.ndim)
    B2 = beta * B1
    u = (B1 * (R / r) ** 3 * (ct * rh - E / 3)
         + B2 * ((R / r) ** 4 - (R / r) ** 2) * (1.5 * ct ** 2 - 0.5) * rh
         + B2 * (R / r) ** 4 * ct * (ct * rh - E))
    u[:, r < R] = np.nan
    return u


def near_field_error(ux, uy, flag, NX, NY, NZ, X, R, B1, beta):
    """Relative L2 error of (ux, uy) against Blake on the x-y plane nearest the
    centre, for 1.25 R < r < 3 R (lbm_post.py:200-221, 350-361), periodic box."""
    Xx, Xy, Xz = X
    zp = int(np.floor(Xz % NZ))
    dz = zp - (Xz % NZ)
    dx = np.arange(NX) - (Xx % NX)
    dx -= NX * np.round(dx / NX)
    dy = np.arange(NY) - Xy
    dy -= NY * np.round(dy / NY)
    ox, oy = np.argsort(dx), np.argsort(dy)
    U = ux.reshape(NZ, NY, NX)[zp][np.ix_(oy, ox)]
    V = uy.reshape(NZ, NY, NX)[zp][np.ix_(oy, ox)]
    F = flag.reshape(NZ, NY, NX)[zp][np.ix_(oy, ox)]
    px, py = dx[ox], dy[oy]
    U = np.where(F == 0, U, np.nan)
    V = np.where(F == 0, V, np.nan)
    PX, PY = np.meshgrid(px, py)
    ub = blake_velocity(PX, PY, np.full_like(PX, dz), R, B1, beta)
    rr = np.sqrt(PX ** 2 + PY ** 2 + dz ** 2)
    mk = (rr > 1.25 * R) & (rr < 3 * R) & ~np.isnan(U)
    return float(np.sqrt(np.sum((U[mk] - ub[0][mk]) ** 2 + (V[mk] - ub[1][mk]) ** 2)
                         / np.sum(ub[0][mk] ** 2 + ub[1][mk] ** 2)))


def read_vtk(path):
    """Legacy binary VTK written by the reference (io.cpp:44-67): big-endian."""
    data = open(path, "rb").read()
    out = {}
    pos = 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    while pos < len(data):
        s = line()
        if s.startswith("DIMENSIONS"):
            out["dims"] = tuple(int(v) for v in s.split()[1:])
        elif s.startswith("POINT_DATA"):
            n = int(s.split()[1])
        elif s.startswith("SCALARS flag"):
            line()
            out["flag"] = np.frombuffer(data, np.uint8, n, pos).copy()
            pos += n + 1
        elif s.startswith("SCALARS density"):
            line()
            out["density"] = np.frombuffer(data, ">f8", n, pos).astype(float)
            pos += 8 * n + 1
        elif s.startswith("VECTORS velocity"):
            out["velocity"] = np.frombuffer(data, ">f8", 3 * n, pos).astype(float).reshape(n, 3)
            pos += 24 * n + 1
    return out


def oscillation_extrema(t, Y, ny, R, U0, t_skip=0.0, prominence=0.02):
    """Lateral extrema of a swimmer in a slit channel with walls in y
    (postprocessing/lbm_post.py:224-238).  The channel width is H = ny - 2
    (wall nodes at y = 0 and ny - 1), its centre line y = (ny - 1) / 2; time
    is scaled as T = t U0 / H and the lateral position as
    yn = (Y - centre) / (H / 2 - R), so |yn| = 1 is a sphere touching a wall.
    Returns (T, yn, idx), idx the sorted indices of the local maxima and
    minima of yn with at least the given prominence (scipy find_peaks) and
    T > t_skip."""
    from scipy.signal import find_peaks
    t, Y = np.asarray(t, float), np.asarray(Y, float)
    H = ny - 2
    T = t * U0 / H
    yn = (Y - 0.5 * (ny - 1)) / (0.5 * H - R)
    hi, _ = find_peaks(yn, prominence=prominence)
    lo, _ = find_peaks(-yn, prominence=prominence)
    idx = np.sort(np.concatenate([hi, lo]))
    return T, yn, idx[T[idx] > t_skip]
