"""Seeded initial fields for ``ic_type='files'``.

The reference reads u, v, w (nx, ny, nz float64, Fortran order, no record
markers) from ``vel_u.dat``/``vel_v.dat``/``vel_w.dat``
(src/initial_conditions.f90:130-165).  These helpers generate the benchmark
initial condition -- a log-law mean u(z) = ln(z/z0)/kappa with seeded
uniform perturbations of +-a of the local mean for z <= 0.5, v = w = 0 --
from a counter-based hash (SplitMix64 of seed and linear index), so any
host regenerates identical fields without files.
"""
from __future__ import annotations

import pathlib
from typing import Tuple

import numpy as np

KAPPA = 0.4
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied to uint64 counters (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform_pm1(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n values in [-1, 1) from counters (seed << 40) + offset + [0, n)."""
    ctr = np.arange(offset, offset + n, dtype=np.uint64) + (np.uint64(seed) << np.uint64(40))
    bits = splitmix64(ctr) >> np.uint64(11)
    return bits.astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0


def log_law_ic(nx: int, ny: int, nz: int, lz: float, z0: float, amp: float, seed: int = 1234,
               z_turb: float = 0.5) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Return u, v, w as (nz, ny, nx) arrays (C order = Fortran (nx, ny, nz))."""
    dz = lz / nz
    k = np.arange(1, nz + 1, dtype=np.float64)
    z = ((1.0 + k) - 1.5) * dz
    mean = (1.0 / KAPPA) * np.log(z / z0)
    xi = uniform_pm1(seed, nx * ny * nz).reshape(nz, ny, nx)
    pert = np.where((z <= z_turb)[:, None, None], amp * xi, 0.0)
    u = mean[:, None, None] * (1.0 + pert)
    v = np.zeros_like(u)
    w = np.zeros_like(u)
    return u, v, w


def write_vel_files(directory, u, v, w) -> None:
    """Write vel_u.dat, vel_v.dat, vel_w.dat the way the reference reads them."""
    d = pathlib.Path(directory)
    for name, a in (("vel_u.dat", u), ("vel_v.dat", v), ("vel_w.dat", w)):
        np.ascontiguousarray(a, dtype="<f8").tofile(d / name)


def read_vel_files(directory, nx: int, ny: int, nz: int):
    d = pathlib.Path(directory)
    return tuple(np.fromfile(d / n, dtype="<f8").reshape(nz, ny, nx) for n in ("vel_u.dat", "vel_v.dat", "vel_w.dat"))
