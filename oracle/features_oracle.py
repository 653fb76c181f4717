"""FP64 numpy restatement of the learned SGS surrogate's input features --
TEST INFRASTRUCTURE ONLY.

Only tests/ import this module.  The features have no counterpart in the
reference (the surrogate is this build's addition); what they are built on
does: the strain rate S_ij on the nodes of the reference's compute_sij
(src/compute_sgs.f90:260-335: the first level on the uvp node with the
wall-law du/dz, dv/dz and w gradients averaged up; other levels on w nodes
with the horizontal gradients averaged over k-1, k), the rotation rate R_ij
with the same node placement, and the spectral cut-off test filters of
src/testfilter.f90:101-184 (G = 1/(nx ny) where kk2 < (pi / (alpha
sqrt(dx dy)))^2, Nyquist row and column zeroed; alpha = 2 and 4).

For every point of SGS levels 1..nz+1 the 12 features are
  f0 = log(|S| + 1e-12) at delta, |S| = sqrt(2 S:S),
  f1 = R:R / (S:S + 1e-20), f2 = tr(S^3) / (|S|^3 + 1e-30),
  f3 = tr(S R^2) / (|S| S:S + 1e-30)                      (at delta),
  f4 = log((|S|_2d + 1e-12) / (|S| + 1e-12)), f5..f7 = f1..f3 at 2 delta,
  f8 = log((|S|_4d + 1e-12) / (|S| + 1e-12)), f9, f10 = f1, f2 at 4 delta,
  f11 = log(z / delta) with z = dz/2 on level 1 and (k-1) dz above,
where the quantities at 2 delta and 4 delta are those of the test-filtered
S_ij and R_ij (each level filtered in its plane).
"""
from __future__ import annotations

import numpy as np

GRADS = ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")


def strain_rotation(g: dict, nz: int):
    """S (6: 11 12 13 22 23 33) and R (3: 12 13 23) on SGS levels 1..nz+1
    from the nine velocity gradients g[name] of shape (nz+2, ny, nx), on the
    nodes of compute_sij.  Returns arrays of shape (6 or 3, nz+1, ny, nx)."""
    h = 0.5
    S = np.zeros((6, nz + 1) + g["dudx"].shape[1:])
    R = np.zeros((3, nz + 1) + g["dudx"].shape[1:])
    for k in range(1, nz + 2):
        if k == 1:
            ux, uy, uz = g["dudx"][1], g["dudy"][1], g["dudz"][1]
            vx, vy, vz = g["dvdx"][1], g["dvdy"][1], g["dvdz"][1]
            wx = h * (g["dwdx"][1] + g["dwdx"][2])
            wy = h * (g["dwdy"][1] + g["dwdy"][2])
            wz = g["dwdz"][1]
        else:
            ux = h * (g["dudx"][k] + g["dudx"][k - 1])
            uy = h * (g["dudy"][k] + g["dudy"][k - 1])
            uz = g["dudz"][k]
            vx = h * (g["dvdx"][k] + g["dvdx"][k - 1])
            vy = h * (g["dvdy"][k] + g["dvdy"][k - 1])
            vz = g["dvdz"][k]
            wx, wy = g["dwdx"][k], g["dwdy"][k]
            wz = h * (g["dwdz"][k] + g["dwdz"][k - 1])
        i = k - 1
        S[:, i] = (ux, h * (uy + vx), h * (uz + wx), vy, h * (vz + wy), wz)
        R[:, i] = (h * (uy - vx), h * (uz - wx), h * (vz - wy))
    return S, R


def test_filter(f: np.ndarray, G: np.ndarray) -> np.ndarray:
    """Spectral cut-off of every plane of f (..., ny, nx) with the table G
    (ny, nx/2+1) that carries the 1/(nx ny) of the unnormalised transforms."""
    nx, ny = f.shape[-1], f.shape[-2]
    return np.fft.irfft2(np.fft.rfft2(f) * (G * (nx * ny)), s=(ny, nx))


def invariants(S: np.ndarray, R: np.ndarray):
    """|S|, R:R/S:S, tr(S^3)/|S|^3 and tr(S R^2)/(|S| S:S) pointwise."""
    s11, s12, s13, s22, s23, s33 = S
    r12, r13, r23 = R
    Sm = np.array([[s11, s12, s13], [s12, s22, s23], [s13, s23, s33]])
    z = np.zeros_like(r12)
    Rm = np.array([[z, r12, r13], [-r12, z, r23], [-r13, -r23, z]])
    SS = np.einsum("ij...,ij...->...", Sm, Sm)
    RR = np.einsum("ij...,ij...->...", Rm, Rm)
    S2 = np.einsum("ik...,kj...->ij...", Sm, Sm)
    R2 = np.einsum("ik...,kj...->ij...", Rm, Rm)
    S3 = np.einsum("ij...,ji...->...", S2, Sm)
    SR2 = np.einsum("ij...,ji...->...", Sm, R2)
    mag = np.sqrt(2.0 * SS)
    return mag, RR / (SS + 1e-20), S3 / (mag ** 3 + 1e-30), SR2 / (mag * SS + 1e-30)


def features(g: dict, G2: np.ndarray, G4: np.ndarray, nz: int, dz: float, delta: float) -> np.ndarray:
    """The 12 features of every point of levels 1..nz+1: (nz+1, ny, nx, 12)."""
    S, R = strain_rotation(g, nz)
    inv0 = invariants(S, R)
    inv2 = invariants(test_filter(S, G2), test_filter(R, G2))
    inv4 = invariants(test_filter(S, G4), test_filter(R, G4))
    k = np.arange(1, nz + 2, dtype=float)
    z = np.where(k == 1, 0.5 * dz, (k - 1) * dz)[:, None, None] * np.ones_like(inv0[0])
    m0 = inv0[0] + 1e-12
    f = [np.log(m0), inv0[1], inv0[2], inv0[3],
         np.log((inv2[0] + 1e-12) / m0), inv2[1], inv2[2], inv2[3],
         np.log((inv4[0] + 1e-12) / m0), inv4[1], inv4[2],
         np.log(z / delta)]
    return np.stack(f, axis=-1)
