"""Pinning the numpy restatement of the surrogate's features
(oracle/features_oracle.py) by closed forms, and showing that each check
catches a plausible mistake planted in a copy of the restatement.

- invariants: an axisymmetric strain S = a diag(2, -1, -1) has |S| = a sqrt12,
  tr S^3 / |S|^3 = 6 / 12^1.5 and no rotation; a simple shear u = g y has
  R:R / S:S = 1 and tr S^3 = tr(S R^2) = 0;
- node placement of S_ij, R_ij (src/compute_sgs.f90:260-335): gradients
  linear in the level index give the averages of the w nodes exactly, level 1
  on the uvp node;
- test filters (src/testfilter.f90:101-184): a shear mode below both cut-offs
  (|k| < nx/4 at 2 delta, nx/8 at 4 delta) keeps |S| at both scales
  (f4 = f8 = 0); one between the cut-offs vanishes at 4 delta only;
- log(z / delta) at z = dz/2 and (k-1) dz.
"""
import importlib.util

import numpy as np
import pytest

from conftest import ORACLE_DIR, load_oracle, load_params

spec = importlib.util.spec_from_file_location("features_oracle", ORACLE_DIR / "features_oracle.py")
FO = importlib.util.module_from_spec(spec)
spec.loader.exec_module(FO)


def _grads(nz, ny, nx, **fields):
    g = {n: np.zeros((nz + 2, ny, nx)) for n in FO.GRADS}
    for n, v in fields.items():
        g[n] = np.broadcast_to(v, (nz + 2, ny, nx)).astype(float).copy()
    return g


def _tables(nx=32, ny=32, nz=8):
    P, O = load_params(), load_oracle()
    prm = P.LesParams(nx=nx, ny=ny, nz=nz)
    o = O.Oracle(prm)
    dz = prm.lz / nz
    delta = (prm.lx / nx * prm.ly / ny * dz) ** (1 / 3)
    return prm, o.table("G2"), o.table("G4"), dz, delta


# each check takes the module (or a copy with a planted mistake)
def check_axisymmetric_strain(fo):
    a = 0.7
    S = np.array([2 * a, 0, 0, -a, 0, -a])[:, None]
    R = np.zeros((3, 1))
    mag, rr, s3, sr2 = fo.invariants(S, R)
    assert abs(mag[0] - a * np.sqrt(12)) < 1e-12
    assert abs(s3[0] - 6 / 12 ** 1.5) < 1e-12 and rr[0] == 0 and sr2[0] == 0


def check_simple_shear(fo):
    gam = 1.3
    S = np.array([0, gam / 2, 0, 0, 0, 0])[:, None]
    R = np.array([gam / 2, 0, 0])[:, None]
    mag, rr, s3, sr2 = fo.invariants(S, R)
    assert abs(mag[0] - gam) < 1e-12 and abs(rr[0] - 1) < 1e-12 and abs(s3[0]) < 1e-15 and abs(sr2[0]) < 1e-15


def check_node_placement(fo):
    nz, ny, nx = 6, 4, 4
    k = np.arange(nz + 2, dtype=float)[:, None, None]
    g = _grads(nz, ny, nx, dudx=k, dudz=2 * k, dwdx=3 * k, dwdz=5 * k, dvdz=7 * k, dwdy=11 * k)
    S, R = fo.strain_rotation(g, nz)
    for kk in range(2, nz + 2):
        i = kk - 1
        assert np.allclose(S[0, i], kk - 0.5)            # dudx averaged over k-1, k
        assert np.allclose(S[2, i], 0.5 * (2 * kk + 3 * kk))  # dudz, dwdx on the w node
        assert np.allclose(S[5, i], 5 * (kk - 0.5))      # dwdz averaged
        assert np.allclose(R[1, i], 0.5 * (2 * kk - 3 * kk))
        assert np.allclose(S[4, i], 0.5 * (7 * kk + 11 * kk))
    # level 1: uvp node, w gradients averaged over levels 1 and 2
    assert np.allclose(S[0, 0], 1.0) and np.allclose(S[2, 0], 0.5 * (2 * 1 + 0.5 * (3 + 6)))
    assert np.allclose(S[5, 0], 5.0)


def check_filter_scales(fo):
    prm, G2, G4, dz, delta = _tables()
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    y = (np.arange(ny) * prm.ly / ny)[None, :, None]
    for m, keep4 in ((3, True), (6, False)):
        # u = cos(m y): du/dy = -m sin(m y), a shear of wavenumber m
        g = _grads(nz, ny, nx, dudy=-m * np.sin(m * y))
        f = fo.features(g, G2, G4, nz, dz, delta)
        mag = np.exp(f[..., 0]) - 1e-12
        big = mag > 1e-3
        assert np.allclose(f[..., 4][big], 0.0, atol=1e-9)  # kept at 2 delta
        assert np.allclose(f[..., 5][big], 1.0, atol=1e-9)  # R:R / S:S of a shear
        if keep4:
            assert np.allclose(f[..., 8][big], 0.0, atol=1e-9)
        else:
            assert np.all(f[..., 8][big] < np.log(1e-9))    # removed at 4 delta


def check_height(fo):
    prm, G2, G4, dz, delta = _tables(nz=8)
    g = _grads(prm.nz, prm.ny, prm.nx, dudz=1.0)
    f = fo.features(g, G2, G4, prm.nz, dz, delta)
    assert np.allclose(f[0, ..., 11], np.log(0.5 * dz / delta))
    for k in range(2, prm.nz + 2):
        assert np.allclose(f[k - 1, ..., 11], np.log((k - 1) * dz / delta))


CHECKS = {f.__name__[6:]: f for f in (check_axisymmetric_strain, check_simple_shear, check_node_placement,
                                      check_filter_scales, check_height)}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_features_oracle_pin(name):
    CHECKS[name](FO)


def _planted(which):
    """A copy of the restatement with one plausible mistake."""
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    if which == "mag_without_2":
        inv = m.invariants

        def invariants(S, R):
            mag, rr, s3, sr2 = inv(S, R)
            return mag / np.sqrt(2.0), rr, s3 * 2 ** 1.5, sr2 * np.sqrt(2.0)
        m.invariants = invariants
    elif which == "filters_swapped":
        feat = m.features
        m.features = lambda g, G2, G4, nz, dz, delta: feat(g, G4, G2, nz, dz, delta)
    elif which == "dudz_averaged":
        sr = m.strain_rotation

        def strain_rotation(g, nz):
            g2 = dict(g)
            d = g["dudz"].copy()
            d[1:] = 0.5 * (g["dudz"][1:] + g["dudz"][:-1])
            g2["dudz"] = d
            return sr(g2, nz)
        m.strain_rotation = strain_rotation
    elif which == "height_k_dz":
        feat = m.features

        def features(g, G2, G4, nz, dz, delta):
            f = feat(g, G2, G4, nz, dz, delta)
            f[..., 11] = np.log(np.arange(1, nz + 2, dtype=float) * dz / delta)[:, None, None]
            return f
        m.features = features
    elif which == "rotation_not_halved":
        sr = m.strain_rotation
        m.strain_rotation = lambda g, nz: (lambda S, R: (S, 2.0 * R))(*sr(g, nz))
    return m


PLANTED = {"mag_without_2": ["axisymmetric_strain"], "filters_swapped": ["filter_scales"],
           "dudz_averaged": ["node_placement"], "height_k_dz": ["height"],
           "rotation_not_halved": ["node_placement", "filter_scales"]}


@pytest.mark.parametrize("which", sorted(PLANTED))
def test_features_planted_mistake_is_caught(which):
    m = _planted(which)
    failed = []
    for name in PLANTED[which]:
        try:
            CHECKS[name](m)
        except AssertionError:
            failed.append(name)
    assert failed, f"planted mistake {which} passes {PLANTED[which]}"
