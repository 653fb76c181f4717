"""Pinning the oracle's running statistics, profile reduction and running
diagnostics (CPU only, no GPU).

The reference has no tests or stored outputs for these routines
(src/IO_statistics.f90, matlab_src/get_profile_struct.m,
src/running_diagnostics.f90), so they are pinned by closed forms on fields
chosen so that each plausible mistake changes the answer: a time variance
from two snapshots, a resolved stress whose level average must use level
k - 1 (not k + 1), a dispersive stress of two cosines, the level-1 resets
and the exclusion of the ghost levels.  A second, independent numpy
statement of the accumulation checks all 26 sums on random fields.
"""
import numpy as np
import pytest

from conftest import load_oracle, load_params

O = load_oracle()
P = load_params()

NX, NY, NZ = 12, 8, 5


def _oracle(dt=2e-4):
    prm = P.LesParams(nx=NX, ny=NY, nz=NZ, dt=dt)
    return O.Oracle(prm), prm


def _put_all(o, vals):
    """vals: name -> array (nz+2, ny, nx) or scalar; unnamed fields zero."""
    shape = (NZ + 2, NY, NX)
    for name in ["u", "v", "w", "p", "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nu_t",
                 "cs_opt2", "dudx", "dvdy", "dwdz"]:
        a = vals.get(name, 0.0)
        o.put(name, np.broadcast_to(np.asarray(a, dtype=np.float64), shape).copy())
    o.set2d("ustar", np.broadcast_to(np.asarray(vals.get("ustar", 0.0)), (NY, NX)).copy())


def test_time_variance_resolved_stress_and_level_one():
    o, _ = _oracle()
    o.stat_reset()
    a, b, c, d, e = 3.0, 0.5, -1.25, 0.75, 0.25
    kk = np.arange(NZ + 2, dtype=np.float64)[:, None, None]
    for sgn in (1.0, -1.0):
        # u(k, t) = a + sgn * k b: the w-node average must use levels k-1, k
        _put_all(o, {"u": a + sgn * kk * b, "v": c + sgn * d, "w": sgn * e, "txz": 1.0 + sgn,
                     "tyz": -2.0, "txx": 4.0, "cs_opt2": (0.25 if sgn > 0 else 0.0625), "nu_t": 2.0 + sgn,
                     "ustar": 1.0 + 0.5 * sgn})
        o.stat_accumulate()
    sums, us, it = o.stat_get()
    assert it == 2 and np.all(us == 2.0)
    k = np.arange(1, NZ + 1)
    assert np.array_equal(sums["u"][:, 0, 0], 2 * a + 0 * k)
    assert np.array_equal(sums["uw"][0], np.zeros((NY, NX)))  # reset at level 1
    pr = o.stat_profiles()
    assert np.allclose(pr["u"], a, rtol=0, atol=1e-15)
    assert np.allclose(pr["var_u"], (k * b) ** 2, rtol=1e-14, atol=1e-14)
    assert np.allclose(pr["var_v"], d * d, rtol=1e-14)
    assert np.allclose(pr["var_w"], e * e, rtol=1e-14)
    # R13(k) = <w 0.5 (u(k-1) + u(k))> - u_w w = e b (k - 1/2) above level 1, 0 at level 1
    r13 = np.where(k == 1, 0.0, e * b * (k - 0.5))
    assert np.allclose(pr["R13"], r13, rtol=1e-14, atol=1e-15)
    assert np.allclose(pr["R23"], np.where(k == 1, 0.0, e * d), rtol=1e-14, atol=1e-15)
    assert np.allclose(pr["txz"], -1.0) and np.allclose(pr["tyz"], 2.0) and np.allclose(pr["txx"], -4.0)
    # uniform fields: no dispersive stress; St13 = R13 + txz
    assert np.allclose(pr["C13"], 0.0, atol=1e-14) and np.allclose(pr["St13"], r13 - 1.0, atol=1e-14)
    # tke with var_w on the uvp node (0 at the top level)
    varw_uvp = np.where(k < NZ, e * e, 0.0)
    assert np.allclose(pr["tke"], 0.5 * ((k * b) ** 2 + d * d + varw_uvp), rtol=1e-14)
    assert np.allclose(pr["cs"], 0.5 * (0.5 + 0.25)) and np.allclose(pr["nut"], 2.0)


def test_dispersive_stress_of_two_cosines():
    o, _ = _oracle()
    o.stat_reset()
    x = 2 * np.pi * np.arange(NX) / NX
    ub, bb, cc = 2.0, 0.5, 0.75
    u = np.broadcast_to(ub + bb * np.cos(x), (NZ + 2, NY, NX))
    w = np.broadcast_to(cc * np.cos(x), (NZ + 2, NY, NX))
    for _ in range(3):
        _put_all(o, {"u": u, "w": w})
        o.stat_accumulate()
    pr = o.stat_profiles()
    k = np.arange(1, NZ + 1)
    # C13 = <u_w w>_xy - <u_w>_xy <w>_xy = b c / 2 above level 1 (u_w = 0 at level 1)
    assert np.allclose(pr["C13"], np.where(k == 1, 0.0, 0.5 * bb * cc), atol=1e-14)
    # steady fields: no resolved (temporal) stress, no variance
    assert np.allclose(pr["R13"], 0.0, atol=1e-14) and np.allclose(pr["var_u"], 0.0, atol=1e-14)
    assert np.allclose(pr["St13"], pr["C13"], atol=1e-14)


def _numpy_sums(snaps):
    """Independent statement of io_3D_stat_averages' sums over snapshots."""
    n = {}
    for f in O.Oracle.STATS:
        n[f] = np.zeros((NZ, NY, NX))
    for s in snaps:
        u, v, w = s["u"], s["v"], s["w"]
        I = slice(1, NZ + 1)
        ui, vi, wi = u[I], v[I], w[I]
        n["u"] = n["u"] + ui
        n["v"] = n["v"] + vi
        n["w"] = n["w"] + wi
        n["p"] = n["p"] + s["p"][I]
        for q, x in (("u", ui), ("v", vi), ("w", wi)):
            n[q + "2"] = n[q + "2"] + x * x
            n[q + "3"] = n[q + "3"] + (x * x) * x
            n[q + "4"] = n[q + "4"] + (x * x) * (x * x)
        n["uv"] = n["uv"] + ui * vi
        uw, vw = np.zeros((NZ, NY, NX)), np.zeros((NZ, NY, NX))
        uw[1:] = n["uw"][1:] + w[2:NZ + 1] * (0.5 * (u[1:NZ] + u[2:NZ + 1]))
        vw[1:] = n["vw"][1:] + w[2:NZ + 1] * (0.5 * (v[1:NZ] + v[2:NZ + 1]))
        n["uw"], n["vw"] = uw, vw
        for f, src in (("txx", "txx"), ("tyy", "tyy"), ("tzz", "tzz"), ("txy", "txy"), ("txz", "txz"),
                       ("tyz", "tyz"), ("dudz", "dudz"), ("dvdz", "dvdz"), ("nut", "nu_t")):
            n[f] = n[f] + s[src][I]
        n["cs"] = n["cs"] + np.sqrt(s["cs_opt2"][I])
    return n


def test_sums_match_independent_statement_on_random_fields():
    o, _ = _oracle()
    o.stat_reset()
    rng = np.random.default_rng(5)
    snaps = []
    for _ in range(3):
        s = {f: rng.standard_normal((NZ + 2, NY, NX)) for f in
             ["u", "v", "w", "p", "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nu_t"]}
        s["cs_opt2"] = rng.random((NZ + 2, NY, NX))
        _put_all(o, s)
        o.stat_accumulate()
        snaps.append(s)
    sums, _, it = o.stat_get()
    ref = _numpy_sums(snaps)
    assert it == 3
    for f in O.Oracle.STATS:
        assert np.array_equal(sums[f], ref[f]), f


def test_reset_clears_sums_and_count():
    o, _ = _oracle()
    o.stat_reset()
    _put_all(o, {"u": 1.0})
    o.stat_accumulate()
    o.stat_reset()
    sums, us, it = o.stat_get()
    assert it == 0 and not np.any(sums["u"]) and not np.any(us)
    with pytest.raises(ValueError):
        o.stat_profiles()


def test_running_diagnostics_closed_forms():
    dt = 2e-4
    o, prm = _oracle(dt)
    x = 2 * np.pi * np.arange(NX) / NX
    a, b, c, e = 5.0, -1.5, 0.75, 0.5
    u = np.broadcast_to(a + b * np.cos(x), (NZ + 2, NY, NX)).copy()
    u[0] = 1e3      # ghost levels are outside every reduction
    u[NZ + 1] = -1e3
    w = np.full((NZ + 2, NY, NX), e)
    w[NZ + 1] = 0.0  # the top level enters only the uvp-node average
    w[0] = 7e2
    dwdz = np.broadcast_to(np.arange(NZ + 2, dtype=np.float64)[:, None, None], (NZ + 2, NY, NX))
    ustar = np.random.default_rng(2).random((NY, NX))
    _put_all(o, {"u": u, "v": c, "w": w, "dudx": 1.0, "dvdy": 2.0, "dwdz": dwdz, "ustar": ustar})
    d = o.diagnostics()
    dx, dy, dz = prm.lx / NX, prm.ly / NY, prm.lz / NZ
    assert d["cfl"] == pytest.approx(dt * ((abs(a) + abs(b)) / dx + abs(c) / dy + e / dz), rel=1e-14)
    wmean = e * (NZ - 0.5) / NZ
    assert d["mke"] == pytest.approx(0.5 * (a * a + c * c + wmean * wmean), rel=1e-13)
    assert d["u_star"] == pytest.approx(ustar.mean(), rel=1e-14)
    k = np.arange(1, NZ + 1)
    assert d["divu"] == pytest.approx(np.sqrt(np.mean((3.0 + k) ** 2)), rel=1e-14)
