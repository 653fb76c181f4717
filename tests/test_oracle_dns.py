"""Pinning the oracle's dns branch of sgs_stress (src/compute_sgs.f90:124-140),
CPU only: the molecular stress tau_ij = 2/Re S_ij on levels 1..nz.

Pinned by a trigonometric velocity whose spectral derivatives are known in
closed form (so the stresses are known to round-off), and by what the branch
must leave alone: Cs^2, nu_t, the S_ij fields and the levels outside 1..nz.
"""
import numpy as np

from conftest import load_oracle, load_params

O = load_oracle()
P = load_params()


def test_dns_stress_of_a_trigonometric_field():
    Re = 2.5e3
    prm = P.LesParams(nx=32, ny=16, nz=5, sgs_model=P.SGS_DNS, re=Re)
    prm.validate()
    o = O.Oracle(prm)
    shape = (prm.nz + 2, prm.ny, prm.nx)
    x = (np.arange(prm.nx) * prm.lx / prm.nx)[None, None, :]
    y = (np.arange(prm.ny) * prm.ly / prm.ny)[None, :, None]
    o.put("u", np.broadcast_to(np.sin(2 * x) + 0.5 * np.cos(y), shape))
    o.put("v", np.broadcast_to(np.cos(3 * x) * np.sin(y), shape))
    o.put("w", np.broadcast_to(0.25 * np.sin(x + y), shape))
    o.op("vel_grad")
    marker = 7.0
    for n in ("txx", "txy", "txz", "tyy", "tyz", "tzz", "nu_t", "cs_opt2", "s11"):
        o.put(n, np.full(shape, marker))
    dudz, dvdz = o.get("dudz"), o.get("dvdz")
    o.sgs_stress(3, 3)
    c = 2.0 / Re
    I = slice(1, prm.nz + 1)
    dudx, dudy = 2 * np.cos(2 * x), -0.5 * np.sin(y)
    dvdx, dvdy = -3 * np.sin(3 * x) * np.sin(y), np.cos(3 * x) * np.cos(y)
    dwdx = dwdy = 0.25 * np.cos(x + y)
    tol = 1e-12 * c
    assert np.allclose(o.get("txx")[I], c * dudx, atol=tol, rtol=0)
    assert np.allclose(o.get("txy")[I], c * 0.5 * (dudy + dvdx), atol=tol, rtol=0)
    assert np.allclose(o.get("tyy")[I], c * dvdy, atol=tol, rtol=0)
    assert np.allclose(o.get("txz")[I], c * 0.5 * (dudz[I] + dwdx), atol=tol, rtol=0)
    assert np.allclose(o.get("tyz")[I], c * 0.5 * (dvdz[I] + dwdy), atol=tol, rtol=0)
    assert np.allclose(o.get("tzz")[I], c * o.get("dwdz")[I], atol=0, rtol=1e-15)
    for n in ("txx", "txy", "txz", "tyy", "tyz", "tzz"):
        a = o.get(n)
        assert np.all(a[0] == marker) and np.all(a[prm.nz + 1] == marker), n
    for n in ("nu_t", "cs_opt2", "s11"):
        assert np.all(o.get(n) == marker), n


def test_dns_steps_run():
    from conftest import load_ic
    IC = load_ic()
    prm = P.LesParams(nx=16, ny=16, nz=6, sgs_model=P.SGS_DNS, re=1e4)
    o = O.Oracle(prm)
    u, v, w = IC.log_law_ic(16, 16, 6, prm.lz, prm.z0, 0.1, seed=3)
    o.set_ic(u, v, w)
    o.step(4)
    assert np.all(np.isfinite(o.get("u"))) and np.all(o.get("cs_opt2") == 0)
