"""Pinning the oracle's other dealiasing options of the convective term (CPU).

The reference picks the dealiasing at compile time (DSLG, src/Makefile:17-40):
TRUNC (2/3 rule, src/compute_convective_trunc.f90), FSMTHG (Fourier
smoothing, src/compute_convective_Fsmthg.f90) and NODLSG (none,
src/compute_convective_nodlsg.f90) beside the default PADD.  None has tests or
stored outputs in the reference.  They are pinned by

* the multiplier tables against the reference's index arithmetic, worked by
  hand for grids where the source and product cut-offs differ (96 x 48);
* band-limited fields (all products below every cut-off): TRUNC and FSMTHG
  equal the alias-free PADD result;
* a single high mode v = B sin 9x on 32 points whose product sin 18x aliases
  to -sin 14x: PADD and TRUNC remove it, FSMTHG keeps it scaled by
  F(9)^2 F(14), NODLSG keeps the raw product times const^3.
"""
import numpy as np
import pytest

from conftest import load_oracle, load_params

O = load_oracle()
P = load_params()

DERIVS = ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")


def _oracle(nx, ny, nz, dealias):
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, dealias=dealias)
    prm.validate()
    return prm, O.Oracle(prm)


def _F(m, lh):
    f = np.exp(-36.0 * (np.asarray(m, dtype=float) / (lh - 1)) ** 36.0)
    return np.where(np.asarray(m) == lh - 1, 0.0, f)


def test_trunc_tables_follow_the_reference_cutoffs():
    # sources: tt = ceiling(n/3) zeroes modes >= tt-1; products: tt = ceiling(2 lh/3)
    prm, o = _oracle(96, 48, 3, P.DEALIAS["TRUNC"])
    ga, gb = o.table("GA"), o.table("GB")
    # x: 96 -> sources keep kx <= 30, products (lhx = 49, tt = 33) keep kx <= 31
    assert ga[0, 30] == 1 and ga[0, 31] == 0 and np.all(ga[0, 31:] == 0)
    assert gb[0, 31] == 1 and gb[0, 32] == 0
    # y: 48 -> sources keep |ky| <= 14, products (lhy = 25, tt = 17) keep |ky| <= 15
    assert ga[14, 0] == 1 and ga[15, 0] == 0 and ga[48 - 14, 0] == 1 and ga[48 - 15, 0] == 0
    assert gb[15, 0] == 1 and gb[16, 0] == 0 and gb[48 - 15, 0] == 1 and gb[48 - 16, 0] == 0
    # separable 0/1 multiplier without normalisation
    assert set(np.unique(ga)) == {0.0, 1.0}
    assert np.array_equal(ga, np.outer(ga[:, 0], ga[0, :]))


def test_fsmthg_tables_are_the_smoothing_function():
    prm, o = _oracle(32, 48, 3, P.DEALIAS["FSMTHG"])
    ga, gb = o.table("GA"), o.table("GB")
    assert np.array_equal(ga, gb)
    lhx, lhy = 17, 25
    fx, fy = _F(np.arange(lhx), lhx), _F(np.arange(lhy), lhy)
    ky = np.minimum(np.arange(48), 48 - np.arange(48))
    np.testing.assert_allclose(ga, np.outer(fy[ky], fx), rtol=1e-15, atol=0)
    assert ga[0, 0] == 1.0 and np.all(ga[:, -1] == 0) and np.all(ga[24, :] == 0)
    assert np.all(np.diff(fx) <= 0) and 0.7 < fx[14] < 0.8


def test_padd_has_no_dealias_tables():
    prm, o = _oracle(16, 16, 3, P.DEALIAS["PADD"])
    with pytest.raises(KeyError):
        o.table("GA")


def _put(o, prm, fields):
    shape = (prm.nz + 2, prm.ny, prm.nx)
    for n in ("u", "v", "w") + DERIVS:
        o.put(n, np.broadcast_to(fields.get(n, 0.0), shape))


def _convect(dealias, fields, nx=32, ny=32, nz=3):
    prm, o = _oracle(nx, ny, nz, dealias)
    _put(o, prm, fields)
    o.op("conv")
    return prm, [o.get(n)[1:nz + 1] for n in ("rhsx", "rhsy", "rhsz")]


def _band_limited(nx, ny, kmax, rng):
    x = (np.arange(nx) * 2 * np.pi / nx)[None, :]
    y = (np.arange(ny) * 2 * np.pi / ny)[:, None]
    f = np.zeros((ny, nx))
    for kx in range(kmax + 1):
        for ky in range(-kmax, kmax + 1):
            a, b = rng.standard_normal(2)
            f = f + a * np.cos(kx * x + ky * y) + b * np.sin(kx * x + ky * y)
    return f


@pytest.mark.parametrize("name", ["TRUNC", "FSMTHG"])
def test_band_limited_fields_match_the_alias_free_padded_result(name):
    rng = np.random.default_rng(1)
    f = {n: _band_limited(32, 32, 3, rng) for n in ("u", "v", "w") + DERIVS}
    # products reach |k| <= 6, below every cut-off (TRUNC keeps <= 9 / <= 10
    # on 32 points; FSMTHG's F(m <= 6) differs from 1 by < 2e-14)
    _, ref = _convect(P.DEALIAS["PADD"], f)
    _, got = _convect(P.DEALIAS[name], f)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_aliased_mode_per_option():
    B = 0.7
    x = (np.arange(32) * 2 * np.pi / 32)[None, None, :]
    f = {"v": B * np.sin(9 * x), "dvdx": 9 * B * np.cos(9 * x)}
    c = 1.0 / (32 * 32)
    raw = (B * np.sin(9 * x)) * (9 * B * np.cos(9 * x))   # v omega_3 = 4.5 B^2 sin 18x
    alias = -4.5 * B * B * np.sin(14 * x)                  # sin 18x on 32 points
    assert np.allclose(raw, alias, atol=1e-13)
    out = {}
    for name in ("PADD", "TRUNC", "FSMTHG", "NODLSG"):
        _, out[name] = _convect(P.DEALIAS[name], f)
        assert np.abs(out[name][1]).max() == 0 and np.abs(out[name][2]).max() == 0, name
    assert np.abs(out["PADD"][0]).max() < 1e-13
    assert np.abs(out["TRUNC"][0]).max() < 1e-13
    smooth = _F(9, 17) ** 2 * _F(14, 17) * alias
    np.testing.assert_allclose(out["FSMTHG"][0], np.broadcast_to(smooth, out["FSMTHG"][0].shape), rtol=0,
                               atol=1e-12 * np.abs(smooth).max())
    assert np.abs(smooth).max() > 1.0
    # no dealiasing: the scaled sources' product times const, as the reference forms it
    np.testing.assert_allclose(out["NODLSG"][0], np.broadcast_to(c ** 3 * raw, out["NODLSG"][0].shape),
                               rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", ["TRUNC", "FSMTHG", "NODLSG"])
def test_steps_run_and_differ_from_padd(name):
    from conftest import load_ic
    IC = load_ic()
    res = {}
    for opt in ("PADD", name):
        prm, o = _oracle(16, 16, 6, P.DEALIAS[opt])
        u, v, w = IC.log_law_ic(16, 16, 6, prm.lz, prm.z0, 0.1, seed=3)
        o.set_ic(u, v, w)
        o.step(4)
        res[opt] = o.get("u")[1:7]
        assert np.all(np.isfinite(res[opt]))
    assert not np.array_equal(res["PADD"], res[name])
