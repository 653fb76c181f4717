"""Pinning the oracle's plane-averaged dynamic Smagorinsky model (CPU only).

compute_DynSmag_coeff / compute_dynamic_mum (src/compute_DynSmag.f90:61-245)
have no tests or stored outputs in the reference.  They are pinned by

* a closed form: u = a cos 3x + e cos 9x on a 32 x 32 plane of width 2 pi
  (2-delta cutoff |k| < 8) gives L11 = e^2/2 + a e cos 6x; s11 = b cos 6x,
  s22 = b sin 6x give |S| = sqrt(2) |b| and M = -6 sqrt(2) delta^2 |b| S, so
  Cs^2 = -(sqrt(2)/24) a e sign(b) / (delta^2 b^2), clipped at 0.  A wrong
  filter width (4 delta keeps |k| < 4 and drops the cos 6x of L), a wrong
  factor 4 on |S~| S~, a missing clip or a per-point instead of plane
  average each change it;
* a brute-force restatement on a tiny random grid with dense DFT matrices
  (no FFT) for the test filter;
* exact power-of-two scalings: u -> 2u multiplies Cs^2 by 4 and S -> 2S
  divides it by 4, bit for bit;
* the schedule: 0.03 before dyn_init, an update only when the local step
  count jt is a multiple of cs_count, the coefficient kept otherwise, and a
  zero M:M plane giving 0 rather than a division by zero.
"""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params

O = load_oracle()
P = load_params()
IC = load_ic()


def _oracle(nx, ny, nz, **kw):
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, sgs_model=P.SGS_DYN_SMAG, **kw)
    prm.validate()
    return prm, O.Oracle(prm)


def _delta(prm):
    return (prm.lx / prm.nx * prm.ly / prm.ny * prm.lz / prm.nz) ** (1.0 / 3.0)


def _put_fields(o, prm, u, v, w, grads):
    shape = (prm.nz + 2, prm.ny, prm.nx)
    o.put("u", np.broadcast_to(u, shape))
    o.put("v", np.broadcast_to(v, shape))
    o.put("w", np.broadcast_to(w, shape))
    for n in ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz"):
        o.put(n, np.broadcast_to(grads.get(n, 0.0), shape))


def _closed_form_case(a, e, b, nz=3):
    prm, o = _oracle(32, 32, nz, cs_count=1)
    x = (np.arange(prm.nx) * prm.lx / prm.nx)[None, None, :]
    u = a * np.cos(3 * x) + e * np.cos(9 * x)
    _put_fields(o, prm, u, 0.0, 0.0, {"dudx": b * np.cos(6 * x), "dvdy": b * np.sin(6 * x)})
    o.sgs_stress(1, 1)
    return prm, o


def test_closed_form_coefficient_and_eddy_viscosity():
    a, e, b = 1.5, -0.8, 0.7
    prm, o = _closed_form_case(a, e, b)
    d = _delta(prm)
    ref = -(np.sqrt(2.0) / 24.0) * a * e / (d * d * b * b)
    cs = o.get("cs_opt2")[1:prm.nz + 2]
    assert ref > 0
    for k in range(prm.nz + 1):
        assert np.all(cs[k] == cs[k, 0, 0]), k  # one value per level
        assert cs[k, 0, 0] == pytest.approx(ref, rel=1e-11), k
    # dynamic models use the filter width itself as mixing length
    nut = o.get("nu_t")[1:prm.nz + 2]
    np.testing.assert_allclose(nut, np.sqrt(2.0) * b * cs * d * d, rtol=1e-12)
    # ghost levels untouched
    assert np.all(o.get("cs_opt2")[0] == 0)


def test_negative_coefficient_clipped_to_zero():
    prm, o = _closed_form_case(1.5, -0.8, -0.7)
    assert np.all(o.get("cs_opt2")[1:prm.nz + 2] == 0.0)
    assert np.all(o.get("nu_t")[1:prm.nz + 2] == 0.0)


def test_zero_strain_gives_zero_not_nan():
    prm, o = _oracle(16, 16, 3, cs_count=1)
    x = (np.arange(prm.nx) * prm.lx / prm.nx)[None, None, :]
    _put_fields(o, prm, np.cos(x) + 0.3 * np.cos(7 * x), 0.2, 0.0, {})
    o.sgs_stress(1, 1)
    cs = o.get("cs_opt2")
    assert np.all(np.isfinite(cs)) and np.all(cs == 0.0)


def _dense_filter(f, G):
    """test_filter by dense DFT matrices: unnormalised forward transform,
    multiplication by the half-spectrum kernel G (which carries 1/(nx ny)),
    Hermitian extension, unnormalised inverse."""
    ny, nx = f.shape
    Wx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    Wy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    F = Wy @ f @ Wx.T
    Gf = np.zeros((ny, nx))
    lhx = nx // 2 + 1
    Gf[:, :lhx] = G
    for kx in range(lhx, nx):
        for j in range(ny):
            Gf[j, kx] = G[(-j) % ny, nx - kx]
    return (np.conj(Wy) @ (F * Gf) @ np.conj(Wx).T).real


def _brute_force_ratio(prm, G, u, v, w, s):
    """compute_dynamic_mum (src/compute_DynSmag.f90:96-243) on levels 1..nz+1,
    before the clip at 0."""
    nz = prm.nz
    d = _delta(prm)
    out = np.zeros(nz + 2)
    for k in range(1, nz + 2):
        if k == 1:
            uf, vf, wf = u[1], v[1], 0.25 * w[2]
        else:
            uf, vf, wf = 0.5 * (u[k] + u[k - 1]), 0.5 * (v[k] + v[k - 1]), w[k]
        vel = [uf, vf, wf]
        fv = [_dense_filter(q, G) for q in vel]
        idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
        Lm = {ij: _dense_filter(vel[ij[0]] * vel[ij[1]], G) - fv[ij[0]] * fv[ij[1]] for ij in idx}
        S = {ij: s[ij][k] for ij in idx}
        smod = np.sqrt(2 * (S[0, 0] ** 2 + S[1, 1] ** 2 + S[2, 2] ** 2
                            + 2 * (S[0, 1] ** 2 + S[0, 2] ** 2 + S[1, 2] ** 2)))
        Sf = {ij: _dense_filter(S[ij], G) for ij in idx}
        sfmod = np.sqrt(2 * (Sf[0, 0] ** 2 + Sf[1, 1] ** 2 + Sf[2, 2] ** 2
                             + 2 * (Sf[0, 1] ** 2 + Sf[0, 2] ** 2 + Sf[1, 2] ** 2)))
        M = {ij: 2 * d * d * (_dense_filter(smod * S[ij], G) - 4 * sfmod * Sf[ij]) for ij in idx}
        wgt = {ij: (1.0 if ij[0] == ij[1] else 2.0) for ij in idx}
        lm = sum(wgt[ij] * Lm[ij] * M[ij] for ij in idx).mean()
        mm = sum(wgt[ij] * M[ij] * M[ij] for ij in idx).mean()
        out[k] = lm / mm
    return out


@pytest.mark.parametrize("seed", [0, 1])
def test_brute_force_dense_dft_on_random_fields(seed):
    # S -> -S flips the sign of M and so of <L:M> exactly: one of the two
    # clipped coefficients is the ratio's magnitude and the other 0, which
    # recovers the unclipped ratio on every level
    prm, o = _oracle(12, 8, 4, cs_count=1)
    rng = np.random.default_rng(seed)
    shape = (prm.nz + 2, prm.ny, prm.nx)
    u, v, w = (rng.standard_normal(shape) for _ in range(3))
    grads = {n: rng.standard_normal(shape) for n in
             ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")}
    cs = {}
    for sgn in (1.0, -1.0):
        _put_fields(o, prm, u, v, w, {n: sgn * g for n, g in grads.items()})
        o.sgs_stress(5, 5)
        cs[sgn] = o.get("cs_opt2")
        if sgn > 0:
            s = {(0, 0): o.get("s11"), (0, 1): o.get("s12"), (0, 2): o.get("s13"), (1, 1): o.get("s22"),
                 (1, 2): o.get("s23"), (2, 2): o.get("s33")}
    ref = _brute_force_ratio(prm, o.table("G2"), u, v, w, s)
    assert np.any(ref[1:] > 0) and np.any(ref[1:] < 0)
    for k in range(1, prm.nz + 2):
        for sgn in (1.0, -1.0):
            assert np.all(cs[sgn][k] == cs[sgn][k, 0, 0])
            assert cs[sgn][k, 0, 0] >= 0
        assert min(cs[1.0][k, 0, 0], cs[-1.0][k, 0, 0]) == 0.0
        got = cs[1.0][k, 0, 0] - cs[-1.0][k, 0, 0]
        assert got == pytest.approx(ref[k], rel=1e-10), (k, got, ref[k])


def test_power_of_two_scalings_are_exact():
    prm, o = _oracle(16, 16, 3, cs_count=1)
    rng = np.random.default_rng(4)
    shape = (prm.nz + 2, prm.ny, prm.nx)
    u, v, w = (rng.standard_normal(shape) for _ in range(3))
    grads = {n: rng.standard_normal(shape) for n in ("dudx", "dudy", "dvdx", "dvdy", "dwdz", "dudz", "dvdz")}

    def cs_of(scale_u, scale_s):
        _put_fields(o, prm, scale_u * u, scale_u * v, scale_u * w, {n: scale_s * g for n, g in grads.items()})
        o.sgs_stress(1, 1)
        return o.get("cs_opt2")[1:prm.nz + 2, 0, 0].copy()

    base = cs_of(1.0, 1.0)
    assert np.any(base > 0)
    assert np.array_equal(cs_of(2.0, 1.0), 4.0 * base)
    assert np.array_equal(cs_of(1.0, 2.0), base / 4.0)


def test_schedule_initial_value_and_local_step_count():
    prm, o = _oracle(16, 16, 4, cs_count=3, dyn_init=4)
    u, v, w = IC.log_law_ic(prm.nx, prm.ny, prm.nz, prm.lz, prm.z0, 0.1, seed=3)
    o.set_ic(u, v, w)
    o.step(1)  # jtt = 1 < dyn_init: Cs^2 = 0.03 on levels 1..nz+1
    cs = o.get("cs_opt2")
    assert np.all(cs[1:prm.nz + 2] == 0.03) and np.all(cs[0] == 0)
    o.step(3)  # jt = 4: not a multiple of cs_count, jtt = 4 >= dyn_init: kept
    assert np.all(o.get("cs_opt2")[1:prm.nz + 2] == 0.03)
    assert o.jt() == 4 and o.jt_total() == 4
    # the update keys on jt, the reset on jtt: after a restart to jtt = 100
    # the coefficient is recomputed when jt (not jtt) reaches a multiple of 3
    o.set_jt_total(100)
    o.step(1)  # jt = 5, jtt = 101
    assert np.all(o.get("cs_opt2")[1:prm.nz + 2] == 0.03)
    o.step(1)  # jt = 6, jtt = 102: update
    cs = o.get("cs_opt2")[1:prm.nz + 2]
    assert not np.all(cs == 0.03)
    assert all(np.all(cs[k] == cs[k, 0, 0]) for k in range(prm.nz + 1))
    assert cs.min() >= 0
