"""Closed-form checks that pin the oracle's LASD, SGS-stress, convection,
pressure and time-advance functions (CPU only).

The reference ships no tests, golden vectors or fixtures for this path and
cannot be built here, so each function of the oracle is pinned by what the
mathematics fixes for it: a closed form on fields chosen so the answer is
known exactly (single Fourier modes below a filter cut-off, fields linear in
an index, uniform fields, a unidirectional shear flow), or a scaling law.
Every check here takes the path of an oracle library, so that
tests/test_oracle_pins.py can run it both on the oracle the other tests use
and on builds of it with a planted mistake (les_oracle.cpp, ORC_PLANT), where
it must fail.  PLANTED maps each planted mistake to the checks that must
catch it.
"""
import numpy as np

from conftest import load_ic, load_oracle, load_params

O = load_oracle()
P = load_params()
IC = load_ic()

KAPPA = 0.4


def _mk(lib, nx=16, ny=16, nz=8, **kw):
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, **kw)
    prm.validate()
    return prm, O.Oracle(prm, lib)


def _shape(prm):
    return (prm.nz + 2, prm.ny, prm.nx)


def _delta(prm):
    return (prm.lx / prm.nx * prm.ly / prm.ny * prm.lz / prm.nz) ** (1.0 / 3.0)


def _xy(prm):
    x = (np.arange(prm.nx) * prm.lx / prm.nx)[None, None, :]
    y = (np.arange(prm.ny) * prm.ly / prm.ny)[None, :, None]
    return x, y


def _close(got, ref, tol, what):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max() / scale
    assert np.isfinite(got).all() and err <= tol, f"{what}: max error {err:.3e} (relative to {scale:.3e}) > {tol}"


def _same(got, ref, what):
    """Bit-for-bit equality (inf and NaN where the reference arithmetic makes
    them, e.g. Cs^2 on a lid with M:M = 0, count as equal)."""
    assert np.array_equal(np.asarray(got), np.asarray(ref), equal_nan=True), f"{what}: not bit-for-bit equal"


def _zero_state(o, prm, names=("u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy",
                               "dwdz")):
    for n in names:
        o.put(n, np.zeros(_shape(prm)))


# ---------------------------------------------------------------- update_interpolag_coeff
# src/compute_LagScaleDep.f90:607-640: T = 1.5 delta (F_AB F_BB)^(-1/8)
# (the Lagrangian time scale of Meneveau, Lund & Cabot 1996 with theta = 1.5),
# eps = (dt cs_count / T) / (1 + dt cs_count / T), F <- eps new + (1 - eps) F.

def _relax_factor(lib, fab, fbb, cs_count=5, dt=2e-4):
    """dt cs_count / T read back from one relaxation towards AB = BB = 0:
    F_BB (not cast) becomes (1 - eps) F_BB."""
    prm, o = _mk(lib, 16, 16, 4, cs_count=cs_count, dt=dt)
    sh = _shape(prm)
    o.put("F_LM", np.full(sh, fab))
    o.put("F_MM", np.full(sh, fbb))
    o.update_F(0, np.zeros(sh), np.zeros(sh))
    one_m_eps = o.get("F_MM")[1:prm.nz + 2] / fbb
    eps = 1.0 - one_m_eps
    return eps / one_m_eps, prm, o


def pin_update_F_fixed_point(lib):
    """AB = F_AB and BB = F_BB leave F unchanged whatever eps is."""
    prm, o = _mk(lib, 16, 16, 4)
    sh = _shape(prm)
    rng = np.random.default_rng(1)
    fab = rng.uniform(1e-3, 1.0, sh).astype(np.float32).astype(np.float64)
    fbb = rng.uniform(1e-3, 1.0, sh)
    o.put("F_LM", fab)
    o.put("F_MM", fbb)
    o.update_F(0, fab, fbb)
    _close(o.get("F_LM")[1:], fab[1:], 0.0, "F_LM fixed point")
    _close(o.get("F_MM")[1:], fbb[1:], 4e-16, "F_MM fixed point")


def pin_update_F_time_scale_exponent(lib):
    """F_AB F_BB times 2^8 halves T: the relaxation factor dt cs_count / T
    doubles (exponent -1/8); times 2^-16 it is quartered."""
    r0, _, _ = _relax_factor(lib, 0.5, 0.25)
    r1, _, _ = _relax_factor(lib, 0.5 * 16.0, 0.25 * 16.0)
    r2, _, _ = _relax_factor(lib, 0.5 / 16.0, 0.25 / 16.0)
    # T is rounded to float32 once (src/compute_LagScaleDep.f90:626)
    _close(r1 / r0, np.full_like(r0, 2.0), 2e-7, "relaxation factor ratio for F_AB F_BB x 2^8")
    _close(r2 / r0, np.full_like(r0, 0.5), 2e-7, "relaxation factor ratio for F_AB F_BB x 2^-8")


def pin_update_F_time_scale_constant(lib):
    """F_AB F_BB = 1: T = 1.5 delta, so dt cs_count / T = dt cs_count / (1.5
    delta); proportional to cs_count and to dt."""
    for cs, dt in ((5, 2e-4), (3, 2e-4), (1, 1e-4)):
        r, prm, _ = _relax_factor(lib, 2.0, 0.5, cs_count=cs, dt=dt)
        T = float(np.float32(1.5 * _delta(prm)))
        # eps / (1 - eps) read back through 1 - F_MM / F_MM(before): ~1e-16 / eps
        _close(r, np.full_like(r, dt * cs / T), 2e-12, f"relaxation factor at cs_count={cs}, dt={dt}")


def pin_update_F_casts_and_clip(lib):
    """F_AB is stored as a float32 value and clipped at float32(1e-24) when the
    blend goes negative; F_BB is not cast."""
    prm, o = _mk(lib, 16, 16, 4)
    sh = _shape(prm)
    rng = np.random.default_rng(2)
    fab = rng.uniform(0.1, 1.0, sh)
    fbb = rng.uniform(0.1, 1.0, sh)
    ab = rng.uniform(-1.0, 1.0, sh) * 1e5
    bb = rng.uniform(0.1, 1.0, sh) * np.pi
    o.put("F_LM", fab)
    o.put("F_MM", fbb)
    o.update_F(0, ab, bb)
    flm = o.get("F_LM")[1:]
    fmm = o.get("F_MM")[1:]
    assert np.array_equal(flm.astype(np.float32).astype(np.float64), flm), "F_LM not float32-valued"
    neg = ab[1:] < -1e4  # eps ~ 2e-3 here, so the blend is negative
    assert neg.any() and np.all(flm[neg] == np.float64(np.float32(1e-24))), "negative blend not clipped to 1e-24"
    assert not np.array_equal(fmm.astype(np.float32).astype(np.float64), fmm), "F_MM was rounded to float32"


# ---------------------------------------------------------------- compute_dynamic_coeff
# src/compute_LagScaleDep.f90:651-707: with the scale-dependent assumption
# Cs^2(2 delta) = beta Cs^2(delta), Cs^2(4 delta) = beta^2 Cs^2(delta):
# beta = c4 / c2 (clipped at 1/8, 1 on the lid) and Cs^2 = c2 / beta.

def pin_dynamic_coeff_closed_form(lib):
    prm, o = _mk(lib, 16, 16, 5)
    sh = _shape(prm)
    # per level (1..nz+1): c2 = F_LM / F_MM, c4 = F_QN / F_NN, exact binary values
    c2 = np.array([0.0, 2.0 ** -5, 2.0 ** -6, 2.0 ** -4, 2.0 ** -5, 2.0 ** -5, 2.0 ** -4])
    c4 = np.array([0.0, 2.0 ** -6, 2.0 ** -7, 2.0 ** -9, 2.0 ** -5 * 0.1875, 2.0 ** -3, 2.0 ** -6])
    mm = 2.0 ** np.arange(-3, 4, dtype=float)
    nn = 2.0 ** np.arange(2, -5, -1, dtype=float)
    o.put("F_LM", np.broadcast_to((c2 * mm)[:, None, None], sh))
    o.put("F_MM", np.broadcast_to(mm[:, None, None], sh))
    o.put("F_QN", np.broadcast_to((c4 * nn)[:, None, None], sh))
    o.put("F_NN", np.broadcast_to(nn[:, None, None], sh))
    o.op("dynamic_coeff")
    beta = o.get("beta")[1:, 0, 0]
    cs = o.get("cs_opt2")[1:, 0, 0]
    for k in range(1, prm.nz + 2):
        ratio = c4[k] / c2[k]
        if k == prm.nz + 1:
            eb, ec = 1.0, c2[k]
        elif ratio >= 0.125:
            eb, ec = ratio, c2[k] ** 2 / c4[k]
        else:
            eb, ec = 0.125, 8.0 * c2[k]
        _close(beta[k - 1], eb, 1e-7, f"beta at level {k}")
        _close(cs[k - 1], ec, 1e-7, f"Cs^2 at level {k}")
    # backscatter: F_LM < 0 clips c2 to float32(1e-24), so Cs^2 = float32(1e-24)
    o.put("F_LM", np.full(sh, -1.0))
    o.op("dynamic_coeff")
    assert np.all(o.get("cs_opt2")[1:prm.nz + 1] == np.float64(np.float32(1e-24)))


# ---------------------------------------------------------------- compute_dynamic_mum
# src/compute_LagScaleDep.f90:485-596 at 2 delta (G_test, tf^2 = 4) and
# 4 delta (G_test_test, tf^2 = 16).  On a 32 x 32 plane of width 2 pi the test
# filters keep |k| < 8 (2 delta) and |k| < 4 (4 delta).

def _germano_case(lib, a, e, b, c, strain):
    prm, o = _mk(lib, 32, 32, 3, cs_count=1)
    sh = _shape(prm)
    x, _ = _xy(prm)
    _zero_state(o, prm, ("u", "v", "w", "s11", "s12", "s13", "s22", "s23", "s33"))
    u = np.broadcast_to(a * np.cos(x) + e * np.cos(2 * x), sh)
    w = np.broadcast_to(c * np.cos(2 * x), sh).copy()
    w[1] = 0.0
    o.put("u", u)
    o.put("w", w)
    if strain == "horizontal":
        o.put("s11", np.broadcast_to(b * np.cos(3 * x), sh))
        o.put("s22", np.broadcast_to(b * np.sin(3 * x), sh))
    else:
        o.put("s13", np.broadcast_to(b * np.cos(3 * x), sh))
        o.put("s23", np.broadcast_to(b * np.sin(3 * x), sh))
    return prm, o, x[0]


def pin_germano_4delta_closed_form(lib):
    """u = a cos x + e cos 2x, S11 = b cos 3x, S22 = b sin 3x (|S| = sqrt2 |b|):
    at 4 delta only the cos 4x part of u^2 is removed, so L11 = -(e^2/2) cos 4x
    and M_ii = 2 delta^2 (1 - 16) sqrt2 |b| S_ii: QN = 15 sqrt2 delta^2 e^2 |b| b
    cos 4x cos 3x, NN = 1800 delta^4 b^4; at 2 delta nothing is removed:
    LM = 0, MM = 72 delta^4 b^4."""
    a, e, b = 0.7, -1.3, 0.45
    prm, o, x = _germano_case(lib, a, e, b, 0.0, "horizontal")
    LM, MM, QN, NN = o.dynamic_mum()
    d2 = _delta(prm) ** 2
    lev = slice(1, prm.nz + 2)
    qn = 15 * np.sqrt(2) * d2 * e ** 2 * abs(b) * b * np.cos(4 * x) * np.cos(3 * x)
    _close(QN[lev], np.broadcast_to(qn, QN[lev].shape), 1e-11, "QN (4 delta)")
    _close(NN[lev], np.full(NN[lev].shape, 1800 * d2 ** 2 * b ** 4), 1e-11, "NN (4 delta)")
    assert np.abs(LM[lev]).max() <= 1e-12 * np.abs(qn).max(), "LM (2 delta) should vanish"
    _close(MM[lev], np.full(MM[lev].shape, 72 * d2 ** 2 * b ** 4), 1e-11, "MM (2 delta)")


def pin_germano_first_node_w(lib):
    """S13 = b cos 3x, S23 = b sin 3x (|S| = 2|b|), u = a cos x + e cos 2x and
    w = c cos 2x on the w levels >= 2: QN = 2 L13 M13 = 60 c e delta^2 |b| b
    cos 4x cos 3x, except at the first node where w_ft = w(2) / 4 gives a
    quarter of it; NN = 7200 delta^4 b^4 and MM = 288 delta^4 b^4."""
    a, e, b, c = 0.9, 0.6, -0.35, 1.7
    prm, o, x = _germano_case(lib, a, e, b, c, "vertical")
    LM, MM, QN, NN = o.dynamic_mum()
    d2 = _delta(prm) ** 2
    qn = 60 * c * e * d2 * abs(b) * b * np.cos(4 * x) * np.cos(3 * x)
    _close(QN[2:prm.nz + 2], np.broadcast_to(qn, QN[2:prm.nz + 2].shape), 1e-11, "QN above the first node")
    _close(QN[1], np.broadcast_to(0.25 * qn, QN[1].shape), 1e-11, "QN at the first node (w_ft = w(2)/4)")
    _close(NN[1:], np.full(NN[1:].shape, 7200 * d2 ** 2 * b ** 4), 1e-11, "NN")
    _close(MM[1:], np.full(MM[1:].shape, 288 * d2 ** 2 * b ** 4), 1e-11, "MM")


# ---------------------------------------------------------------- Lagrangian interpolation
# src/compute_LagScaleDep.f90:309-474.  Trilinear interpolation reproduces a
# field linear in i, j, k exactly, so with uniform u_lag, v_lag, w_lag the
# interpolated F is F + bx xp + by yp + bz zp with xp = -u dt cs_count / dx
# (u_lag holding cs_count samples), except where the reference's special
# cases act: the first w node (zp doubled towards the wall at level 2,
# zp(1) <- clamp(zp(2))), the periodic x halo and the lid.

def _lagr_case(lib, xp, yp, zp, F0, bx, by, bz, cs_count=5):
    prm, o = _mk(lib, 16, 16, 8, cs_count=cs_count)
    sh = _shape(prm)
    dx, dy, dz = prm.lx / prm.nx, prm.ly / prm.ny, prm.lz / prm.nz
    cs, dt = prm.cs_count, prm.dt
    # u_lag holds cs_count samples of u: u = -xp dx / (dt cs)
    o.put("u_lag", np.full(sh, cs * (-xp * dx / (dt * cs))))
    o.put("v_lag", np.full(sh, cs * (-yp * dy / (dt * cs))))
    o.put("w_lag", np.full(sh, cs * (-zp * dz / (dt * cs))))
    k = np.arange(prm.nz + 2, dtype=float)[:, None, None]
    j = np.arange(1, prm.ny + 1, dtype=float)[None, :, None]
    i = np.arange(1, prm.nx + 1, dtype=float)[None, None, :]
    F = np.broadcast_to(F0 + bx * i + by * j + bz * k, sh).copy()
    o.put("F_LM", F)
    o.op("lagrangian")
    return prm, o.get("F_LM"), F


def pin_lagrangian_linear_interior(lib):
    xp, yp, zp = -0.3, 0.2, -0.4
    bx, by, bz = 0.37, -0.11, 0.53
    prm, got, F = _lagr_case(lib, xp, yp, zp, 5.0, bx, by, bz)
    inner = (slice(3, prm.nz + 1), slice(1, prm.ny - 1), slice(1, prm.nx - 1))
    _close(got[inner], F[inner] + bx * xp + by * yp + bz * zp, 1e-14, "interpolated linear field")


def pin_lagrangian_periodic_x(lib):
    """xp < 0 at i = 1 reads the periodic neighbour i = nx."""
    xp, bx = -0.3, 0.37
    prm, got, F = _lagr_case(lib, xp, 0.0, 0.0, 5.0, bx, 0.0, 0.0)
    lev = slice(3, prm.nz + 1)
    _close(got[lev, :, 0], (1 - abs(xp)) * F[lev, :, 0] + abs(xp) * F[lev, :, prm.nx - 1], 1e-14, "x halo at i = 1")
    _close(got[lev, :, 1:], F[lev, :, 1:] + bx * xp, 1e-14, "x interior")


def pin_lagrangian_first_nodes(lib):
    """Upward w (zp < 0): level 2 uses 2 zp towards level 1.  Downward w
    (zp > 0): level 1 uses zp(2) towards level 2 (zp(1) <- clamp(zp(2)),
    although w_lag(1) = w_lag(2) / 4)."""
    bz = 0.53
    zp = -0.2
    prm, got, F = _lagr_case(lib, 0.0, 0.0, zp, 5.0, 0.0, 0.0, bz)
    _close(got[2], F[2] + bz * 2 * zp, 1e-14, "level 2, upward w")
    _close(got[1], F[1], 1e-14, "level 1, upward w")
    _close(got[3:prm.nz + 2], F[3:prm.nz + 2] + bz * zp, 1e-14, "levels >= 3, upward w")
    zp = 0.3
    prm, got, F = _lagr_case(lib, 0.0, 0.0, zp, 5.0, 0.0, 0.0, bz)
    _close(got[1], F[1] + bz * zp, 1e-14, "level 1, downward w (zp(1) <- zp(2))")
    _close(got[2:prm.nz + 1], F[2:prm.nz + 1] + bz * zp, 1e-14, "levels 2..nz, downward w")
    _close(got[prm.nz + 1], F[prm.nz + 1], 1e-14, "lid, downward w")


# ---------------------------------------------------------------- LASD composition
# src/compute_LagScaleDep.f90:244-287: Germano at both widths, then on the
# first update F_MM = MM, F_LM = 0.03 MM (F_NN, F_QN alike) and on later ones
# the Lagrangian interpolation of the previous F, then the relaxation, then
# the coefficients.  Composed here from the parts pinned above.

def _lasd_state(lib, lagran_init):
    prm, o = _mk(lib, 16, 16, 8, cs_count=2, lagran_init=lagran_init)
    u, v, w = IC.log_law_ic(prm.nx, prm.ny, prm.nz, prm.lz, prm.z0, 0.1, seed=7)
    o.set_ic(u, v, w)
    o.op("wall_law")
    o.op("vel_grad")
    o.op("sij")
    return prm, o


def pin_lasd_update_composition(lib):
    for lagran_init in (0, 1):
        prm, o = _lasd_state(lib, lagran_init)
        sh = _shape(prm)
        rng = np.random.default_rng(3)
        Fs = {n: rng.uniform(0.5, 2.0, sh) * s for n, s in (("F_LM", 1e-3), ("F_MM", 1e-2), ("F_QN", 1e-3),
                                                              ("F_NN", 1e-2))}
        lag = {n: rng.uniform(-1.0, 1.0, sh) * 300.0 for n in ("u_lag", "v_lag", "w_lag")}
        state = {n: o.get(n) for n in ("u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx",
                                        "dwdy", "dwdz", "s11", "s12", "s13", "s22", "s23", "s33")}
        LM, MM, QN, NN = o.dynamic_mum()
        if lagran_init:
            for n, a in {**Fs, **lag}.items():
                o.put(n, a)
            # u_lag gains this step's u before the update
            for n, f in (("u_lag", "u"), ("v_lag", "v"), ("w_lag", "w")):
                o.put(n, lag[n] + state[f])
            o.op("lagrangian")
        else:
            o.put("F_LM", 0.03 * MM)
            o.put("F_MM", MM)
            o.put("F_QN", 0.03 * NN)
            o.put("F_NN", NN)
        o.update_F(0, LM, MM)
        o.update_F(1, QN, NN)
        o.op("dynamic_coeff")
        expect = {n: o.get(n) for n in ("F_LM", "F_MM", "F_QN", "F_NN", "beta", "cs_opt2")}
        prm2, o2 = _mk(lib, 16, 16, 8, cs_count=2, lagran_init=lagran_init)
        for n, a in state.items():
            o2.put(n, a)
        if lagran_init:
            for n, a in {**Fs, **lag}.items():
                o2.put(n, a)
        o2.sgs_stress(2, 2)
        for n, a in expect.items():
            _same(o2.get(n)[1:], a[1:], f"{n} after the {'second' if lagran_init else 'first'} update")


# ---------------------------------------------------------------- sgs_stress staggering
# src/compute_sgs.f90:176-242: nu_t on the w nodes; txx, txy, tyy, tzz on the
# uvp nodes with 2 * (nu_t(k) + nu_t(k+1)) / 2, except level 1 with nu_t(1);
# txz, tyz on the w nodes with nu_t(k).

def pin_tau_staggering(lib):
    """dudx = (-1)^k a (likewise dvdy, dwdz) averages to S11 = 0 above level 1,
    and dudz = g0 + g1 z on the w nodes gives |S| = dudz there, so nu_t is
    linear in z: txx at a uvp level is 2 nu_t(z_uvp) dudx."""
    prm, o = _mk(lib, 16, 16, 8, cs_count=5)
    sh = _shape(prm)
    _zero_state(o, prm)
    dz = prm.lz / prm.nz
    kk = np.arange(prm.nz + 2)
    sgn = ((-1.0) ** kk)[:, None, None]
    a, b, c = 0.8, -0.5, 0.3
    g0, g1 = 2.0, 3.0
    zw = (kk - 1) * dz
    o.put("dudx", np.broadcast_to(a * sgn, sh))
    o.put("dvdy", np.broadcast_to(b * sgn, sh))
    o.put("dwdz", np.broadcast_to(c * sgn, sh))
    o.put("dudz", np.broadcast_to((g0 + g1 * zw)[:, None, None], sh))
    cs2 = 0.02
    o.put("cs_opt2", np.full(sh, cs2))
    o.sgs_stress(1, 1)  # jtt = 1 < cs_count: no update, Cs^2 kept
    d2 = _delta(prm) ** 2
    nz = prm.nz
    nu = lambda z: cs2 * d2 * (g0 + g1 * z)  # noqa: E731
    k = np.arange(2, nz + 1)
    zuvp = (k - 0.5) * dz
    txx = o.get("txx")[:, 0, 0]
    _close(txx[2:nz + 1], 2 * nu(zuvp) * a * sgn[2:nz + 1, 0, 0], 1e-13, "txx on the uvp levels")
    s1 = np.sqrt(2 * (a * a + b * b + c * c) + g0 * g0)
    _close(txx[1], 2 * cs2 * d2 * s1 * (-a), 1e-13, "txx at level 1 (nu_t(1) alone)")
    txz = o.get("txz")[:, 0, 0]
    _close(txz[2:nz + 2], cs2 * d2 * (g0 + g1 * zw[2:nz + 2]) ** 2, 1e-13, "txz on the w levels")
    tzz = o.get("tzz")[:, 0, 0]
    assert tzz[0] == tzz[1], "tzz(0) = tzz(1)"


# ---------------------------------------------------------------- cst_Smag mixing length
# src/compute_sgs.f90:142-168: 1/l^2 = Cs0^2 / (kappa z)^2 + 1/delta^2
# (Mason & Thomson 1992 with n = 2), z = dz/2 at level 1 and (k-1) dz above;
# Cs^2 = Cs0^2; nu_t = Cs^2 l^2 |S|.

def pin_cst_smag_mixing_length(lib):
    prm, o = _mk(lib, 16, 16, 8, sgs_model=P.SGS_CST_SMAG, cs0=0.16)
    sh = _shape(prm)
    _zero_state(o, prm)
    G = 2.5
    o.put("dudz", np.full(sh, G))  # |S| = G at every node
    o.sgs_stress(1, 1)
    dz, d = prm.lz / prm.nz, _delta(prm)
    z = np.concatenate([[0.5 * dz], (np.arange(2, prm.nz + 2) - 1) * dz])
    cs0 = prm.cs0
    nut = cs0 ** 2 * G / (cs0 ** 2 / (KAPPA * z) ** 2 + 1 / d ** 2)
    got = o.get("nu_t")[1:prm.nz + 2, 0, 0]
    _close(got, nut, 1e-13, "nu_t with the damped mixing length")
    assert np.all(o.get("cs_opt2")[1:prm.nz + 2] == cs0 ** 2)
    # limits: Prandtl's mixing length (kappa z)^2 |S| near the wall, Cs0^2 delta^2 |S| far away
    assert abs(got[0] / ((KAPPA * z[0]) ** 2 * G) - 1) < 0.5
    assert abs(got[-1] / (cs0 ** 2 * d ** 2 * G) - 1) < 0.1


# ---------------------------------------------------------------- convection
# src/compute_convective_padd.f90:184-247, rotational form u x omega with
# w-node averages and the wall row.  For u, v, w functions of z only the
# horizontal transforms are exact (mean mode), and with omega_1 = -dv/dz,
# omega_2 = du/dz, omega_3 = 0:
#   cx(k) = -(W_{k+1} du/dz_{k+1} + W_k du/dz_k) / 2  (k >= 2), -W_2 du/dz_2 / 2 (k = 1)
#   cy(k) = -(W_{k+1} dv/dz_{k+1} + W_k dv/dz_k) / 2  (k >= 2), -W_2 dv/dz_2 / 2 (k = 1)
#   cz(k) = ((U_k + U_{k-1}) du/dz_k + (V_k + V_{k-1}) dv/dz_k) / 2 (k >= 2), 0 (k = 1)

def pin_convection_vertical_structure(lib):
    prm, o = _mk(lib, 16, 16, 8)
    sh = _shape(prm)
    _zero_state(o, prm)
    nz, dz = prm.nz, prm.lz / prm.nz
    rng = np.random.default_rng(4)
    U = rng.uniform(5, 15, nz + 2)
    V = rng.uniform(-2, 2, nz + 2)
    W = rng.uniform(-1, 1, nz + 2)
    W[1] = W[nz + 1] = 0.0
    uz = np.zeros(nz + 2)
    vz = np.zeros(nz + 2)
    uz[1:] = (U[1:] - U[:-1]) / dz
    vz[1:] = (V[1:] - V[:-1]) / dz
    for n, prof in (("u", U), ("v", V), ("w", W), ("dudz", uz), ("dvdz", vz)):
        o.put(n, np.broadcast_to(prof[:, None, None], sh))
    o.op("conv")
    cx, cy, cz = (o.get(n)[:, 0, 0] for n in ("rhsx", "rhsy", "rhsz"))
    k = np.arange(2, nz + 1)
    ex = -np.concatenate([[0.5 * W[2] * uz[2]], 0.5 * (W[k + 1] * uz[k + 1] + W[k] * uz[k])])
    ey = -np.concatenate([[0.5 * W[2] * vz[2]], 0.5 * (W[k + 1] * vz[k + 1] + W[k] * vz[k])])
    ez = np.concatenate([[0.0], 0.5 * ((U[k] + U[k - 1]) * uz[k] + (V[k] + V[k - 1]) * vz[k])])
    _close(cx[1:nz + 1], ex, 1e-12, "x convection")
    _close(cy[1:nz + 1], ey, 1e-12, "y convection")
    _close(cz[1:nz + 1], ez, 1e-12, "z convection")


# ---------------------------------------------------------------- div_stress + forcing
# src/compute_derivatives.f90:464-577 and src/LES_engine.f90:509-572:
# rhs = conv + d/dx tx + d/dy ty + dz tz (+ pfx, pfy, Coriolis); dz is ddz_w
# (forward, uvp output) for x and y, ddz_uv (backward, w output) for z;
# bottom_wk = full divergence of (txz, tyz, tzz) at level 1, top_wk = its
# horizontal part at level nz+1.

def pin_div_stress_closed_form(lib):
    f, ug, vg, U0, V0 = 1.3, 0.7, -0.4, 2.0, 0.5
    prm, o = _mk(lib, 16, 16, 8, coriolis=1, coriol=f, ug=ug, vg=vg, pfx=0.9, pfy=-0.2)
    sh = _shape(prm)
    _zero_state(o, prm)
    o.put("u", np.full(sh, U0))  # uniform: omega = 0, no convection
    o.put("v", np.full(sh, V0))
    nz, dz = prm.nz, prm.lz / prm.nz
    x, y = _xy(prm)
    kk = np.arange(nz + 2, dtype=float)[:, None, None]
    A, B, C = 1 + 0.1 * kk, 0.5 - 0.05 * kk ** 2, 0.3 * kk
    D, E, F = 0.2 + 0.3 * kk ** 2, 1.1 - 0.2 * kk, 0.4 + 0.01 * kk ** 3
    tau = {"txx": A * np.cos(2 * x + y), "txy": B * np.sin(x - 3 * y), "tyy": C * np.cos(3 * y),
           "txz": D * np.sin(2 * x) * np.cos(y), "tyz": E * np.cos(x + 2 * y), "tzz": F * np.sin(x) * np.sin(2 * y)}
    for n, a in tau.items():
        o.put(n, np.broadcast_to(a, sh))
    rng = np.random.default_rng(5)
    r0 = {n: rng.standard_normal(sh) for n in ("rhsx", "rhsy", "rhsz")}
    for n, a in r0.items():
        o.put(n, a)
    o.op("rhs")
    dxtxx = -2 * A * np.sin(2 * x + y)
    dytxy = -3 * B * np.cos(x - 3 * y)
    dxtxy = B * np.cos(x - 3 * y)
    dytyy = -3 * C * np.sin(3 * y)
    dxtxz = 2 * D * np.cos(2 * x) * np.cos(y)
    dytyz = -2 * E * np.sin(x + 2 * y)
    tx, ty, tz = (np.broadcast_to(tau[n], sh) for n in ("txz", "tyz", "tzz"))
    L = slice(1, nz + 1)
    Lp = slice(2, nz + 2)
    Lm = slice(0, nz)
    ex = (dxtxx + dytxy)[L] + (tx[Lp] - tx[L]) / dz + prm.pfx + f * V0 - f * vg
    ey = (dxtxy + dytyy)[L] + (ty[Lp] - ty[L]) / dz + prm.pfy - f * U0 + f * ug
    ez = np.broadcast_to(dxtxz + dytyz, sh)[L] + (tz[L] - tz[Lm]) / dz
    _close(o.get("rhsx")[L], ex, 1e-12, "rhs x")
    _close(o.get("rhsy")[L], ey, 1e-12, "rhs y")
    _close(o.get("rhsz")[L], ez, 1e-12, "rhs z")
    for n, f_ in (("rhsx", "rhsx_f"), ("rhsy", "rhsy_f"), ("rhsz", "rhsz_f")):
        _close(o.get(f_)[L], r0[n][L], 0.0, f"{f_} = previous rhs")
    hb = np.broadcast_to(dxtxz + dytyz, sh)
    _close(o.get2d("bottom_wk"), hb[1] + (tz[1] - tz[0]) / dz, 1e-12, "bottom_wk")
    _close(o.get2d("top_wk"), hb[nz + 1], 1e-12, "top_wk")


# ---------------------------------------------------------------- pressure boundary rows
# src/compute_pressure.f90:152-165,270-286,448-493: with u = v = w = 0 the
# pressure solves the discrete Laplace equation with (p(1) - p(0)) / dz =
# bottom_wk and (p(nz+1) - p(nz)) / dz = top_wk; the mean mode is pinned to
# 0 at the ghost level 0 and its bottom row is replaced, so mean p(k) = k dz
# mean(top_wk).

def pin_pressure_boundary_rows(lib):
    prm, o = _mk(lib, 16, 16, 12)
    _zero_state(o, prm)
    nz, dz = prm.nz, prm.lz / prm.nz
    x, y = _xy(prm)
    B0, B1, T0, T1 = 0.8, 1.5, -0.6, 0.9
    bot = B0 + B1 * np.cos(2 * x[0] + y[0])
    top = T0 + T1 * np.sin(3 * y[0] - x[0])
    o.set2d("bottom_wk", np.broadcast_to(bot, (prm.ny, prm.nx)))
    o.set2d("top_wk", np.broadcast_to(top, (prm.ny, prm.nx)))
    o.op("pressure", 1.5)
    p = o.get("p")[:nz + 1]
    ph = np.fft.rfft2(p) / (prm.nx * prm.ny)
    # mean mode
    _close(ph[:, 0, 0].real, np.arange(nz + 1) * dz * T0, 1e-12, "mean pressure profile")
    # the bottom's mode (kx, ky) = (2, 1) and the top's (1, -3); the top row
    # through the level-nz equation p(nz-1) - (2 + k^2 dz^2) p(nz) + p_top = 0
    bt = np.fft.rfft2(bot) / (prm.nx * prm.ny)
    tt = np.fft.rfft2(top) / (prm.nx * prm.ny)
    amp = max(B1, T1)
    for (j, i), kk2 in (((1, 2), 5.0), ((prm.ny - 3, 1), 10.0)):
        pk = ph[:, j, i]
        assert np.abs(pk).max() > 1e-3 * amp * dz, f"mode {(i, j)} missing"
        assert abs((pk[1] - pk[0]) / dz - bt[j, i]) <= 1e-10 * amp, f"bottom row of mode {(i, j)}"
        lap = pk[:-2] - (2 + kk2 * dz * dz) * pk[1:-1] + pk[2:]
        assert np.abs(lap).max() <= 1e-12 * np.abs(pk).max(), f"Laplace equation of mode {(i, j)}"
        ptop = (2 + kk2 * dz * dz) * pk[nz] - pk[nz - 1]
        assert abs((ptop - pk[nz]) / dz - tt[j, i]) <= 1e-10 * amp, f"top row of mode {(i, j)}"


# ---------------------------------------------------------------- time advance
# src/LES_engine.f90:253-352: Euler on the first step (jt_total = 1), AB2
# (1.5, -0.5) after it, and rhs <- rhs - grad p after the projection, so the
# rhs_f of the next step holds -grad p.  A unidirectional shear flow
# u = U0 + a sin y, v = w = 0 with no SGS stress (cst_Smag, Cs0 = 0): the
# convective term u x omega = grad(u^2/2) is a discrete gradient (the
# products are alias-free), removed by the projection, so away from the wall
# u advances by dt pfx per step, v and w stay 0, and p = u^2/2 + c, with
# c making the ghost level's mean 0 (p(0) = p(1)).  Were rhs_f missing
# -grad p, the second step's pressure would be (1.5 p_2 - 0.5 p_1) / 1.5.

def pin_time_advance(lib):
    prm, o = _mk(lib, 16, 16, 16, sgs_model=P.SGS_CST_SMAG, cs0=0.0, pfx=3.0, dt=2e-3)
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    _, y = _xy(prm)
    U0, a = 8.0, 1.5
    u0 = np.broadcast_to(U0 + a * np.sin(y[0]), (nz, ny, nx))
    o.set_ic(u0, np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx)))
    dt, pf = prm.dt, prm.pfx
    ustart = o.get("u")
    for n in (1, 2):
        o.step(1)
        u = o.get("u")
        _close(u[2:nz + 1], np.broadcast_to(u0 + n * dt * pf, u0.shape)[1:], 1e-13,
               f"u above the wall after step {n}")
        assert np.abs(o.get("v")[1:nz + 1]).max() < 1e-11 and np.abs(o.get("w")[1:nz + 1]).max() < 1e-11
        # pressure of this step from the velocity it started from
        q = 0.5 * ustart[1:nz + 1] ** 2
        pexp = q - q[0].mean()
        p = o.get("p")
        _close(p[1:nz + 1], pexp, 1e-10, f"pressure of step {n}")
        _close(p[0], pexp[0], 1e-10, f"pressure ghost level of step {n}")
        ustart = u


PINS = {f.__name__[4:]: f for f in (
    pin_update_F_fixed_point, pin_update_F_time_scale_exponent, pin_update_F_time_scale_constant,
    pin_update_F_casts_and_clip, pin_dynamic_coeff_closed_form, pin_germano_4delta_closed_form,
    pin_germano_first_node_w, pin_lagrangian_linear_interior, pin_lagrangian_periodic_x,
    pin_lagrangian_first_nodes, pin_lasd_update_composition, pin_tau_staggering, pin_cst_smag_mixing_length,
    pin_convection_vertical_structure, pin_div_stress_closed_form, pin_pressure_boundary_rows, pin_time_advance)}

# planted mistake (les_oracle.cpp ORC_PLANT = n) -> (what it is, the checks that must catch it)
PLANTED = {
    1: ("relaxation time exponent -1/4 instead of -1/8", ["update_F_time_scale_exponent"]),
    2: ("relaxation time 1.0 delta instead of 1.5 delta", ["update_F_time_scale_constant"]),
    3: ("Lagrangian dt without cs_count", ["update_F_time_scale_constant"]),
    4: ("relaxation weights swapped", ["update_F_time_scale_constant", "update_F_time_scale_exponent"]),
    5: ("F_AB not clipped at 1e-24", ["update_F_casts_and_clip"]),
    6: ("beta = c2 / c4", ["dynamic_coeff_closed_form"]),
    7: ("beta clipped at 1/4 instead of 1/8", ["dynamic_coeff_closed_form"]),
    8: ("beta not 1 on the lid", ["dynamic_coeff_closed_form"]),
    9: ("Cs^2 = c2 beta", ["dynamic_coeff_closed_form"]),
    10: ("tf^2 = 4 at 4 delta", ["germano_4delta_closed_form", "germano_first_node_w"]),
    11: ("2-delta filter used at 4 delta", ["germano_4delta_closed_form", "germano_first_node_w"]),
    12: ("M_ij constant delta^2 instead of 2 delta^2", ["germano_4delta_closed_form", "germano_first_node_w"]),
    13: ("w_ft(1) = w(2) / 2", ["germano_first_node_w"]),
    14: ("departure point sign flipped", ["lagrangian_linear_interior"]),
    15: ("u_lag not divided by cs_count", ["lagrangian_linear_interior"]),
    16: ("zp(1) clamped from itself", ["lagrangian_first_nodes"]),
    17: ("zp(2) not doubled towards the wall", ["lagrangian_first_nodes"]),
    18: ("x halo a shifted copy instead of periodic", ["lagrangian_periodic_x"]),
    19: ("first update F_LM = MM", ["lasd_update_composition"]),
    20: ("Lagrangian interpolation after the relaxation", ["lasd_update_composition"]),
    21: ("uvp stresses with nu_t(k) instead of the average", ["tau_staggering"]),
    22: ("level-1 stresses with the average nu_t", ["tau_staggering"]),
    23: ("w-node stresses with an averaged nu_t", ["tau_staggering"]),
    24: ("mixing length at z = k dz", ["cst_smag_mixing_length"]),
    25: ("Cs^2 = Cs0", ["cst_smag_mixing_length"]),
    26: ("wall row of the x convection without 1/2", ["convection_vertical_structure"]),
    27: ("z convection with 2 u_k instead of u_k + u_k-1", ["convection_vertical_structure"]),
    28: ("x convection averaged over w levels k-1, k", ["convection_vertical_structure"]),
    29: ("bottom pressure row without dz", ["pressure_boundary_rows"]),
    30: ("top pressure row from bottom_wk", ["pressure_boundary_rows"]),
    31: ("mean mode's bottom row not replaced", ["pressure_boundary_rows"]),
    32: ("AB2 on the first step", ["time_advance"]),
    33: ("rhs_f without -grad p", ["time_advance"]),
    34: ("AB2 second weight +0.5", ["time_advance"]),
    35: ("wall_inst filtered", ["wall_inst_unfiltered"]),
    36: ("log law at dz instead of dz/2", ["wall_law_log_law"]),
    37: ("bottom_wk without the vertical part", ["div_stress_closed_form"]),
    38: ("ddz_w / ddz_uv swapped in div_stress", ["div_stress_closed_form"]),
}


# ---------------------------------------------------------------- wall law
# src/compute_wall_law.f90:107-254: u* = kappa |u~| / ln(0.5 dz / z0) with u~
# the 2-delta filtered first-level velocity (wall_filter) or the velocity
# itself (wall_inst).

def _wall_case(lib, wall_model):
    prm, o = _mk(lib, 16, 16, 4, wall_model=wall_model)
    sh = _shape(prm)
    x, _ = _xy(prm)
    U0, a = 7.0, 2.0  # cos 5x is above the 2-delta cut-off |k| < 4, cos 3x below
    o.put("u", np.broadcast_to(U0 + a * np.cos(5 * x) + 0.5 * np.cos(3 * x), sh))
    o.put("v", np.zeros(sh))
    o.op("wall_law")
    dz = prm.lz / prm.nz
    return prm, o, x[0], U0, a, np.log(0.5 * dz / prm.z0)


def pin_wall_inst_unfiltered(lib):
    prm, o, x, U0, a, lg = _wall_case(lib, 1)
    u1 = U0 + a * np.cos(5 * x) + 0.5 * np.cos(3 * x)
    _close(o.get2d("ustar"), np.broadcast_to(KAPPA * np.abs(u1) / lg, (prm.ny, prm.nx)), 1e-13, "u* (wall_inst)")
    _close(o.get("txz")[1], np.broadcast_to((KAPPA * u1 / lg) ** 2, (prm.ny, prm.nx)), 1e-13, "txz(1) (wall_inst)")


def pin_wall_law_log_law(lib):
    prm, o, x, U0, a, lg = _wall_case(lib, 0)
    u1 = U0 + 0.5 * np.cos(3 * x)
    _close(o.get2d("ustar"), np.broadcast_to(KAPPA * np.abs(u1) / lg, (prm.ny, prm.nx)), 1e-12, "u* (wall_filter)")
    dz = prm.lz / prm.nz
    _close(o.get2d("lbc_u"), np.broadcast_to(KAPPA * u1 / lg / (0.5 * dz * KAPPA), (prm.ny, prm.nx)), 1e-12,
           "du/dz at the wall")


PINS.update({f.__name__[4:]: f for f in (pin_wall_inst_unfiltered, pin_wall_law_log_law)})


# ---------------------------------------------------------------- Coriolis forcing
# src/LES_engine.f90:545-562: rhs_x += pfx + f v - f vg, rhs_y += pfy - f u + f ug.
# A uniform velocity (no vorticity, no stress) leaves only the forcing in the
# right-hand side of levels 1..nz.

def pin_coriolis_forcing(lib):
    f, ug, vg, pfx, pfy = 2.0 ** -7, 10.0, 3.0, 1.0, 0.5
    prm, o = _mk(lib, 16, 16, 6, coriolis=1, coriol=f, ug=ug, vg=vg, pfx=pfx, pfy=pfy)
    sh = _shape(prm)
    _zero_state(o, prm)
    for n in ("txx", "txy", "txz", "tyy", "tyz", "tzz", "rhsx", "rhsy", "rhsz"):
        o.put(n, np.zeros(sh))
    U, V = 7.0, -2.0
    o.put("u", np.full(sh, U))
    o.put("v", np.full(sh, V))
    o.op("rhs")
    nz = prm.nz
    _close(o.get("rhsx")[1:nz + 1], np.full((nz, prm.ny, prm.nx), pfx + f * V - f * vg), 1e-14, "rhs_x forcing")
    _close(o.get("rhsy")[1:nz + 1], np.full((nz, prm.ny, prm.nx), pfy - f * U + f * ug), 1e-14, "rhs_y forcing")


PINS["coriolis_forcing"] = pin_coriolis_forcing
PLANTED[39] = ("Coriolis term with the opposite sign", ["coriolis_forcing"])
PLANTED[40] = ("geostrophic wind components swapped", ["coriolis_forcing"])


# ---------------------------------------------------------------- TKE budget
# src/tke_budget1D.f90:209-440: plane averages (over x, y) of products of the
# fields on the uvp nodes of levels 1..nz, added to per-level sums.  On
# single Fourier modes the averages are closed forms: <cos> = 0,
# <cos^2> = 1/2, <cos^3> = 0, <cos sin> = 0.

_TKE_FIELDS = ("u", "v", "w", "p", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
               "txx", "txy", "txz", "tyy", "tyz", "tzz", "cs_opt2")


def _tke_case(lib, fields, nx=16, ny=16, nz=6, calls=1):
    prm, o = _mk(lib, nx, ny, nz)
    sh = _shape(prm)
    for n in _TKE_FIELDS:
        o.put(n, np.broadcast_to(fields.get(n, 0.0), sh))
    o.tke_reset()
    for _ in range(calls):
        o.tke_accumulate()
    return prm, o.tke_get()


def _small(got, scale, tol, what):
    """|got| <= tol * scale: a vanishing average, measured on the scale of
    the terms it averages."""
    got = np.asarray(got, dtype=np.float64)
    err = np.abs(got).max() / scale
    assert np.isfinite(got).all() and err <= tol, f"{what}: {err:.3e} of {scale:.3e} > {tol}"


def _levels(prm):
    return np.arange(prm.nz + 2, dtype=np.float64)[:, None, None]


def pin_tke_cleaned_pressure(lib):
    """p - (u^2 + v^2 + W^2) / 2, filtered, and its derivatives: with u, v,
    w uniform and p = P0 + beta k + A cos 3x the mean is P0 - (U^2 + V^2 +
    c^2) / 2 + beta k, <p dudx> = A / 2 for dudx = cos 3x, and dpdz = beta /
    dz between the ghost planes (which are 0: level 1 reads Pf(0) = 0, level
    nz reads Pf(nz+1) = 0)."""
    U, V, c, P0, beta, A = 1.5, -0.5, 0.25, 3.0, 0.125, 0.75
    prm, o0 = _mk(lib, 16, 16, 6)
    x, _ = _xy(prm)
    k = _levels(prm)
    _, t = _tke_case(lib, {"u": U, "v": V, "w": c, "p": P0 + beta * k + A * np.cos(3 * x),
                           "dudx": np.cos(3 * x) + 0 * k})
    nz, dz = prm.nz, prm.lz / prm.nz
    kk = np.arange(1, nz + 1, dtype=np.float64)
    pm = P0 - 0.5 * (U * U + V * V + c * c) + beta * kk
    _close(t["p"], pm, 1e-13, "<p'>")
    _close(t["pu"], U * pm, 1e-13, "<p' u>")
    _close(t["pw"], c * pm, 1e-13, "<p' W>")
    _close(t["pdudx"], np.full(nz, A / 2), 1e-12, "<p' du/dx>")
    _small(t["dpdx"], 3 * A, 1e-13, "<dp'/dx>")
    dpdz = np.full(nz, beta / dz)
    dpdz[0] = 0.5 * (pm[0] + beta) / dz
    dpdz[-1] = 0.5 * (beta - pm[-1]) / dz
    _close(t["dpdz"], dpdz, 1e-12, "<dp'/dz>")


def pin_tke_transport_and_interpolation(lib):
    """u = U + a cos x, w(k) = (w0 + w1 k) cos x on the w nodes, so on the uvp
    node W = b_k cos x with b_k = w0 + w1 (k + 1/2): <W^2> = b^2 / 2,
    <u W> = a b / 2, <u^2 W> = U a b, <u W^2> = U b^2 / 2, <W^3> = 0."""
    U, a, w0, w1 = 2.0, 0.5, 0.3, 0.2
    prm, _ = _mk(lib, 16, 16, 6)
    x, _ = _xy(prm)
    k = _levels(prm)
    _, t = _tke_case(lib, {"u": U + a * np.cos(x) + 0 * k, "w": (w0 + w1 * k) * np.cos(x)})
    b = w0 + w1 * (np.arange(1, prm.nz + 1) + 0.5)
    _close(t["ww"], b * b / 2, 1e-13, "<W W>")
    _close(t["uw"], a * b / 2, 1e-13, "<u W>")
    _close(t["uuw"], U * a * b, 1e-13, "<u u W>")
    _close(t["uww"], U * b * b / 2, 1e-13, "<u W W>")
    _small(t["www"], b.max() ** 3, 1e-14, "<W W W>")
    _close(t["u"], np.full(prm.nz, U), 1e-14, "<u>")


def pin_tke_wall_level_stresses(lib):
    """txz(k) = t0 + t1 k, dudz(k) = g0 + g1 k, dwdx(k) = h0 + h1 k uniform
    in the plane, u = U: the SGS transport u txz and dissipation on the uvp
    node interpolate between k and k+1, except at level 1 (the wall law),
    where txz(1) and dudz(1) + (dwdx(1) + dwdx(2)) / 2 are used."""
    U, t0, t1, g0, g1, h0, h1 = 1.25, 0.5, 0.25, 2.0, -0.5, 0.125, 0.375
    prm, _ = _mk(lib, 16, 16, 6)
    k = _levels(prm)
    _, t = _tke_case(lib, {"u": U + 0 * k, "txz": t0 + t1 * k, "dudz": g0 + g1 * k, "dwdx": h0 + h1 * k})
    kk = np.arange(1, prm.nz + 1, dtype=np.float64)
    Txz = t0 + t1 * (kk + 0.5)
    Txz[0] = t0 + t1
    S13 = (g0 + g1 * (kk + 0.5)) + (h0 + h1 * (kk + 0.5))
    S13[0] = (g0 + g1) + (h0 + h1 * 1.5)
    _close(t["utxz"], U * Txz, 1e-13, "<u txz>")
    _close(t["dxz"], 0.5 * S13 * Txz, 1e-13, "<S13 txz>")
    _close(t["txz"], t0 + t1 * (kk + 0.5), 1e-13, "<txz> (interpolated at every level)")


def pin_tke_dissipation_single_modes(lib):
    """dudx = a cos x, dudy = b sin x, dvdx = c sin x, txy = d sin x:
    <dudx^2> = a^2 / 2, <(dudy + dvdx) / 2 txy> = (b + c) d / 4, and two
    calls add (the window sums)."""
    a, b, c, d = 0.75, 0.5, -0.25, 2.0
    prm, _ = _mk(lib, 16, 16, 6)
    x, _ = _xy(prm)
    z = 0 * _levels(prm)
    _, t = _tke_case(lib, {"dudx": a * np.cos(x) + z, "dudy": b * np.sin(x) + z, "dvdx": c * np.sin(x) + z,
                           "txy": d * np.sin(x) + z}, calls=2)
    nz = prm.nz
    _close(t["dudx2"], np.full(nz, 2 * a * a / 2), 1e-13, "<dudx^2>")
    _close(t["dxy"], np.full(nz, 2 * (b + c) * d / 4), 1e-13, "<S12 txy>")
    _close(t["dudy2"], np.full(nz, 2 * b * b / 2), 1e-13, "<dudy^2>")


PINS.update({f.__name__[4:]: f for f in (pin_tke_cleaned_pressure, pin_tke_transport_and_interpolation,
                                          pin_tke_wall_level_stresses, pin_tke_dissipation_single_modes)})
PLANTED[41] = ("cleaned pressure without the W^2 term", ["tke_cleaned_pressure"])
PLANTED[42] = ("dp/dz left on the w node (no interpolation)", ["tke_cleaned_pressure"])
PLANTED[43] = ("w not interpolated to the uvp node", ["tke_transport_and_interpolation"])
PLANTED[44] = ("u W^2 computed as u^2 W", ["tke_transport_and_interpolation"])
PLANTED[45] = ("wall-level txz interpolated like the other levels", ["tke_wall_level_stresses"])
PLANTED[46] = ("S12 without the factor 1/2", ["tke_dissipation_single_modes"])
PLANTED[47] = ("window sums overwritten instead of added", ["tke_dissipation_single_modes"])
