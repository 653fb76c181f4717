"""Pinning the CPU oracle (CPU only, no GPU).

The reference holds no tests, golden vectors or fixtures for this path, and
it cannot be built here (it needs the external 2DECOMP&FFT library), so the
oracle is pinned by what the mathematics fixes: library routines (numpy's
FFT), closed forms on single Fourier modes and linear fields, dense linear
solves, invariants (Parseval, idempotence, divergence-free projection,
partition of unity of the Lagrangian weights) and brute-force convolution on
tiny grids.  Each check is chosen so a plausible mistake (dropped term,
wrong sign, transposed index) fails it.
"""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params, rel_l2

O = load_oracle()
P = load_params()
IC = load_ic()

LENGTHS = [6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768]


# ------------------------------------------------------------------ FFTs
@pytest.mark.parametrize("n", LENGTHS)
def test_fft1_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x)
    assert np.abs(O.fft1(x) - ref).max() <= 1e-14 * np.abs(ref).max() * np.log2(n)
    # backward transform is unnormalised (FFTW convention)
    assert np.abs(O.fft1(x, True) - np.fft.ifft(x) * n).max() <= 1e-14 * np.abs(ref).max() * np.log2(n)


def test_fft1_rejects_other_lengths():
    with pytest.raises(ValueError):
        O.fft1(np.ones(10))


@pytest.mark.parametrize("nx,ny", [(16, 16), (64, 64), (96, 48), (48, 96), (256, 256)])
def test_fft2_matches_numpy_rfft2(nx, ny):
    rng = np.random.default_rng(nx * ny)
    a = rng.standard_normal((3, ny, nx))
    s = O.fft2_fwd(a, 0.5)
    ref = np.fft.rfft2(a * 0.5)
    assert rel_l2(s, ref) < 1e-15
    back = O.fft2_inv(ref, nx)
    assert rel_l2(back, np.fft.irfft2(ref, s=(ny, nx)) * nx * ny) < 1e-15


def test_fft2_single_mode_closed_form():
    nx = ny = 32
    x = np.arange(nx) * 2 * np.pi / nx
    y = np.arange(ny) * 2 * np.pi / ny
    f = np.cos(3 * x)[None, :] * np.cos(5 * y)[:, None]
    s = O.fft2_fwd(f[None], 1.0)[0]
    expect = np.zeros_like(s)
    expect[5, 3] = expect[ny - 5, 3] = nx * ny / 4
    assert np.abs(s - expect).max() < 1e-12


def test_c2r_ignores_imaginary_dc_and_nyquist():
    # FFTW c2r treats the input as Hermitian: Im of the kx = 0 and kx = nx/2
    # columns' self-conjugate entries does not reach the output
    nx = ny = 16
    rng = np.random.default_rng(3)
    a = rng.standard_normal((1, ny, nx))
    s = np.fft.rfft2(a)
    s2 = s.copy()
    s2[0, 0, 0] += 7j
    s2[0, 0, nx // 2] += 3j
    assert np.array_equal(O.fft2_inv(s, nx), O.fft2_inv(s2, nx))


def test_parseval():
    rng = np.random.default_rng(11)
    z = rng.standard_normal(384) + 1j * rng.standard_normal(384)
    Z = O.fft1(z)
    assert abs(np.vdot(Z, Z).real / 384 - np.vdot(z, z).real) < 1e-10 * np.vdot(z, z).real


def _bandlimited(nx, ny, seed):
    rng = np.random.default_rng(seed)
    A = np.fft.rfft2(rng.standard_normal((ny, nx)))
    A[:, nx // 2] = 0
    A[ny // 2, :] = 0
    return np.fft.irfft2(A, s=(ny, nx))


def test_pad_is_exact_trig_interpolation():
    nx, ny = 32, 16
    a = _bandlimited(nx, ny, 1)
    got = O.pad_plane(a) / (nx * ny)  # both unnormalised passes scale by n
    # evaluate the trigonometric polynomial on the 3/2 grid directly
    A = np.fft.fft2(a) / (nx * ny)
    kx = np.fft.fftfreq(nx, 1 / nx)
    ky = np.fft.fftfreq(ny, 1 / ny)
    xf = np.arange(3 * nx // 2) * 2 * np.pi / (3 * nx // 2)
    yf = np.arange(3 * ny // 2) * 2 * np.pi / (3 * ny // 2)
    ex = np.exp(1j * np.outer(kx, xf))
    ey = np.exp(1j * np.outer(ky, yf))
    ref = (ey.T @ A @ ex).real
    assert np.abs(got - ref).max() < 1e-12


def test_padded_product_is_alias_free():
    # 3/2 padding makes the quadratic product exact up to the truncation to
    # |k| < n/2 (src/compute_convective_padd.f90): brute-force convolution
    nx = ny = 16
    a, b = _bandlimited(nx, ny, 2), _bandlimited(nx, ny, 3)
    pa, pb = O.pad_plane(a) / (nx * ny), O.pad_plane(b) / (nx * ny)
    prod = pa * pb / (9 * nx * ny / 4)
    got = O.trunc_plane(prod)
    A, B = np.fft.fft2(a) / (nx * ny), np.fft.fft2(b) / (nx * ny)
    C = np.zeros((ny, nx), complex)
    ks = np.fft.fftfreq(nx, 1 / nx).astype(int)
    for j1 in range(ny):
        for i1 in range(nx):
            if A[j1, i1] == 0:
                continue
            for j2 in range(ny):
                for i2 in range(nx):
                    ky, kx = ks[j1] + ks[j2], ks[i1] + ks[i2]
                    if abs(kx) < nx // 2 and abs(ky) < ny // 2:
                        C[ky % ny, kx % nx] += A[j1, i1] * B[j2, i2]
    ref = np.fft.ifft2(C).real * nx * ny
    assert np.abs(got - ref).max() < 1e-11 * np.abs(ref).max()


# ------------------------------------------------------------------ state-level operators
def _oracle(nx=16, ny=16, nz=8, **kw):
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, **kw)
    return prm, O.Oracle(prm)


def _grid(prm):
    x = np.arange(prm.nx) * prm.lx / prm.nx
    y = np.arange(prm.ny) * prm.ly / prm.ny
    return x[None, None, :], y[None, :, None]


def test_wavenumbers_and_test_filters():
    prm, o = _oracle(32, 32, 4)
    kkx, kky, kk2 = o.table("kkx"), o.table("kky"), o.table("kk2")
    lhx = prm.nx // 2 + 1
    kx = np.arange(lhx, dtype=float)
    kx[-1] = 0
    ky = ((np.arange(prm.ny) + prm.ny // 2) % prm.ny - prm.ny // 2).astype(float)
    assert np.array_equal(kkx[0], kx) and np.all(kkx[prm.ny // 2] == 0)
    assert np.array_equal(kky[:, 0][np.arange(prm.ny) != prm.ny // 2], ky[np.arange(prm.ny) != prm.ny // 2])
    assert np.allclose(kk2, kkx ** 2 + kky ** 2)
    for name, alpha in (("G2", 2.0), ("G4", 4.0)):
        G = o.table(name)
        kc2 = (np.pi / (alpha * np.sqrt(prm.lx / prm.nx * prm.ly / prm.ny))) ** 2
        mask = (kk2 < kc2).astype(float) / (prm.nx * prm.ny)
        mask[:, -1] = 0
        mask[prm.ny // 2, :] = 0
        assert np.array_equal(G, mask)
        # at Lx = 2 pi the cutoff keeps |k| < nx / (2 alpha) in integer units
        assert G[0, int(prm.nx / (2 * alpha)) - 1] > 0 and G[0, int(prm.nx / (2 * alpha))] == 0


def test_gradients_of_trig_polynomial_and_linear_z():
    prm, o = _oracle(32, 16, 8)
    x, y = _grid(prm)
    dz = prm.lz / prm.nz
    z = (np.arange(prm.nz + 2) - 0.5)[:, None, None] * dz
    u = (np.sin(2 * x) + np.cos(3 * y) + 0.3 * np.sin(x + 2 * y)) * (1 + z) + 2.5 * z
    w = np.cos(x) * np.sin(y) * z
    u = u * np.ones((prm.nz + 2, prm.ny, prm.nx))
    w = w * np.ones_like(u)
    o.put("u", u)
    o.put("v", np.zeros_like(u))
    o.put("w", w)
    o.op("vel_grad")
    dudx = (2 * np.cos(2 * x) + 0.3 * np.cos(x + 2 * y)) * (1 + z)
    dudy = (-3 * np.sin(3 * y) + 0.6 * np.cos(x + 2 * y)) * (1 + z)
    assert np.abs(o.get("dudx") - dudx).max() < 1e-12
    assert np.abs(o.get("dudy") - dudy).max() < 1e-12
    # the field has no Nyquist content, so filtering leaves it unchanged
    assert np.abs(o.get("u")[1:-1] - u[1:-1]).max() < 1e-12
    # staggered FD of a field linear in z is exact: dudz at w-node k
    uz = o.get("dudz")
    ref = (np.sin(2 * x) + np.cos(3 * y) + 0.3 * np.sin(x + 2 * y)) + 2.5
    assert np.abs(uz[2:prm.nz + 1] - ref).max() < 1e-11
    dwdz = o.get("dwdz")
    assert np.abs(dwdz[2:prm.nz] - np.cos(x) * np.sin(y)).max() < 1e-11


def test_gradient_filter_removes_nyquist():
    prm, o = _oracle(16, 16, 4)
    x, y = _grid(prm)
    u = (np.cos(8 * x) + np.cos(8 * y) + np.sin(x)) * np.ones((prm.nz + 2, prm.ny, prm.nx))
    o.put("u", u)
    o.op("vel_grad")
    got = o.get("u")
    assert np.abs(got[2:prm.nz] - np.sin(x)).max() < 1e-12


def test_pressure_tridiagonal_matches_dense_solve():
    prm, o = _oracle(16, 16, 10)
    rng = np.random.default_rng(5)
    nz = prm.nz
    dz = prm.lz / nz
    kkx, kky = o.table("kkx"), o.table("kky")
    for (i, j) in [(0, 0), (1, 0), (3, 5), (7, 15), (2, 12)]:
        r = rng.standard_normal(nz + 2) + 1j * rng.standard_normal(nz + 2)
        if (i, j) == (0, 0):
            r[0] = 0
        k2 = kkx[j, i] ** 2 + kky[j, i] ** 2
        A = np.zeros((nz + 2, nz + 2))
        A[0, 0], A[0, 1] = (1.0, 0.0) if (i, j) == (0, 0) else (-1.0, 1.0)
        for k in range(1, nz + 1):
            A[k, k - 1] = A[k, k + 1] = 1 / dz ** 2
            A[k, k] = -(k2 + 2 / dz ** 2)
        A[nz + 1, nz], A[nz + 1, nz + 1] = -1.0, 1.0
        ref = np.linalg.solve(A, r)
        got = o.tridiag(i, j, r)
        assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max()


def _ic_oracle(nx=16, ny=16, nz=8, amp=0.1, **kw):
    prm, o = _oracle(nx, ny, nz, **kw)
    u, v, w = IC.log_law_ic(nx, ny, nz, prm.lz, prm.z0, amp, seed=3)
    o.set_ic(u, v, w)
    return prm, o


def _div(prm, o):
    """Discrete divergence of the post-projection velocity at uvp levels
    2..nz-1 in spectral space (non-Nyquist modes)."""
    nz, dz = prm.nz, prm.lz / prm.nz
    U, V, W = (np.fft.rfft2(o.get(n)) for n in ("u", "v", "w"))
    kkx, kky = o.table("kkx"), o.table("kky")
    d = 1j * kkx * U[2:nz] + 1j * kky * V[2:nz] + (W[3:nz + 1] - W[2:nz]) / dz
    d[:, :, -1] = 0
    d[:, prm.ny // 2, :] = 0
    scale = np.abs(1j * kkx * U[2:nz]).max()
    return np.abs(d).max() / scale


def test_projection_makes_velocity_divergence_free():
    prm, o = _ic_oracle()
    o.step(3)
    assert _div(prm, o) < 1e-12


def test_wall_law_uniform_closed_form():
    prm, o = _oracle(16, 16, 4, wall_model=0)
    U0, V0 = 7.0, -2.0
    u = np.full((prm.nz + 2, prm.ny, prm.nx), U0)
    v = np.full_like(u, V0)
    o.put("u", u)
    o.put("v", v)
    o.op("wall_law")
    dz = prm.lz / prm.nz
    mag = np.hypot(U0, V0)
    us = 0.4 * mag / np.log(0.5 * dz / prm.z0)
    assert np.allclose(o.get2d("ustar"), us, rtol=1e-14)
    assert np.allclose(o.get2d("lbc_u"), us / (0.5 * dz * 0.4) * U0 / mag, rtol=1e-13)
    assert np.allclose(o.get("txz")[1], us * us * U0 / mag, rtol=1e-13)
    assert np.allclose(o.get("tyz")[1], us * us * V0 / mag, rtol=1e-13)


def test_lagrangian_weights_preserve_uniform_F():
    prm, o = _oracle(16, 16, 6)
    rng = np.random.default_rng(9)
    shape = (prm.nz + 2, prm.ny, prm.nx)
    for n in ("u_lag", "v_lag", "w_lag"):
        o.put(n, rng.uniform(-1, 1, shape) * 2000.0)  # |displacement| < 1 cell
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        o.put(n, np.full(shape, 0.37))
    o.op("lagrangian")
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        assert np.abs(o.get(n)[1:prm.nz + 2] - 0.37).max() < 1e-15


def test_lagrangian_zero_velocity_keeps_F():
    prm, o = _oracle(16, 16, 6)
    rng = np.random.default_rng(10)
    shape = (prm.nz + 2, prm.ny, prm.nx)
    F = rng.uniform(0.1, 1.0, shape)
    o.put("F_LM", F)
    o.op("lagrangian")
    assert np.array_equal(o.get("F_LM")[1:prm.nz + 2], F[1:prm.nz + 2])


def test_lagrangian_one_rank_halo_is_shifted_copy():
    # a uniform +y displacement of 0.5 cells at the row j = 1 reads the halo
    # row 0, which a one-rank run fills with row 2 (src/decomp_2d_custom.f90:963-991)
    for mode, src_row in ((0, 2), (1, 16)):
        prm, o = _oracle(16, 16, 6, halo_mode=mode)
        shape = (prm.nz + 2, prm.ny, prm.nx)
        dy = prm.ly / prm.ny
        v = 0.5 * dy / (prm.dt * prm.cs_count)  # yp = -0.5 -> backward neighbour
        o.put("v_lag", np.full(shape, v * prm.cs_count))
        rows = np.arange(1, prm.ny + 1, dtype=float)[None, :, None] * np.ones(shape)
        o.put("F_LM", rows)
        o.op("lagrangian")
        got = o.get("F_LM")[3, 0, 0]
        assert abs(got - 0.5 * (1 + src_row)) < 1e-12, (mode, got)


def test_lasd_update_casts_and_clips():
    prm, o = _ic_oracle(16, 16, 8, cs_count=2)
    o.step(4)  # updates at steps 2 and 4
    cs, beta = o.get("cs_opt2")[1:prm.nz + 2], o.get("beta")[1:prm.nz + 2]
    assert np.array_equal(cs.astype(np.float32).astype(np.float64), cs)
    assert np.array_equal(beta.astype(np.float32).astype(np.float64), beta)
    flm = o.get("F_LM")[1:prm.nz + 2]
    assert np.array_equal(flm.astype(np.float32).astype(np.float64), flm)
    assert beta.min() >= 0.125 and np.all(beta[-1] == 1.0)
    assert cs.min() >= np.float32(1e-24)


def test_first_steps_use_zero_cs_before_first_update():
    prm, o = _ic_oracle(16, 16, 8)
    o.step(1)
    assert np.all(o.get("cs_opt2") == 0) and np.all(o.get("nu_t") == 0)
    o.step(4)
    assert o.get("cs_opt2")[1:prm.nz + 2].min() > 0


def test_spectral_derivatives_carry_no_nyquist():
    # grad_velocity_filt zeroes the kx = nx/2 and ky = ny/2 modes before
    # differentiating; the vertical stress derivative (physical space) may
    # put Nyquist content back into u, as in the reference
    prm, o = _ic_oracle(16, 16, 8)
    o.step(6)
    for n in ("dudx", "dudy", "dvdx", "dwdy"):
        U = np.fft.rfft2(o.get(n)[1:prm.nz + 1])
        ref = np.abs(U[:, :, :-1]).max()
        assert np.abs(U[:, :, -1]).max() < 1e-13 * ref
        assert np.abs(U[:, prm.ny // 2, :]).max() < 1e-13 * ref


def test_step_is_deterministic_and_bounded():
    prm, o = _ic_oracle(16, 16, 8)
    o.step(12)
    a = o.get("u")
    prm2, o2 = _ic_oracle(16, 16, 8)
    o2.step(12)
    assert np.array_equal(a, o2.get("u"))
    assert np.isfinite(a).all() and 0 < a[1:-1].mean() < 40
