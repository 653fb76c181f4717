"""CPU checks of the FP64 MLP reference used to judge the device training
kernels: its backprop against central finite differences of its own loss,
and the bf16 rounding helper against a direct bit-level construction."""
import numpy as np

from conftest import ORACLE_DIR  # noqa: F401  (puts oracle/ on the import path)
import importlib.util

spec = importlib.util.spec_from_file_location("mlp_oracle", ORACLE_DIR / "mlp_oracle.py")
M = importlib.util.module_from_spec(spec)
spec.loader.exec_module(M)


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    v = M.init_params(1)
    X = rng.standard_normal((32, 12))
    T = rng.standard_normal((32, 2))
    loss, g, _ = M.loss_and_grad(v, X, T, "exact")
    idx = rng.choice(v.size, 60, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(v[i]))
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        fd = (M.loss_and_grad(vp, X, T)[0] - M.loss_and_grad(vm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float64)
    got = M.bf16(x)
    assert got[0] == 1.0
    assert got[1] == 1.0  # halfway: ties to even mantissa
    assert got[2] == 1.0 + 2 ** -7
    assert got[3] == -2.5
    assert abs(got[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_gelu_derivative():
    x = np.linspace(-4, 4, 81)
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.abs(fd - M.gelu_d(x)).max() < 1e-8
