"""Host-side logic of the measurement tools, on the CPU: the benchmark's
refusal of more than one GPU, the constant-coefficient variant of the
surrogate's weights, and the kernel-name keys that pair the library's
per-kernel HIP-event statistics with rocprofv3's kernel names."""
import importlib.util

import numpy as np

from conftest import PKG_DIR, ROOT

_spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
bench = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bench)
_spec2 = importlib.util.spec_from_file_location("profile_summary", PKG_DIR / "profile_summary.py")
ps = importlib.util.module_from_spec(_spec2)
_spec2.loader.exec_module(ps)


def test_bench_refuses_more_than_one_gpu(capsys):
    assert bench.main(["--gpus", "2"]) == 2
    assert "one GPU" in capsys.readouterr().err


def test_constant_weights_zero_the_output_layer_only():
    # compact layout: W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
    # W4[2][64] b4[2], then the normalisers
    sizes = [64 * 12, 64, 64 * 64, 64, 64 * 64, 64, 2 * 64, 2, 12, 12, 2, 2, 1]
    w = np.arange(1, sum(sizes) + 1, dtype=np.float32)
    c = bench.constant_weights(w)
    w4 = sum(sizes[:6])
    assert np.all(c[w4:w4 + 2 * 64 + 2] == 0.0)
    assert np.array_equal(c[:w4], w[:w4]) and np.array_equal(c[w4 + 130:], w[w4 + 130:])
    assert np.all(w != 0.0)  # the input is left alone


def test_kernel_keys_pair_library_and_rocprof_names():
    cases = {
        "void k_cols_c16<float, 1, 2, 2>(ColArgs<float>, Cplx<float> const*)": "k_cols_c16<float>",
        "k_rows_c2r_proj(RowC2RArgs<double, double>, ProjEpi3, FftPlan, Cplx<double> const*)": "k_rows_c2r_proj",
        "void k_rows_r2c<double, float>(RowR2CArgs<double, float>, FftPlan, Cplx<float> const*)":
            "k_rows_r2c<double, float>",
        "void k_nyq_apply_ddz<float, 8, true>(NyqArgs, BcDdzArgs<float>)": "k_nyq_apply_ddz<float>",
        "void k_rows_resample<float, float, double>(RowResArgs<float, float, double>, FftPlan, FftPlan)":
            "k_rows_resample<float, float>",
        "k_tridiag(TriArgs)": "k_tridiag",
    }
    for name, key in cases.items():
        assert ps.rocprof_key(name) == key, name
    # the library's names carry a variant tag that the key drops
    assert bench.kernel_key("k_cols_c16<float> [GRAD]") == "k_cols_c16<float>"
    assert bench.kernel_key("k_tridiag") == "k_tridiag"


def test_counter_summaries_keep_the_last_cycle():
    vals = list(range(10))
    assert ps._last(vals, 3) == [7, 8, 9]
    assert ps._last(vals, 0) == vals
