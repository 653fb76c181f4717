"""Host-side logic of the measurement tools, on the CPU: the benchmark's
refusal of more than one GPU, the constant-coefficient variant of the
surrogate's weights, the kernel-name keys that pair the library's
per-kernel HIP-event statistics with rocprofv3's kernel names, and the
counter summaries' selection of the HIP-event cycle's dispatches."""
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


def _counter_csv(path, rows):
    path.mkdir(parents=True)
    with open(path / "1_counter_collection.csv", "w") as f:
        f.write("Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n")
        for did, name, counter, value in rows:
            f.write(f'{did},"{name}",{counter},{value}\n')


def test_stall_summary_ratios_over_the_last_cycle(tmp_path):
    k = "void k_sgs<float>(SgsArgs<float>)"
    # dispatch 1 belongs to the warm-up (not in the cycle), 2 and 3 to the cycle
    wait = []
    for did, wc in ((1, 1000.0), (2, 100.0), (3, 300.0)):
        wait += [(did, k, "SQ_WAVE_CYCLES", wc), (did, k, "SQ_WAIT_ANY", 0.5 * wc),
                 (did, k, "SQ_WAIT_INST_ANY", 0.25 * wc), (did, k, "SQ_WAIT_INST_LDS", 0.0)]
    lds = [(did, k, "SQ_LDS_IDX_ACTIVE", 30.0) for did in (1, 2, 3)] + \
          [(did, k, "SQ_LDS_BANK_CONFLICT", 10.0) for did in (1, 2, 3)]
    _counter_csv(tmp_path / "a", wait)
    _counter_csv(tmp_path / "b", lds)
    bj = tmp_path / "bench.json"
    bj.write_text('{"kernels": {"k_sgs<float>": {"calls": 2, "ms": 0.6}}}\n')
    args = ps.argparse.Namespace(dir_wait=str(tmp_path / "a"), dir_lds=str(tmp_path / "b"), bench_json=str(bj),
                                 title="t", command=["c"], out=None)
    row = [ln for ln in ps.stalls(args).splitlines() if ln.startswith("| `k_sgs<float>`")]
    assert row == ["| `k_sgs<float>` | 0.50 | 0.25 | 0.000 | 0.50 |"]
