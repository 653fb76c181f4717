"""les_stat_profiles_noalloc: the allocation-free entry point gives the
profiles of les_stat_profiles bit for bit (same kernels and summation order,
partial sums in the library's fixed device array), on two calls in a row, the
second on a smaller grid, whose call leaves the first call's values in the
array beyond its own levels; and it refuses a grid beyond its limit."""
import numpy as np
import pytest

from conftest import load_ic, load_params, load_pkg

pytestmark = pytest.mark.gpu


def _profiles_both(nx, ny, nz, precision, steps):
    pkg, P, ic = load_pkg(), load_params(), load_ic()
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, precision=precision)
    prm.validate()
    u, v, w = ic.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.1, seed=nz)
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        g.stat_init()
        for _ in range(steps):
            g.step(1)
            g.stat_accumulate()
        na = g.stat_profiles_noalloc()
        ref = g.stat_profiles()
    return na, ref


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_profiles_equal_the_allocating_entry_point(precision):
    for grid, steps in (((32, 32, 16), 6), ((16, 16, 8), 3)):
        print(f"noalloc profiles {grid}, precision {precision}", flush=True)
        na, ref = _profiles_both(*grid, precision, steps)
        for name in ref:
            assert np.array_equal(na[name], ref[name]), (grid, name)
        assert np.all(np.isfinite(ref["u"])) and ref["u"].max() > 0


def test_noalloc_profiles_refuse_nz_beyond_the_limit():
    pkg, P = load_pkg(), load_params()
    prm = P.LesParams(nx=8, ny=8, nz=1025)
    with pkg.LesGpu(prm) as g:
        g.stat_init()
        with pytest.raises(pkg.LesError) as e:
            g.stat_profiles_noalloc()
        assert e.value.code == 10 and "1024" in str(e.value)
