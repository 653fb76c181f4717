"""The TKE budget sums (src/tke_budget1D.f90:209-440) of the device against
the oracle's on the same state: the 61 plane-averaged quantities after a few
steps, accumulated over several calls.  The per-point values are the same
operations in the same order in the FP64 mode; the plane sums run in another
order (slices and threads on the device, one running sum per level in the
oracle, as compute_planar_average), so each sum is compared within 1e-12 of
the same sum of |value| (the bound of a reordered sum's rounding, with
margin).  The mixed mode's sums follow the FP32 fields: within 1e-4 of that
scale."""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params, load_pkg

pytestmark = pytest.mark.gpu


def _run(grid, prec, steps=6, first=2, seed=7):
    pkg, P, IC, O = load_pkg(), load_params(), load_ic(), load_oracle()
    nx, ny, nz = grid
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, precision=prec)
    ref = P.LesParams(nx=nx, ny=ny, nz=nz)
    u, v, w = IC.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.1, seed=seed)
    o = O.Oracle(ref)
    o.set_ic(u, v, w)
    o.tke_reset()
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        g.tke_init()
        for s in range(1, steps + 1):
            g.step(1)
            o.step(1)
            if s >= first:
                g.tke_accumulate()
                o.tke_accumulate()
        dev = g.tke_download()
        assert not g.check_nan()
    orc, ab = o.tke_get(with_abs=True)
    return dev, orc, ab


def _compare(dev, orc, ab, tol):
    worst = ("", 0.0)
    for n in orc:
        assert np.isfinite(dev[n]).all(), n
        scale = max(ab[n].max(), 1e-300)
        err = np.abs(dev[n] - orc[n]).max() / scale
        if err > worst[1]:
            worst = (n, err)
        assert err <= tol, f"{n}: {err:.3e} of its |value| sum {scale:.3e}"
    print(f"  worst: {worst[0]} {worst[1]:.2e}", flush=True)


@pytest.mark.parametrize("grid", [(32, 32, 16), (256, 256, 8), (24, 32, 13)])
def test_tke_sums_match_oracle(grid):
    print(f"tke budget FP64 {grid}", flush=True)
    P = load_params()
    dev, orc, ab = _run(grid, P.PREC_FP64, steps=6 if grid[0] == 32 else 3, first=2)
    _compare(dev, orc, ab, 1e-12)


def test_tke_sums_mixed_precision():
    print("tke budget mixed 32x32x16", flush=True)
    P = load_params()
    dev, orc, ab = _run((32, 32, 16), P.PREC_MIXED)
    _compare(dev, orc, ab, 1e-4)


def test_tke_reset_and_errors():
    pkg, P = load_pkg(), load_params()
    prm = P.LesParams(nx=16, ny=16, nz=8)
    with pkg.LesGpu(prm) as g:
        with pytest.raises(pkg.LesError):
            g.tke_accumulate()  # before les_tke_init
        g.tke_init()
        g.step(1)
        g.tke_accumulate()
        assert np.abs(g.tke_download()["uu"]).max() > 0
        g.tke_reset()
        assert all(np.all(a == 0) for a in g.tke_download().values())
