"""Running statistics, their profiles and the running diagnostics on the GPU.

FP64: the device sums are accumulated in lockstep with the CPU oracle from
the same initial field and must be equal bit for bit (the step itself is
bit-identical); the profile and diagnostic reductions sum in another order,
so they are held to 1e-12 relative plus 1e-14 absolute (velocities are
O(1-10) in u* units; a near-zero mean such as w carries summation-order noise
of that size), the CFL, a maximum, exactly.  Mixed
precision: the device sums must equal an accumulation of the downloaded
fields (widened to float64) bit for bit."""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params, load_pkg

pytestmark = pytest.mark.gpu


def _setup(nx, ny, nz, prec, seed=3):
    P, ic = load_params(), load_ic()
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, precision=prec)
    prm.validate()
    return prm, ic.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.1, seed=seed)


def _close(a, b, rtol, atol=1e-14):
    return np.abs(a - b).max() <= rtol * np.abs(b).max() + atol


@pytest.mark.parametrize("grid", [(32, 32, 16), (96, 48, 6)])
def test_fp64_statistics_and_diagnostics_match_oracle(grid):
    pkg, O, P = load_pkg(), load_oracle(), load_params()
    nx, ny, nz = grid
    prm, (u, v, w) = _setup(nx, ny, nz, P.PREC_FP64)
    print(f"statistics fp64 {grid}", flush=True)
    o = O.Oracle(prm)
    o.set_ic(u, v, w)
    o.step(2)
    o.stat_reset()
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        g.step(2)
        with pytest.raises(pkg.LesError):
            g.stat_accumulate()  # before les_stat_init
        g.stat_init()
        with pytest.raises(pkg.LesError):
            g.stat_profiles()    # nothing accumulated yet
        for s in range(8):       # steps 3..10: two LASD updates inside the window
            o.step(1)
            g.step(1)
            o.stat_accumulate()
            g.stat_accumulate()
            dg, do = g.running_diagnostics(), o.diagnostics()
            assert dg["cfl"] == do["cfl"], (dg, do)
            for key in ("mke", "u_star", "divu"):
                assert dg[key] == pytest.approx(do[key], rel=1e-12, abs=1e-300), (key, dg, do)
        sg, usg, itg = g.stat_download()
        so, uso, ito = o.stat_get()
        assert itg == ito == 8
        assert np.array_equal(usg, uso)
        for name in so:
            assert np.array_equal(sg[name], so[name]), name
        pg, po = g.stat_profiles(), o.stat_profiles()
        for name in po:
            print(f"  {name:6s} max|dev| {np.abs(pg[name] - po[name]).max():.2e}", flush=True)
            assert _close(pg[name], po[name], 1e-12), name
        # the acceptance quantities are non-trivial here
        assert np.abs(po["R13"]).max() > 0 and np.abs(po["txz"]).max() > 0
        g.stat_reset()
        sg, usg, itg = g.stat_download()
        assert itg == 0 and not np.any(sg["u"]) and not np.any(usg)


def test_mixed_statistics_equal_accumulated_downloads():
    pkg, P = load_pkg(), load_params()
    prm, (u, v, w) = _setup(32, 32, 16, P.PREC_MIXED, seed=9)
    nz = prm.nz
    print("statistics mixed 32x32x16", flush=True)
    names = {"u": "u", "v": "v", "w": "w", "p": "p", "txx": "txx", "tyy": "tyy", "tzz": "tzz", "txy": "txy",
             "txz": "txz", "tyz": "tyz", "dudz": "dudz", "dvdz": "dvdz", "nut": "nu_t"}
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        g.stat_init()
        acc = {k: 0.0 for k in pkg.STATS}
        us_acc = 0.0
        I = slice(1, nz + 1)
        for s in range(6):
            g.step(1)
            g.stat_accumulate()
            f = {k: g.download(src) for k, src in names.items()}
            cs = g.download("cs_opt2")
            uu, vv, ww = f["u"][I], f["v"][I], f["w"][I]
            for k in ("u", "v", "w", "p", "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nut"):
                acc[k] = acc[k] + f[k][I]
            for q, x in (("u", uu), ("v", vv), ("w", ww)):
                acc[q + "2"] = acc[q + "2"] + x * x
                acc[q + "3"] = acc[q + "3"] + (x * x) * x
                acc[q + "4"] = acc[q + "4"] + (x * x) * (x * x)
            acc["uv"] = acc["uv"] + uu * vv
            for key, a in (("uw", f["u"]), ("vw", f["v"])):
                new = np.zeros_like(uu)
                prev = acc[key] if isinstance(acc[key], np.ndarray) else np.zeros_like(uu)
                new[1:] = prev[1:] + f["w"][2:nz + 1] * (0.5 * (a[1:nz] + a[2:nz + 1]))
                acc[key] = new
            acc["cs"] = acc["cs"] + np.sqrt(cs[I])
            us_acc = us_acc + g.layer("ustar")
        sg, usg, itg = g.stat_download()
        assert itg == 6 and np.array_equal(usg, us_acc)
        for k in pkg.STATS:
            assert np.array_equal(sg[k], acc[k]), k
        pr = g.stat_profiles()
        np.testing.assert_allclose(pr["u"], (acc["u"] / 6).mean(axis=(1, 2)), rtol=1e-12)
        d = g.running_diagnostics()
        uu = g.download("u")[I]
        assert d["cfl"] > 0 and d["divu"] >= 0
        assert d["mke"] == pytest.approx(0.5 * (uu.mean() ** 2 + g.download("v")[I].mean() ** 2
                                                + (0.5 * (g.download("w")[1:nz + 1] + g.download("w")[2:nz + 2])).mean() ** 2),
                                         rel=1e-10)
