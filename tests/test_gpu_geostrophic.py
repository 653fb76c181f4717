"""A turning geostrophic wind (les_set_geostrophic, the reference's
compute_geostrophic_correct, src/compute_geostrophic.f90:24-45) reaches the
Coriolis forcing of the next step: the FP64 device, given a new (ug, vg)
before each step, equals the oracle given the same bit for bit, and differs
from a run that keeps the first wind."""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params, load_pkg

pytestmark = pytest.mark.gpu


def test_geostrophic_wind_update_matches_oracle():
    pkg, P, IC, O = load_pkg(), load_params(), load_ic(), load_oracle()
    nx, ny, nz = 32, 32, 16
    f = float(np.float32(1.067e-4)) * 1000.0 / 0.45
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, coriolis=1, coriol=f, ug=10.0 / 0.45, vg=0.0)
    u, v, w = IC.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.1, seed=5)
    winds = [(prm.ug * np.cos(a), prm.ug * np.sin(a)) for a in np.linspace(0.0, 0.5, 6)]
    o = O.Oracle(prm)
    o.set_ic(u, v, w)
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        for wg in winds:
            g.set_geostrophic(*wg)
            o.set_geostrophic(*wg)
            g.step(1)
            o.step(1)
        turned = {n: g.download(n) for n in ("u", "v", "w")}
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        g.step(len(winds))
        fixed = g.download("v")
    for n in ("u", "v", "w"):
        assert np.array_equal(turned[n][1:nz + 1], o.get(n)[1:nz + 1]), n
    assert not np.array_equal(turned["v"][1:nz + 1], fixed[1:nz + 1])
