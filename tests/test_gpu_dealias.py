"""The other dealiasing options of the convective term on the GPU against the
CPU oracle: TRUNC (src/compute_convective_trunc.f90), FSMTHG
(src/compute_convective_Fsmthg.f90) and NODLSG
(src/compute_convective_nodlsg.f90), chosen by les_params.dealias.

FP64: sources, products and the spectral multipliers run the same operations
in the same order as the oracle (the transforms are those of the test filter,
exact against the oracle), so the convective term is compared bit for bit at
the operator level and whole steps to the 1e-11 relative-L2 bar.  The mixed
path is held to the bounds of the padded mixed path."""
import numpy as np
import pytest

from conftest import load_ic, load_oracle, load_params, load_pkg, rel_l2

pytestmark = pytest.mark.gpu

P = load_params()
OPTIONS = ["TRUNC", "FSMTHG", "NODLSG"]


def _setup(nx, ny, nz, name, seed=7, **kw):
    ic = load_ic()
    prm = P.LesParams(nx=nx, ny=ny, nz=nz, dealias=P.DEALIAS[name], **kw)
    prm.validate()
    u, v, w = ic.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.1, seed=seed)
    return prm, u, v, w


@pytest.mark.parametrize("name", OPTIONS)
def test_convective_term_bitwise(name):
    pkg, O = load_pkg(), load_oracle()
    for grid in ((32, 32, 8), (96, 48, 6), (256, 256, 3)):
        prm, u, v, w = _setup(*grid, name, seed=11)
        print(f"{name} convection {grid}", flush=True)
        o = O.Oracle(prm)
        o.set_ic(u, v, w)
        o.step(3)
        with pkg.LesGpu(prm) as g:
            for n in ("u", "v", "w", "dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy"):
                g.upload(n, o.get(n))
            o.op("conv")
            g.compute_dealias_convec()
            nz = prm.nz
            for n in ("rhsx", "rhsy", "rhsz"):
                a, b = g.download(n)[1:nz + 1], o.get(n)[1:nz + 1]
                assert np.array_equal(a, b), (name, grid, n, np.abs(a - b).max(), np.abs(b).max())
            assert np.abs(o.get("rhsx")[1:nz + 1]).max() > 0


@pytest.mark.parametrize("name", OPTIONS)
def test_steps_match_oracle(name):
    pkg, O = load_pkg(), load_oracle()
    prm, u, v, w = _setup(32, 32, 16, name)
    print(f"{name} step parity", flush=True)
    o = O.Oracle(prm)
    o.set_ic(u, v, w)
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        for n in (1, 5):
            o.step(n)
            g.step(n)
            nz = prm.nz
            for f in ("u", "v", "w", "rhsx", "rhsy", "rhsz", "cs_opt2"):
                a = g.download(f)[1:nz + 1]
                b = o.get(f)[1:nz + 1]
                e = rel_l2(a, b)
                print(f"  after {o.jt_total()} steps {f:6s} relL2={e:.3e}", flush=True)
                assert e <= 1e-11, (f, e)
        assert not g.check_nan()


@pytest.mark.parametrize("name", ["TRUNC", "FSMTHG"])
def test_mixed_precision_error_vs_fp64_oracle(name):
    pkg, O = load_pkg(), load_oracle()
    prm, u, v, w = _setup(32, 32, 16, name, seed=5)
    o = O.Oracle(prm)
    o.set_ic(u, v, w)
    prm.precision = P.PREC_MIXED
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        o.step(12)
        g.step(12)
        nz = prm.nz
        errs = {n: rel_l2(g.download(n)[1:nz + 1], o.get(n)[1:nz + 1]) for n in ("u", "v", "w")}
        print(f"{name} mixed vs FP64 oracle after 12 steps:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert errs["u"] < 1e-6 and errs["v"] < 1e-4 and errs["w"] < 1e-4


def test_init_rejects_unknown_option():
    pkg = load_pkg()
    prm = P.LesParams(nx=16, ny=16, nz=8)
    prm.dealias = 4
    with pytest.raises(pkg.LesError):
        pkg.LesGpu(prm)
