"""What the trained surrogate gains over the method without it, measured on
the LES itself: the 256^3 learned step with the committed weights and the
same step with the network's last layer zeroed (the constant coefficients
Cs^2, beta of the training means above the near-wall band, with the same
band, fallback and judge schedule) are both run for 3000 steps from the
FP64 run's initial field, and their time-averaged mean-velocity and
resolved-stress profiles over steps 2401-3000 (the reference's
post-processing of its running averages, reduced on the device) are
compared with the FP64 run's.  The band here is z / delta < 4, where the
surrogate covers 95% of the levels; with the default band (z / delta < 16)
both variants stay near the FP64 noise floor (3000 steps: surrogate 0.25% /
2.4%, constant 0.29% / 3.0%).

Measured (MI355X): surrogate 0.31% / 3.7%, constant 0.88% / 14.5% (largest
deviation of mean u, relative; of R13 + C13, relative to its peak).  Over
10000 steps (window 8001-10000) the figures are 2.2-2.7% / 12-13% against
5.0% / 38%, with two FP64 runs whose initial fields differ by 1e-9
relative apart by 0.57% / 3.0% (the noise floor of the window)."""
import numpy as np
import pytest

from conftest import ROOT, load_ic, load_params, load_pkg

pytestmark = pytest.mark.gpu

STEPS, WINDOW = 3000, 600
N_W4 = 64 * 12 + 64 + 2 * (64 * 64 + 64)  # offset of W4 in the compact weight layout


def _profiles(pkg, prm, u, v, w, weights=None):
    with pkg.LesGpu(prm) as g:
        g.set_ic(u, v, w)
        if weights is not None:
            g.mlp_set_weights(weights)
        left = STEPS - WINDOW
        while left > 0:
            n = min(left, 600)
            g.step(n)
            g.sync()
            left -= n
            print(f"  step {g.jt_total}", flush=True)
        g.stat_init()
        for _ in range(WINDOW):
            g.step(1)
            g.stat_accumulate()
        assert not g.check_nan()
        pr = g.stat_profiles()
    return pr["u"], pr["R13"] + pr["C13"]


def test_surrogate_beats_constant_coefficients_in_the_les():
    pkg, P, ic = load_pkg(), load_params(), load_ic()
    weights = np.load(ROOT / "new-repo_amd" / "weights" / "sgs_mlp.npy").astype(np.float32)
    const = weights.copy()
    const[N_W4:N_W4 + 2 * 64 + 2] = 0.0
    nx = ny = nz = 256
    dev = {}
    ref = None
    for name, prec, wts in (("fp64", P.PREC_FP64, None), ("learned", P.PREC_LEARNED, weights),
                            ("constant", P.PREC_LEARNED, const)):
        prm = P.LesParams(nx=nx, ny=ny, nz=nz, dt=1e-4, precision=prec, judge_every=50, learned_zmin=4.0)
        u, v, w = ic.log_law_ic(nx, ny, nz, prm.lz, prm.z0, 0.10, seed=1234)
        print(f"{name}: {STEPS} steps at 256^3", flush=True)
        U, R = _profiles(pkg, prm, u, v, w, wts)
        if ref is None:
            ref = (U, R)
            continue
        du = np.abs(U - ref[0]) / np.abs(ref[0])
        dr = np.abs(R - ref[1]) / np.max(np.abs(ref[1]))
        dev[name] = (float(du.max()), float(dr.max()))
        print(f"  {name}: mean u {du.max():.3%}, resolved stress {dr.max():.3%}", flush=True)
    (lu, lr), (cu, cr) = dev["learned"], dev["constant"]
    # the matched-error criterion on the mean velocity
    assert lu < 0.02
    # the trained network halves the constant coefficients' deviations (at
    # least); measured: a factor 2.9 on mean u and 3.9 on the stress
    assert lu < 0.5 * cu and lr < 0.5 * cr
