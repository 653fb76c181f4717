"""The restatement of the reference's initial conditions (oracle/ref_ic.py),
the checker of the Fortran host's set_initial_conditions, on the CPU: its
generator equals the reference's own ran3 (src/ran3.f90, compiled from that
file into oracle/_ref/ where the reference is present) draw for draw,
through the restarts the noise makes per level; the log law and Poiseuille
profiles without noise are the closed forms; the noise stays inside its
envelope 3 vnf u*^2 (lz - 0.9 z) / lz and only below z_turb lz."""
import ctypes
import importlib.util
import math
import pathlib

import numpy as np
import pytest

from conftest import ROOT, load_params

_spec = importlib.util.spec_from_file_location("ref_ic", ROOT / "oracle" / "ref_ic.py")
ref_ic = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(ref_ic)
REF_RAN3 = ROOT / "oracle" / "_ref" / "libran3.so"


def test_generator_equals_the_reference_ran3():
    if not REF_RAN3.exists():
        pytest.skip("oracle/_ref/libran3.so not built (the reference is absent)")
    lib = ctypes.CDLL(str(REF_RAN3))
    lib.ran3_.restype = ctypes.c_double
    lib.ran3_.argtypes = [ctypes.POINTER(ctypes.c_int)]
    gen = ref_ic.Ran3()
    for seed0 in (-85, -1, -161803397, -12345, -999):
        s_ref = ctypes.c_int(seed0)
        s = seed0
        for n in range(3000):
            r_ref = lib.ran3_(ctypes.byref(s_ref))
            r, s = gen(s)
            assert r == r_ref and s == s_ref.value, (seed0, n)
            assert 0.0 <= r < 1.0


def test_profiles_without_noise_are_closed_forms():
    P = load_params()
    prm = P.LesParams(nx=8, ny=8, nz=16, pfx=3.0, pfy=4.0)
    z = (np.arange(1, 17) - 0.5) * prm.lz / 16
    u, v, w = ref_ic.reference_ic(prm, "log_law", u_star=0.0)
    law = np.log(z / prm.z0) / 0.4
    assert np.allclose(u[:, 0, 0], 0.6 * law * 0.0, atol=0)  # u* = 0: no flow, no noise
    u, v, w = ref_ic.reference_ic(prm, "log_law", u_star=2.0, z_turb=0.0)
    assert np.allclose(u, (0.6 * law * 2.0)[:, None, None], rtol=1e-15)
    assert np.allclose(v, (0.8 * law * 2.0)[:, None, None], rtol=1e-15)
    assert np.all(w == 0)
    u, _, _ = ref_ic.reference_ic(prm, "poiseuille", u_star=0.0)
    assert np.allclose(u, (prm.re * z * (1 - 0.5 * z))[:, None, None], rtol=1e-15)


def test_noise_envelope_and_depth():
    P = load_params()
    prm = P.LesParams(nx=16, ny=8, nz=16)
    u_star, vnf, zt = 0.5, 2.0, 0.4
    u, v, w = ref_ic.reference_ic(prm, "constant", u_star=u_star, z_turb=zt, vnf=vnf, u0=1.0)
    z = (np.arange(1, 17) - 0.5) * prm.lz / 16
    env = 3 * vnf * u_star ** 2 * (prm.lz - 0.9 * z) / prm.lz
    below = z <= zt * prm.lz
    for x, base in ((u, 1.0), (v, 0.0), (w, 0.0)):
        d = x - base
        assert np.all(np.abs(d[below]) <= env[below][:, None, None])
        assert np.all(d[~below] == 0)
        # uniform on (-1, 1) times the envelope: its spread reaches most of it
        assert np.abs(d[below] / env[below][:, None, None]).max() > 0.9
    # u, v and w draw from different seeds
    assert not np.array_equal(u - 1.0, v)


def test_fortran_host_generator_equals_the_restatement(tmp_path):
    """The Fortran host's generator (new-repo_amd/fortran/initial_conditions.f90,
    gen_start / gen_next), built with flang and run on the CPU, draws the
    restatement's values for the seeds the noise uses."""
    import shutil
    import subprocess
    flang = next((c for c in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang")) if c and pathlib.Path(c).exists()),
                 None)
    lib = ROOT / "new-repo_amd" / "liblesgpu.so"
    if flang is None or not lib.exists():
        pytest.skip("flang or liblesgpu.so missing")
    fsrc = ROOT / "new-repo_amd" / "fortran"
    prog = tmp_path / "gen.f90"
    prog.write_text("""program gen
  use initial_conditions
  implicit none
  type(sub_gen) :: g
  integer :: s, n
  do s = 84, 100, 4
     call gen_start(g, -s)
     do n = 1, 500
        write(*, '(ES25.17)') gen_next(g)
     end do
  end do
end program gen
""")
    srcs = [str(fsrc / f) for f in ("les_gpu_abi.f90", "parameters_IO.f90", "initial_conditions.f90")] + [str(prog)]
    exe = tmp_path / "gen"
    r = subprocess.run([flang, "-O2", "-J", str(tmp_path), *srcs, "-L", str(lib.parent), "-llesgpu",
                        f"-Wl,-rpath,{lib.parent}", "-o", str(exe)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr[-2000:]
    got = [float(x) for x in out.stdout.split()]
    want = []
    for s in range(84, 101, 4):
        gen = ref_ic.Ran3()
        seed = -s
        for _ in range(500):
            x, seed = gen(seed)
            want.append(x)
    assert got == want
