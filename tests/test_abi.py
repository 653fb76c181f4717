"""C-ABI and host-logic checks that need no GPU: the library loads and
exports every entry point declared in include/les_gpu.h, the ctypes view of
les_params matches the C layout, the namelist reader follows the
reference's nondimensionalisation, and the seeded initial field is
reproducible and has the requested statistics."""
import ctypes
import pathlib
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, load_ic, load_params, load_pkg

HEADER = ROOT / "include" / "les_gpu.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(les_\w+)\s*\(", text, re.M)))


def test_header_declares_core_entry_points():
    names = header_functions()
    for n in ("les_init", "les_finalize", "les_step", "les_grad_velocity_filt", "les_press_stag_array",
              "les_sgs_stress", "les_compute_dealias_convec", "les_wall_law", "les_fft2d_r2c"):
        assert n in names


def test_library_exports_every_header_symbol():
    pkg = load_pkg()
    lib = pkg.library()  # loads without a GPU; no compute call is made
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(pkg.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (les_\w+)", out))
    assert set(header_functions()) <= exported


def test_params_struct_layout_matches_c():
    pkg = load_pkg()
    src = f"""
#include <stdio.h>
#include <stddef.h>
#include "{HEADER}"
int main(void) {{
  printf("%zu %zu %zu %zu %zu\\n", sizeof(les_params), offsetof(les_params, judge_every),
         offsetof(les_params, dealias), offsetof(les_params, lx), offsetof(les_params, ubc_w));
  return 0;
}}
"""
    with tempfile.TemporaryDirectory() as d:
        c = pathlib.Path(d) / "t.c"
        c.write_text(src)
        exe = pathlib.Path(d) / "t"
        subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
        vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    P = pkg._Params
    assert vals == [ctypes.sizeof(P), P.judge_every.offset, P.dealias.offset, P.lx.offset, P.ubc_w.offset]


def test_field_ids_match_header_enum():
    pkg = load_pkg()
    text = HEADER.read_text()
    body = re.search(r"enum les_field \{(.*?)\};", text, re.S).group(1)
    names = [n.strip().split("=")[0].strip() for n in body.replace("\n", " ").split(",") if n.strip()]
    names = [n for n in names if n != "LES_NFIELDS"]
    assert len(names) == len(pkg.FIELDS)
    for cname, pname in zip(names, pkg.FIELDS):
        assert cname == "LES_" + pname.upper(), (cname, pname)


NAMELIST = """
&simulation_param  p_on_y=1 p_on_z=1 nx=64 ny=64 nz=64 nt=200 dt=2.0d-4 dim_dt_flag=.false. restart_flag=.false. /
&output_param      out_path='./' output=.true. step_out=200 stat_flag=.false. start_stats=0 running_period=100
                   tke_flag=.false. start_tke=0 tke_period=100 checkpoint=200 scheckpoint=100000 /
&physical_param    XY_pi=.true. lx=2.0d0 ly=2.0d0 lz=1.0d0 u_scale=1.0d0 zi=1.0d0 lbc_dz0=1.0d-4 Re=1.0d5
                   pfx=1.0d0 pfy=0.0d0 coriolis_flag=.false. ug_dim=0.0d0 vg_dim=0.0d0 geostr_rot=.false. geostr_start=100000 /
&sgs_param         Cs0=0.16d0 sgs_model='dyn_LagScaleDep' cs_count=5 dyn_init=1 /
&momentum_bc       lbc_model='std' lbc_special='wall_law' ubc_model='std' ubc_special='std' wall_model='wall_filter' surface_files=.false. /
&momentum_ic       ic_type='files' u0=0.0d0 v0=0.0d0 w0=0.0d0 u_star=1.0d0 dudz0=0.0d0 z_turb=0.5d0 vnf=1.0d0 /
"""


def test_namelist_reader_nondimensionalises_like_the_reference():
    P = load_params()
    p = P.params_from_namelist(NAMELIST)
    assert (p.nx, p.ny, p.nz, p.cs_count, p.dyn_init) == (64, 64, 64, 5, 1)
    assert p.lx == 2.0 * P.PI and p.ly == 2.0 * P.PI and p.lz == 1.0
    assert p.dt == 2e-4 and p.z0 == 1e-4 and p.pfx == 1.0 and p.sgs_model == P.SGS_LASD
    # dimensional dt and zi scaling (src/parameters_IO.f90:246-272)
    q = P.params_from_namelist(NAMELIST.replace("dim_dt_flag=.false.", "dim_dt_flag=.true.")
                               .replace("zi=1.0d0", "zi=1000.0d0").replace("u_scale=1.0d0", "u_scale=0.5d0"))
    assert abs(q.dt - 2e-4 * 0.5 / 1000.0) < 1e-20 and abs(q.lz - 1e-3) < 1e-18
    with pytest.raises(ValueError):
        P.params_from_namelist(NAMELIST.replace("dyn_LagScaleDep", "dyn_Foo"))


def test_seeded_initial_field():
    ic = load_ic()
    u, v, w = ic.log_law_ic(32, 32, 16, 1.0, 1e-4, 0.1, seed=5)
    u2, _, _ = ic.log_law_ic(32, 32, 16, 1.0, 1e-4, 0.1, seed=5)
    assert np.array_equal(u, u2) and not v.any() and not w.any()
    dz = 1.0 / 16
    z = (np.arange(1, 17) - 0.5) * dz
    mean = np.log(z / 1e-4) / 0.4
    ratio = u / mean[:, None, None]
    low = z <= 0.5
    assert np.abs(ratio[~low] - 1).max() < 1e-15
    assert 0.99 < ratio[low].max() - 1 <= 0.1 + 1e-12 or ratio[low].max() - 1 <= 0.1
    assert abs((ratio[low] - 1).mean()) < 0.01 and (ratio[low] - 1).min() >= -0.1
    with tempfile.TemporaryDirectory() as d:
        ic.write_vel_files(d, u, v, w)
        r = ic.read_vel_files(d, 32, 32, 16)
        assert np.array_equal(r[0], u)
        assert (pathlib.Path(d) / "vel_u.dat").stat().st_size == 32 * 32 * 16 * 8


def test_splitmix_known_values():
    ic = load_ic()
    # SplitMix64 reference outputs for counters 0 and 1 (Steele et al. 2014)
    out = ic.splitmix64(np.array([0, 1], dtype=np.uint64))
    assert int(out[0]) == 0xE220A8397B1DCDAF and int(out[1]) == 0x910A2DEC89025CC1
