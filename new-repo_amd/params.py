"""Run parameters of the LES step and the reference's namelist reader.

The namelist groups and defaults follow the reference's ``read_parameters``
(src/parameters_IO.f90:128-322): the same six groups are read from
``parameters.nlst`` and nondimensionalised the same way (domain lengths x pi
when ``XY_pi``, divided by ``zi``; ``dt`` scaled when ``dim_dt_flag``).  Only
the options the time step supports are accepted; anything else raises, the
way the reference aborts with code 1 or 22.
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import Dict, List

import numpy as np

PI = 3.1415926535897932384626433

SGS_CST_SMAG = 0
SGS_DYN_SMAG = 1
SGS_LASD = 2
SGS_DNS = 3
# convection dealiasing, the reference's compile-time DSLG choice (src/Makefile:17-40)
DEALIAS = {"PADD": 0, "TRUNC": 1, "FSMTHG": 2, "NODLSG": 3}
PREC_FP64 = 0
PREC_MIXED = 1
PREC_LEARNED = 2


@dataclasses.dataclass
class LesParams:
    nx: int = 64
    ny: int = 64
    nz: int = 64
    nt: int = 200
    dt: float = 2.0e-4          # nondimensional
    lx: float = 2.0 * PI        # nondimensional domain lengths
    ly: float = 2.0 * PI
    lz: float = 1.0
    z0: float = 1.0e-4          # lbc_dz0 / zi
    pfx: float = 1.0
    pfy: float = 0.0
    cs0: float = 0.16
    sgs_model: int = SGS_LASD
    cs_count: int = 5
    dyn_init: int = 1
    lagran_init: int = 0
    wall_model: int = 0         # 0 wall_filter, 1 wall_inst
    ubc_u_neu: int = 1          # stress-free lid: u, v Neumann 0, w Dirichlet 0
    ubc_v_neu: int = 1
    ubc_u: float = 0.0
    ubc_v: float = 0.0
    ubc_w: float = 0.0
    halo_mode: int = 0          # 0 one-rank reference halo, 1 periodic
    coriolis: int = 0
    dealias: int = 0            # 0 PADD (3/2 rule), 1 TRUNC (2/3 rule), 2 FSMTHG, 3 NODLSG
    coriol: float = 0.0
    ug: float = 0.0
    vg: float = 0.0
    precision: int = PREC_FP64
    judge_every: int = 50       # learned SGS: plain LASD judge cadence
    learned_zmin: float = 16.0  # learned SGS: plain LASD on levels with z / delta below this
    re: float = 1.0e5           # physical_param Re (the dns stress uses 2/Re)

    # ---- vectors shared with the CPU oracle (orc_init argument order)
    def int_vector(self) -> List[int]:
        return [self.nx, self.ny, self.nz, self.sgs_model, self.cs_count, self.dyn_init, self.lagran_init,
                self.wall_model, self.ubc_u_neu, self.ubc_v_neu, self.halo_mode, self.coriolis, self.dealias]

    def real_vector(self) -> List[float]:
        return [self.lx, self.ly, self.lz, self.dt, self.z0, self.pfx, self.pfy, self.cs0, self.coriol,
                self.ug, self.vg, self.ubc_u, self.ubc_v, self.ubc_w, self.re]

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz

    def validate(self) -> None:
        for n in (self.nx, self.ny):
            m = n
            while m % 2 == 0:
                m //= 2
            while m % 3 == 0:
                m //= 3
            if m != 1 or n % 4:
                raise ValueError(f"horizontal size {n} must be 2^a 3^b and a multiple of 4")
            if 3 * n // 2 > 1024:
                raise ValueError(f"horizontal size {n}: the 3/2-padded transform length is limited to 1024")
        if self.nz < 3:
            raise ValueError("nz must be >= 3")
        if self.sgs_model not in (SGS_CST_SMAG, SGS_DYN_SMAG, SGS_LASD, SGS_DNS):
            raise ValueError("sgs_model must be cst_Smag, dyn_Smag, dyn_LagScaleDep or dns")
        if self.precision == PREC_LEARNED and self.sgs_model != SGS_LASD:
            raise ValueError("the learned SGS surrogate replaces dyn_LagScaleDep only")
        if self.cs_count < 1:
            raise ValueError("cs_count must be >= 1")
        if self.dealias not in DEALIAS.values():
            raise ValueError("dealias must be 0 PADD, 1 TRUNC, 2 FSMTHG or 3 NODLSG")


# ---------------------------------------------------------------- namelist
_TOKEN = re.compile(r"(\w+)\s*=\s*('[^']*'|\"[^\"]*\"|[^\s,/]+)")


def _value(tok: str):
    t = tok.strip()
    if t[0] in "'\"":
        return t[1:-1]
    low = t.lower()
    if low in (".true.", "t", ".t."):
        return True
    if low in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        return float(low.replace("d", "e"))


def read_namelist(text: str) -> Dict[str, Dict[str, object]]:
    """Parse Fortran namelist groups ``&name ... /`` into dicts (keys lower-case)."""
    groups: Dict[str, Dict[str, object]] = {}
    body = re.sub(r"!.*", "", text)
    for m in re.finditer(r"&(\w+)(.*?)/", body, re.S):
        g: Dict[str, object] = {}
        for k, v in _TOKEN.findall(m.group(2)):
            g[k.lower()] = _value(v)
        groups[m.group(1).lower()] = g
    return groups


def params_from_namelist(text: str, precision: int = PREC_FP64) -> LesParams:
    """Build LesParams from a parameters.nlst text, following
    src/parameters_IO.f90:240-298 for the derived and nondimensional values."""
    nl = read_namelist(text)
    sim, phys = nl.get("simulation_param", {}), nl.get("physical_param", {})
    sgs, bc = nl.get("sgs_param", {}), nl.get("momentum_bc", {})
    p = LesParams()
    p.nx, p.ny, p.nz = int(sim["nx"]), int(sim["ny"]), int(sim["nz"])
    p.nt = int(sim.get("nt", 1))
    if int(sim.get("p_on_y", 1)) * int(sim.get("p_on_z", 1)) != 1:
        pass  # the whole domain lives on one device; the pencil grid is ignored
    zi = float(phys.get("zi", 1.0))
    u_scale = float(phys.get("u_scale", 1.0))
    lx, ly, lz = float(phys.get("lx", 2.0)), float(phys.get("ly", 2.0)), float(phys.get("lz", 1.0))
    if phys.get("xy_pi", False):
        lx, ly = lx * PI, ly * PI
    p.lx, p.ly, p.lz = lx / zi, ly / zi, lz / zi
    dt = float(sim.get("dt", 2e-4))
    p.dt = dt * u_scale / zi if sim.get("dim_dt_flag", False) else dt
    p.z0 = float(phys.get("lbc_dz0", 1e-4)) / zi
    p.pfx, p.pfy = float(phys.get("pfx", 0.0)), float(phys.get("pfy", 0.0))
    p.re = float(phys.get("re", p.re))
    if phys.get("coriolis_flag", False):
        p.coriolis = 1
        p.ug = float(phys.get("ug_dim", 0.0)) / u_scale
        p.vg = float(phys.get("vg_dim", 0.0)) / u_scale
        p.coriol = float(np.float32(1.067e-4)) * zi / u_scale  # the reference's float32 literal
    model = str(sgs.get("sgs_model", "cst_Smag"))
    if model == "dyn_LagScaleDep":
        p.sgs_model = SGS_LASD
    elif model == "dyn_Smag":
        p.sgs_model = SGS_DYN_SMAG
    elif model == "dns":
        p.sgs_model = SGS_DNS
    elif model == "cst_Smag":
        p.sgs_model = SGS_CST_SMAG
    else:
        raise ValueError(f"Invalid SGS model {model!r}")
    p.cs0 = float(sgs.get("cs0", 0.16))
    p.cs_count = int(sgs.get("cs_count", 5))
    p.dyn_init = int(sgs.get("dyn_init", 1))
    if str(bc.get("lbc_special", "wall_law")) != "wall_law":
        raise ValueError("only lbc_special='wall_law' is supported")
    wm = str(bc.get("wall_model", "wall_filter"))
    if wm not in ("wall_filter", "wall_inst"):
        raise ValueError(f"Invalid wall law type {wm!r}")
    p.wall_model = 0 if wm == "wall_filter" else 1
    p.precision = precision
    dslg = str(nl.get("les_gpu_param", {}).get("dealias", "PADD")).upper()
    if dslg not in DEALIAS:
        raise ValueError(f"Invalid dealiasing option {dslg!r}")
    p.dealias = DEALIAS[dslg]
    p.validate()
    return p


# ---------------------------------------------------------------- presets
def config(name: str) -> LesParams:
    """The benchmark grids: 'c0' 64^3 FP64, 'c1' 256x256x128 FP64,
    'c2' 256^3 mixed precision, 'c3' 256^3 mixed + learned SGS."""
    p = LesParams()
    if name == "c0":
        p.nx = p.ny = p.nz = 64
        p.nt = 200
    elif name == "c1":
        p.nx = p.ny = 256
        p.nz = 128
        p.nt = 1000
    elif name in ("c2", "c3"):
        p.nx = p.ny = p.nz = 256
        p.nt = 500
        p.precision = PREC_MIXED if name == "c2" else PREC_LEARNED
    elif name == "tiny":
        p.nx = p.ny = 16
        p.nz = 8
        p.nt = 20
    else:
        raise KeyError(name)
    p.validate()
    return p


def ic_amplitude(name: str) -> float:
    return 0.05 if name == "c1" else 0.10


def dz_of(p: LesParams) -> float:
    return p.lz / p.nz


def delta_of(p: LesParams) -> float:
    return (p.lx / p.nx * p.ly / p.ny * p.lz / p.nz) ** (1.0 / 3.0)


def log_law(z: float, z0: float) -> float:
    return (1.0 / 0.4) * math.log(z / z0)
