"""ctypes view of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It wraps liblesoracle.so (built from les_oracle.cpp by the
Makefile beside it), the FP64 CPU restatement of the reference LES time step.
"""
import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liblesoracle.so"

# field ids: same numbering as the product's les_field enum
FIELDS = ["u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
          "p", "dpdx", "dpdy", "dpdz", "rhsx", "rhsy", "rhsz", "rhsx_f", "rhsy_f", "rhsz_f",
          "txx", "txy", "txz", "tyy", "tyz", "tzz", "cs_opt2", "nu_t", "F_LM", "F_MM", "F_QN", "F_NN",
          "beta", "u_lag", "v_lag", "w_lag", "s11", "s12", "s13", "s22", "s23", "s33"]
FID = {n: i for i, n in enumerate(FIELDS)}

_libs = {}


def lib(path=None):
    """The oracle library (liblesoracle.so beside this file), or, for the
    oracle's own pinning tests, a build of it at `path` (one with a planted
    mistake).  Each library holds its own oracle state."""
    path = str(path or LIB)
    if path not in _libs:
        if not pathlib.Path(path).exists():
            raise RuntimeError(f"oracle library missing: {path} (run make in {HERE})")
        _lib = ctypes.CDLL(path)
        d = ctypes.POINTER(ctypes.c_double)
        i = ctypes.POINTER(ctypes.c_int)
        for name, args in {
            "orc_init": [i, d], "orc_set_ic": [d, d, d], "orc_step": [ctypes.c_int],
            "orc_get": [ctypes.c_int, d], "orc_put": [ctypes.c_int, d], "orc_get_2d": [ctypes.c_int, d],
            "orc_set_2d": [ctypes.c_int, d],
            "orc_fft2_fwd": [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, ctypes.c_double],
            "orc_fft2_inv": [ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d],
            "orc_fft1": [ctypes.c_int, ctypes.c_int, d],
            "orc_pad_plane": [ctypes.c_int, ctypes.c_int, d, d],
            "orc_trunc_plane": [ctypes.c_int, ctypes.c_int, d, d],
            "orc_tridiag": [ctypes.c_int, ctypes.c_int, d, d], "orc_table": [ctypes.c_int, d],
            "orc_op": [ctypes.c_int, ctypes.c_double], "orc_set_jt_total": [ctypes.c_int],
            "orc_set_threads": [ctypes.c_int], "orc_jt_total": [], "orc_jt": [],
            "orc_sgs_stress": [ctypes.c_int, ctypes.c_int],
            "orc_stat_reset": [], "orc_stat_accumulate": [], "orc_stat_get": [d, d, i],
            "orc_stat_profiles": [d], "orc_diagnostics": [d],
            "orc_update_F": [ctypes.c_int, d, d], "orc_dynamic_mum": [d, d, d, d],
            "orc_tke_reset": [], "orc_tke_accumulate": [], "orc_tke_get": [d, d],
            "orc_set_geostrophic": [ctypes.c_double, ctypes.c_double],
        }.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _libs[path] = _lib
    return _libs[path]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fft1(x, inverse=False):
    z = np.ascontiguousarray(x, dtype=np.complex128).copy()
    rc = lib().orc_fft1(len(z), int(inverse), _dp(z.view(np.float64)))
    if rc:
        raise ValueError("unsupported length")
    return z


def fft2_fwd(planes, scale=1.0):
    """planes: (nplanes, ny, nx) float64 -> (nplanes, ny, nx//2+1) complex128."""
    a = np.ascontiguousarray(planes, dtype=np.float64)
    npl, ny, nx = a.shape
    out = np.zeros((npl, ny, nx // 2 + 1), dtype=np.complex128)
    rc = lib().orc_fft2_fwd(nx, ny, npl, _dp(a), _dp(out.view(np.float64)), scale)
    if rc:
        raise ValueError("unsupported size")
    return out


def fft2_inv(spec, nx):
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    npl, ny, lhx = s.shape
    out = np.zeros((npl, ny, nx), dtype=np.float64)
    rc = lib().orc_fft2_inv(nx, ny, npl, _dp(s.view(np.float64)), _dp(out))
    if rc:
        raise ValueError("unsupported size")
    return out


def pad_plane(plane):
    a = np.ascontiguousarray(plane, dtype=np.float64)
    ny, nx = a.shape
    out = np.zeros((3 * ny // 2, 3 * nx // 2))
    lib().orc_pad_plane(nx, ny, _dp(a), _dp(out))
    return out


def trunc_plane(plane):
    a = np.ascontiguousarray(plane, dtype=np.float64)
    ny2, nx2 = a.shape
    nx, ny = 2 * nx2 // 3, 2 * ny2 // 3
    out = np.zeros((ny, nx))
    lib().orc_trunc_plane(nx, ny, _dp(a), _dp(out))
    return out


class Oracle:
    """FP64 CPU restatement of the reference step (one global instance)."""

    def __init__(self, prm, lib_path=None):
        self.prm = prm
        self.L = lib(lib_path)
        ip = np.array(prm.int_vector(), dtype=np.int32)
        rp = np.array(prm.real_vector(), dtype=np.float64)
        rc = self.L.orc_init(ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(rp))
        if rc:
            raise ValueError(f"oracle init failed ({rc})")
        self.shape = (prm.nz + 2, prm.ny, prm.nx)

    @staticmethod
    def threads(n=0):
        return lib().orc_set_threads(n)

    def set_ic(self, u, v, w):
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        self.L.orc_set_ic(_dp(u), _dp(v), _dp(w))

    def step(self, n=1):
        self.L.orc_step(n)

    def get(self, name):
        out = np.zeros(self.shape)
        rc = self.L.orc_get(FID[name], _dp(out))
        if rc:
            raise KeyError(name)
        return out

    def put(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        self.L.orc_put(FID[name], _dp(a))

    def get2d(self, which):
        idx = {"lbc_u": 0, "lbc_v": 1, "ustar": 2, "bottom_wk": 3, "top_wk": 4}[which]
        out = np.zeros(self.shape[1:])
        self.L.orc_get_2d(idx, _dp(out))
        return out

    def set2d(self, which, arr):
        """A surface layer, or "z0": a roughness map (x, y) for the wall law."""
        idx = {"lbc_u": 0, "lbc_v": 1, "ustar": 2, "bottom_wk": 3, "top_wk": 4, "z0": 5}[which]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.L.orc_set_2d(idx, _dp(a))

    def table(self, which):
        idx = {"kkx": 0, "kky": 1, "kk2": 2, "G2": 3, "G4": 4, "GA": 5, "GB": 6}[which]
        out = np.zeros((self.prm.ny, self.prm.nx // 2 + 1))
        if self.L.orc_table(idx, _dp(out)):
            raise KeyError(which)
        return out

    def tridiag(self, i, j, r):
        rr = np.zeros(self.prm.nz + 3, dtype=np.complex128)
        rr[1:self.prm.nz + 3] = r
        x = np.zeros_like(rr)
        self.L.orc_tridiag(i, j, _dp(rr.view(np.float64)), _dp(x.view(np.float64)))
        return x[1:self.prm.nz + 3]

    def op(self, which, arg=0.0):
        code = {"wall_law": 0, "vel_grad": 1, "sgs": 2, "rhs": 3, "pressure": 4, "sij": 5, "conv": 6,
                "lagrangian": 7, "set_bc": 8, "dynamic_coeff": 9}[which]
        self.L.orc_op(code, float(arg))

    def update_F(self, which, AB, BB):
        """update_interpolag_coeff of (F_LM, F_MM) (which 0) or (F_QN, F_NN)
        (which 1) with the fields AB, BB."""
        a = np.ascontiguousarray(np.broadcast_to(AB, self.shape), dtype=np.float64)
        b = np.ascontiguousarray(np.broadcast_to(BB, self.shape), dtype=np.float64)
        self.L.orc_update_F(int(which), _dp(a), _dp(b))

    def dynamic_mum(self):
        """compute_dynamic_mum at 2 delta and 4 delta from the current u, v,
        w and S_ij: (LM, MM, QN, NN) on levels 1..nz+1."""
        out = [np.zeros(self.shape) for _ in range(4)]
        self.L.orc_dynamic_mum(*(_dp(x) for x in out))
        return tuple(out)

    def set_jt_total(self, jt):
        self.L.orc_set_jt_total(jt)

    def jt_total(self):
        return self.L.orc_jt_total()

    def jt(self):
        """Steps taken since initialisation (the reference's local count jt)."""
        return self.L.orc_jt()

    def sgs_stress(self, jt, jtt):
        self.L.orc_sgs_stress(int(jt), int(jtt))

    # ---- running statistics and diagnostics
    STATS = ["u", "v", "w", "p", "u2", "v2", "w2", "u3", "v3", "w3", "u4", "v4", "w4", "uv", "uw", "vw",
             "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nut", "cs"]
    PROFILES = ["u", "v", "w", "p", "var_u", "var_v", "var_w", "txx", "tyy", "tzz", "R13", "txz", "St13",
                "R23", "tyz", "St23", "tke", "cs", "nut", "C13", "C23"]

    def stat_reset(self):
        self.L.orc_stat_reset()

    def stat_accumulate(self):
        self.L.orc_stat_accumulate()

    def stat_get(self):
        nz, ny, nx = self.prm.nz, self.prm.ny, self.prm.nx
        s = np.zeros((len(self.STATS), nz, ny, nx))
        us = np.zeros((ny, nx))
        it = ctypes.c_int()
        self.L.orc_stat_get(_dp(s), _dp(us), ctypes.byref(it))
        return dict(zip(self.STATS, s)), us, it.value

    def stat_profiles(self):
        out = np.zeros((len(self.PROFILES), self.prm.nz))
        if self.L.orc_stat_profiles(_dp(out)):
            raise ValueError("no accumulated statistics")
        return dict(zip(self.PROFILES, out))

    def set_geostrophic(self, ug, vg):
        self.L.orc_set_geostrophic(float(ug), float(vg))

    # ---- TKE budget (src/tke_budget1D.f90): the file names of its records
    TKE = ["u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
           "uu", "vv", "ww", "uv", "uw", "vw", "txx", "tyy", "tzz", "txy", "txz", "tyz", "cs",
           "uuw", "vvw", "www", "uww", "wtzz", "utxz", "vtyz", "dxx", "dyy", "dzz", "dxy", "dxz", "dyz",
           "dudx2", "dudy2", "dudz2", "dvdx2", "dvdy2", "dvdz2", "dwdx2", "dwdy2", "dwdz2",
           "dudxdwdx", "dudydwdy", "dudzdwdz", "p", "dpdx", "dpdy", "dpdz", "pw", "pu",
           "pdudx", "pdvdy", "pdwdz", "pdwdx", "pdudz"]

    def tke_reset(self):
        self.L.orc_tke_reset()

    def tke_accumulate(self):
        self.L.orc_tke_accumulate()

    def tke_get(self, with_abs=False):
        """The sums by name; with_abs: also the same sums of |value| (the
        scale of each sum's rounding)."""
        out = np.zeros((len(self.TKE), self.prm.nz))
        ab = np.zeros_like(out)
        self.L.orc_tke_get(_dp(out), _dp(ab))
        s = dict(zip(self.TKE, out))
        return (s, dict(zip(self.TKE, ab))) if with_abs else s

    def diagnostics(self):
        out = np.zeros(4)
        self.L.orc_diagnostics(_dp(out))
        return {"cfl": out[0], "mke": out[1], "u_star": out[2], "divu": out[3]}
