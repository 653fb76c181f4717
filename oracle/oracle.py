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

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"oracle library missing: {LIB} (run make in {HERE})")
        _lib = ctypes.CDLL(str(LIB))
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
        }.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
    return _lib


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

    def __init__(self, prm):
        self.prm = prm
        ip = np.array(prm.int_vector(), dtype=np.int32)
        rp = np.array(prm.real_vector(), dtype=np.float64)
        rc = lib().orc_init(ip.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(rp))
        if rc:
            raise ValueError(f"oracle init failed ({rc})")
        self.shape = (prm.nz + 2, prm.ny, prm.nx)

    @staticmethod
    def threads(n=0):
        return lib().orc_set_threads(n)

    def set_ic(self, u, v, w):
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        lib().orc_set_ic(_dp(u), _dp(v), _dp(w))

    def step(self, n=1):
        lib().orc_step(n)

    def get(self, name):
        out = np.zeros(self.shape)
        rc = lib().orc_get(FID[name], _dp(out))
        if rc:
            raise KeyError(name)
        return out

    def put(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        lib().orc_put(FID[name], _dp(a))

    def get2d(self, which):
        idx = {"lbc_u": 0, "lbc_v": 1, "ustar": 2, "bottom_wk": 3, "top_wk": 4}[which]
        out = np.zeros(self.shape[1:])
        lib().orc_get_2d(idx, _dp(out))
        return out

    def set2d(self, which, arr):
        idx = {"lbc_u": 0, "lbc_v": 1, "ustar": 2, "bottom_wk": 3, "top_wk": 4}[which]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().orc_set_2d(idx, _dp(a))

    def table(self, which):
        idx = {"kkx": 0, "kky": 1, "kk2": 2, "G2": 3, "G4": 4, "GA": 5, "GB": 6}[which]
        out = np.zeros((self.prm.ny, self.prm.nx // 2 + 1))
        if lib().orc_table(idx, _dp(out)):
            raise KeyError(which)
        return out

    def tridiag(self, i, j, r):
        rr = np.zeros(self.prm.nz + 3, dtype=np.complex128)
        rr[1:self.prm.nz + 3] = r
        x = np.zeros_like(rr)
        lib().orc_tridiag(i, j, _dp(rr.view(np.float64)), _dp(x.view(np.float64)))
        return x[1:self.prm.nz + 3]

    def op(self, which, arg=0.0):
        code = {"wall_law": 0, "vel_grad": 1, "sgs": 2, "rhs": 3, "pressure": 4, "sij": 5, "conv": 6,
                "lagrangian": 7, "set_bc": 8}[which]
        lib().orc_op(code, float(arg))

    def set_jt_total(self, jt):
        lib().orc_set_jt_total(jt)

    def jt_total(self):
        return lib().orc_jt_total()

    def jt(self):
        """Steps taken since initialisation (the reference's local count jt)."""
        return lib().orc_jt()

    def sgs_stress(self, jt, jtt):
        lib().orc_sgs_stress(int(jt), int(jtt))

    # ---- running statistics and diagnostics
    STATS = ["u", "v", "w", "p", "u2", "v2", "w2", "u3", "v3", "w3", "u4", "v4", "w4", "uv", "uw", "vw",
             "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nut", "cs"]
    PROFILES = ["u", "v", "w", "p", "var_u", "var_v", "var_w", "txx", "tyy", "tzz", "R13", "txz", "St13",
                "R23", "tyz", "St23", "tke", "cs", "nut", "C13", "C23"]

    def stat_reset(self):
        lib().orc_stat_reset()

    def stat_accumulate(self):
        lib().orc_stat_accumulate()

    def stat_get(self):
        nz, ny, nx = self.prm.nz, self.prm.ny, self.prm.nx
        s = np.zeros((len(self.STATS), nz, ny, nx))
        us = np.zeros((ny, nx))
        it = ctypes.c_int()
        lib().orc_stat_get(_dp(s), _dp(us), ctypes.byref(it))
        return dict(zip(self.STATS, s)), us, it.value

    def stat_profiles(self):
        out = np.zeros((len(self.PROFILES), self.prm.nz))
        if lib().orc_stat_profiles(_dp(out)):
            raise ValueError("no accumulated statistics")
        return dict(zip(self.PROFILES, out))

    def diagnostics(self):
        out = np.zeros(4)
        lib().orc_diagnostics(_dp(out))
        return {"cfl": out[0], "mke": out[1], "u_star": out[2], "divu": out[3]}
