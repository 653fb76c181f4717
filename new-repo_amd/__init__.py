"""MI355X LES time step: Python view of the C-ABI in include/les_gpu.h.

The product is liblesgpu.so (hand-written HIP kernels for gfx950) driven
through its C-ABI.  The Fortran host modules under fortran/ bind the same
entry points with iso_c_binding; this module binds them with ctypes for the
tests and the benchmark.  Nothing here computes: every call goes to the
device library, and a missing library raises instead of falling back.

Field arrays on the host are float64 with shape (nz+2, ny, nx) -- C order of
the reference's Fortran (nx, ny, 0:nz+1) arrays.
"""
from __future__ import annotations

import ctypes
import pathlib
from typing import Dict, Optional

import numpy as np

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "liblesgpu.so"

FIELDS = ["u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
          "p", "dpdx", "dpdy", "dpdz", "rhsx", "rhsy", "rhsz", "rhsx_f", "rhsy_f", "rhsz_f",
          "txx", "txy", "txz", "tyy", "tyz", "tzz", "cs_opt2", "nu_t", "F_LM", "F_MM", "F_QN", "F_NN",
          "beta", "u_lag", "v_lag", "w_lag", "s11", "s12", "s13", "s22", "s23", "s33"]
FID: Dict[str, int] = {n: i for i, n in enumerate(FIELDS)}
LAYERS = {"lbc_u": 0, "lbc_v": 1, "ustar": 2, "bottom_wk": 3, "top_wk": 4, "z0": 5}
# running sums, in the order of the reference's RAV_<jt>_3D.dat
STATS = ["u", "v", "w", "p", "u2", "v2", "w2", "u3", "v3", "w3", "u4", "v4", "w4", "uv", "uw", "vw",
         "txx", "tyy", "tzz", "txy", "txz", "tyz", "dudz", "dvdz", "nut", "cs"]
# rows of les_stat_profiles
PROFILES = ["u", "v", "w", "p", "var_u", "var_v", "var_w", "txx", "tyy", "tzz", "R13", "txz", "St13",
            "R23", "tyz", "St23", "tke", "cs", "nut", "C13", "C23"]

# TKE budget sums (les_tke_download rows), the names of the reference's
# ta1_profiles/<name>.out files
TKE = ["u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz",
       "uu", "vv", "ww", "uv", "uw", "vw", "txx", "tyy", "tzz", "txy", "txz", "tyz", "cs",
       "uuw", "vvw", "www", "uww", "wtzz", "utxz", "vtyz", "dxx", "dyy", "dzz", "dxy", "dxz", "dyz",
       "dudx2", "dudy2", "dudz2", "dvdx2", "dvdy2", "dvdz2", "dwdx2", "dwdy2", "dwdz2",
       "dudxdwdx", "dudydwdy", "dudzdwdz", "p", "dpdx", "dpdy", "dpdz", "pw", "pu",
       "pdudx", "pdvdy", "pdwdz", "pdwdx", "pdudz"]

STATUS = {0: "ok", 1: "invalid option or model", 3: "out of memory", 4: "double initialisation",
          5: "HIP runtime error", 9: "NaN or singular solve", 10: "bad shape or id"}


class LesError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}: {STATUS.get(code, '?')}] {msg}")
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("nx", "ny", "nz", "sgs_model", "cs_count", "dyn_init", "lagran_init", "wall_model",
                 "ubc_u_neu", "ubc_v_neu", "halo_mode", "coriolis", "precision", "judge_every", "dealias")] + \
               [(n, ctypes.c_double) for n in
                ("lx", "ly", "lz", "dt", "z0", "pfx", "pfy", "cs0", "coriol", "ug", "vg",
                 "ubc_u", "ubc_v", "ubc_w", "re", "learned_zmin")]


_lib = None


def library() -> ctypes.CDLL:
    """Load liblesgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with new-repo_amd/build.py")
        # torch ships its own HIP/HSA runtime under the same sonames as
        # /opt/rocm's.  Loaded first, it satisfies this library's dependency,
        # so one runtime serves both and device pointers can be exchanged;
        # loaded after, a second HSA runtime cannot open the GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(str(LIB_PATH))
        vp, dp, ip, lp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), \
            ctypes.POINTER(ctypes.c_long)
        sig = {
            "les_init": [ctypes.POINTER(_Params)], "les_finalize": [], "les_set_ic": [dp, dp, dp],
            "les_upload_field": [ctypes.c_int, dp], "les_download_field": [ctypes.c_int, dp],
            "les_upload_layer": [ctypes.c_int, dp], "les_download_layer": [ctypes.c_int, dp],
            "les_field_ptr": [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p), ip],
            "les_step": [ctypes.c_int], "les_sync": [], "les_jt_total": [], "les_set_jt_total": [ctypes.c_int],
            "les_check_nan": [ip], "les_timers": [dp, ctypes.c_int], "les_set_timing": [ctypes.c_int], "les_kernel_stat": [ctypes.c_int, ip, ctypes.c_char_p, ctypes.c_int, dp], "les_set_overlap": [ctypes.c_int],
            "les_wall_law": [], "les_compute_vel_grad": [],
            "les_grad_velocity_filt": [ctypes.c_int] * 3, "les_ddz_uv": [ctypes.c_int] * 2,
            "les_ddz_w": [ctypes.c_int] * 2, "les_sgs_stress": [ctypes.c_int] * 2,
            "les_compute_dealias_convec": [ctypes.c_int] * 3, "les_div_stress": [ctypes.c_int] * 6,
            "les_compute_rhs_adv": [], "les_press_stag_array": [ctypes.c_double],
            "les_test_filter": [ctypes.c_int] * 2,
            "les_ab2": [ctypes.c_double, ctypes.c_double], "les_projection": [ctypes.c_double],
            "les_fft2d_r2c": [vp, vp, ctypes.c_int, ctypes.c_double], "les_fft2d_c2r": [vp, vp, ctypes.c_int],
            "les_sgs_samples": [vp, vp, ctypes.c_long, lp, ctypes.c_int],
            "les_mlp_set_weights": [ctypes.POINTER(ctypes.c_float), ctypes.c_long],
            "les_mlp_layout": [lp], "les_mlp_judge_stats": [dp],
            "les_mlp_get_weights": [ctypes.POINTER(ctypes.c_float)],
            "les_mlp_prepare": [vp, vp, ctypes.c_long, vp, vp],
            "les_mlp_grad": [vp, vp, ctypes.c_long, ctypes.POINTER(ctypes.c_float), dp],
            "les_mlp_train_step": [vp, vp, ctypes.c_long, ctypes.c_double, dp],
            "les_mlp_forward": [vp, ctypes.c_long, vp],
            "les_stat_init": [], "les_stat_reset": [], "les_stat_accumulate": [],
            "les_stat_download": [dp, dp, ip], "les_stat_profiles": [dp], "les_stat_profiles_noalloc": [dp], "les_running_diagnostics": [dp],
            "les_tke_init": [], "les_tke_reset": [], "les_tke_accumulate": [], "les_tke_download": [dp],
            "les_set_geostrophic": [ctypes.c_double, ctypes.c_double],
        }
        for name, args in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        lib.les_last_error.argtypes = []
        lib.les_last_error.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def check(code: int) -> None:
    if code != 0:
        raise LesError(code, library().les_last_error().decode())


class LesGpu:
    """Device-resident LES state (one per process, like the reference's modules)."""

    def __init__(self, prm):
        self.prm = prm
        self.lib = library()
        c = _Params()
        for name, _ in _Params._fields_:
            setattr(c, name, getattr(prm, name))
        check(self.lib.les_init(ctypes.byref(c)))
        self.shape = (prm.nz + 2, prm.ny, prm.nx)
        self._open = True

    # ---- lifetime
    def close(self) -> None:
        if self._open:
            self.lib.les_finalize()
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- data
    def set_ic(self, u, v, w) -> None:
        u, v, w = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w))
        check(self.lib.les_set_ic(_dp(u), _dp(v), _dp(w)))

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        check(self.lib.les_download_field(FID[name], _dp(out)))
        return out

    def upload(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        check(self.lib.les_upload_field(FID[name], _dp(a)))

    def layer(self, name: str) -> np.ndarray:
        out = np.empty(self.shape[1:])
        check(self.lib.les_download_layer(LAYERS[name], _dp(out)))
        return out

    def set_layer(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape[1:])
        check(self.lib.les_upload_layer(LAYERS[name], _dp(a)))

    def field_ptr(self, name: str):
        p = ctypes.c_void_p()
        eb = ctypes.c_int()
        check(self.lib.les_field_ptr(FID[name], ctypes.byref(p), ctypes.byref(eb)))
        return p.value, eb.value

    # ---- stepping
    def step(self, n: int = 1) -> None:
        check(self.lib.les_step(n))

    def sync(self) -> None:
        check(self.lib.les_sync())

    @property
    def jt_total(self) -> int:
        return self.lib.les_jt_total()

    def set_jt_total(self, jt: int) -> None:
        check(self.lib.les_set_jt_total(jt))

    def check_nan(self) -> bool:
        found = ctypes.c_int(0)
        rc = self.lib.les_check_nan(ctypes.byref(found))
        if rc not in (0, 9):
            check(rc)
        return bool(found.value)

    def set_timing(self, on: bool) -> None:
        check(self.lib.les_set_timing(int(on)))

    def set_overlap(self, on: bool) -> None:
        """les_step on two streams (default) or one."""
        check(self.lib.les_set_overlap(int(on)))

    def timers(self, reset: bool = False) -> np.ndarray:
        t = np.zeros(5 + 3 * 9)
        check(self.lib.les_timers(_dp(t), int(reset)))
        return t

    def kernel_stats(self) -> Dict[str, Dict[str, float]]:
        """Per-kernel HIP-event statistics of the timing mode since the last
        timers(reset=True): name -> ms, calls, bytes, family."""
        n = ctypes.c_int()
        check(self.lib.les_kernel_stat(-1, ctypes.byref(n), None, 0, None))
        out = {}
        buf = ctypes.create_string_buffer(200)
        v = np.zeros(4)
        for i in range(n.value):
            check(self.lib.les_kernel_stat(i, ctypes.byref(n), buf, 200, _dp(v)))
            out[buf.value.decode()] = {"ms": float(v[0]), "calls": int(v[1]), "bytes": float(v[2]), "family": int(v[3])}
        return out

    # ---- the compute_* procedures
    def wall_law(self) -> None:
        check(self.lib.les_wall_law())

    def compute_vel_grad(self) -> None:
        check(self.lib.les_compute_vel_grad())

    def grad_velocity_filt(self, f: str, dfdx: str, dfdy: str) -> None:
        check(self.lib.les_grad_velocity_filt(FID[f], FID[dfdx], FID[dfdy]))

    def ddz_uv(self, dfdz: str, f: str) -> None:
        check(self.lib.les_ddz_uv(FID[dfdz], FID[f]))

    def ddz_w(self, dfdz: str, f: str) -> None:
        check(self.lib.les_ddz_w(FID[dfdz], FID[f]))

    def sgs_stress(self, jt: int, jtt: int) -> None:
        check(self.lib.les_sgs_stress(jt, jtt))

    def compute_dealias_convec(self, cx="rhsx", cy="rhsy", cz="rhsz") -> None:
        check(self.lib.les_compute_dealias_convec(FID[cx], FID[cy], FID[cz]))

    def div_stress(self, rhs: str, tx: str, ty: str, tz: str, stag: bool, save_wk: bool = False) -> None:
        check(self.lib.les_div_stress(FID[rhs], FID[tx], FID[ty], FID[tz], int(stag), int(save_wk)))

    def compute_rhs_adv(self) -> None:
        check(self.lib.les_compute_rhs_adv())

    def press_stag_array(self, tadv1: float) -> None:
        check(self.lib.les_press_stag_array(float(tadv1)))

    def ab2(self, tadv1: float, tadv2: float) -> None:
        check(self.lib.les_ab2(float(tadv1), float(tadv2)))

    def projection(self, tadv1: float) -> None:
        check(self.lib.les_projection(float(tadv1)))

    def test_filter(self, f: str, alpha: int) -> None:
        check(self.lib.les_test_filter(FID[f], int(alpha)))

    # ---- learned SGS surrogate
    def mlp_size(self) -> int:
        n = ctypes.c_long()
        check(self.lib.les_mlp_layout(ctypes.byref(n)))
        return n.value

    def mlp_set_weights(self, w) -> None:
        a = np.ascontiguousarray(w, dtype=np.float32)
        check(self.lib.les_mlp_set_weights(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), a.size))

    def mlp_get_weights(self) -> np.ndarray:
        out = np.zeros(self.mlp_size(), dtype=np.float32)
        check(self.lib.les_mlp_get_weights(out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def mlp_prepare(self, d_raw: int, d_tgt: int, n: int, d_x: int, d_y: int) -> None:
        check(self.lib.les_mlp_prepare(ctypes.c_void_p(d_raw), ctypes.c_void_p(d_tgt), n, ctypes.c_void_p(d_x),
                                       ctypes.c_void_p(d_y)))

    def mlp_grad(self, d_x: int, d_y: int, n: int):
        g = np.zeros(self.mlp_size() - 29, dtype=np.float32)
        loss = ctypes.c_double()
        check(self.lib.les_mlp_grad(ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), n,
                                    g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)))
        return g, loss.value

    def mlp_train_step(self, d_x: int, d_y: int, n: int, lr: float) -> float:
        loss = ctypes.c_double()
        check(self.lib.les_mlp_train_step(ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), n, float(lr), ctypes.byref(loss)))
        return loss.value

    def mlp_forward(self, d_x: int, n: int, d_out: int) -> None:
        check(self.lib.les_mlp_forward(ctypes.c_void_p(d_x), n, ctypes.c_void_p(d_out)))

    def sgs_samples(self, d_x: int, d_y: int, max_samples: int, stride: int) -> int:
        n = ctypes.c_long()
        check(self.lib.les_sgs_samples(ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), max_samples, ctypes.byref(n), stride))
        return n.value

    def judge_stats(self) -> np.ndarray:
        out = np.zeros(4)
        check(self.lib.les_mlp_judge_stats(_dp(out)))
        return out

    # ---- running statistics and diagnostics
    def stat_init(self) -> None:
        check(self.lib.les_stat_init())

    def stat_reset(self) -> None:
        check(self.lib.les_stat_reset())

    def stat_accumulate(self) -> None:
        check(self.lib.les_stat_accumulate())

    def stat_download(self):
        """(sums, ustar_sum, steps): sums[name] has shape (nz, ny, nx)."""
        nz, ny, nx = self.prm.nz, self.prm.ny, self.prm.nx
        s = np.empty((len(STATS), nz, ny, nx))
        us = np.empty((ny, nx))
        it = ctypes.c_int()
        check(self.lib.les_stat_download(_dp(s), _dp(us), ctypes.byref(it)))
        return dict(zip(STATS, s)), us, it.value

    def stat_profiles(self) -> Dict[str, np.ndarray]:
        out = np.empty((len(PROFILES), self.prm.nz))
        check(self.lib.les_stat_profiles(_dp(out)))
        return dict(zip(PROFILES, out))

    def stat_profiles_noalloc(self) -> Dict[str, np.ndarray]:
        """les_stat_profiles_noalloc: the same profiles, partial sums in the
        library's fixed device array (nothing allocated)."""
        out = np.empty((len(PROFILES), self.prm.nz))
        check(self.lib.les_stat_profiles_noalloc(_dp(out)))
        return dict(zip(PROFILES, out))

    def set_geostrophic(self, ug: float, vg: float) -> None:
        check(self.lib.les_set_geostrophic(float(ug), float(vg)))

    # ---- TKE budget
    def tke_init(self) -> None:
        check(self.lib.les_tke_init())

    def tke_reset(self) -> None:
        check(self.lib.les_tke_reset())

    def tke_accumulate(self) -> None:
        check(self.lib.les_tke_accumulate())

    def tke_download(self) -> Dict[str, np.ndarray]:
        out = np.empty((len(TKE), self.prm.nz))
        check(self.lib.les_tke_download(_dp(out)))
        return dict(zip(TKE, out))

    def running_diagnostics(self) -> Dict[str, float]:
        out = np.zeros(4)
        check(self.lib.les_running_diagnostics(_dp(out)))
        return {"cfl": out[0], "mke": out[1], "u_star": out[2], "divu": out[3]}

    def fft2d_r2c(self, d_in: int, d_out: int, nplanes: int, scale: float = 1.0) -> None:
        check(self.lib.les_fft2d_r2c(ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), nplanes, float(scale)))

    def fft2d_c2r(self, d_in: int, d_out: int, nplanes: int) -> None:
        check(self.lib.les_fft2d_c2r(ctypes.c_void_p(d_in), ctypes.c_void_p(d_out), nplanes))


def load_params_module():
    import importlib.util
    import sys
    name = "new_repo_amd_params"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "params.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_ic_module():
    import importlib.util
    import sys
    name = "new_repo_amd_ic"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "ic.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod
