"""ctypes binding of libwpf_hip.so (the C ABI declared in include/wpf.h).

There is no CPU fallback: if the library is missing or no GPU is visible,
every entry point fails loudly.
"""
from __future__ import annotations

import ctypes
import os
import re

from .params import WpfDiag, WpfParams, WpfWell

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwpf_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "wpf.h")

WPF_OK = 0
WPF_ERR_ARG = 1
WPF_ERR_HIP = 2
WPF_ERR_NOCONV = 3
WPF_ERR_NAN = 4
WPF_ERR_NODEV = 5
WPF_ERR_STATE = 6

FIELD = {"permeab_rock": 0, "pressure": 1, "phase": 2, "concent": 3, "dphasedt": 4,
         "dconcentdt": 5, "permeab": 6, "velocity": 7, "face_t": 8, "face_v": 9}

SOLVER = {"cg": 0, "mgcg": 1, "mixed": 2}
KERNEL = {"perm": 0, "step_fields": 1, "smooth": 2, "matvec": 3, "true_resid": 4,
          "restrict": 5, "prolong": 6, "cg_update": 7}
FLAG_LEARNED_WARM = 1
FLAG_LEARNED_FCG = 2
FLAG_EXTRAPOLATE = 4
FLAG_CAPTURE = 8
FLAG_CHEBYSHEV = 16
FLAG_OP16 = 32
FLAG_F32_STENCIL = 64
FLAG_PROJECT = 128
FLAG_SPLIT_TAIL = 256


class WpfSolverOpts(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("maxit", ctypes.c_int32),
                ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("mg_levels", ctypes.c_int32), ("mg_sweeps", ctypes.c_int32),
                ("mg_omega", ctypes.c_double), ("mg_coarse_scale", ctypes.c_double),
                ("inner_rtol", ctypes.c_double), ("inner_maxit", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class WpfSolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("outer", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("learned", ctypes.c_int32),
                ("rnorm0", ctypes.c_double), ("rnorm", ctypes.c_double),
                ("bnorm", ctypes.c_double), ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class WpfEnsembleOpts(ctypes.Structure):
    _fields_ = [("stages", ctypes.c_int32), ("inner_iters", ctypes.c_int32),
                ("grid_blocks", ctypes.c_int32), ("tile_depth", ctypes.c_int32)]


class WpfEnsembleStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("fallbacks", ctypes.c_int32),
                ("iterations", ctypes.c_int64), ("seconds", ctypes.c_double),
                ("stages", ctypes.c_int32), ("inner_iters", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class WpfError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


_lib = None
P = ctypes.c_void_p
D = ctypes.POINTER(ctypes.c_double)
F = ctypes.POINTER(ctypes.c_float)
I32 = ctypes.c_int32


def header_symbols():
    """Names of every function declared in include/wpf.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wpf_[a-z0-9_]+)\s*\(", txt)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "wpf_create": ([ctypes.POINTER(WpfParams), ctypes.POINTER(ctypes.c_int32),
                        ctypes.POINTER(WpfWell), I32, I32, ctypes.POINTER(P)], I32),
        "wpf_destroy": ([P], I32),
        "wpf_device_count": ([ctypes.POINTER(I32)], I32),
        "wpf_status_string": ([I32], ctypes.c_char_p),
        "wpf_last_error": ([], ctypes.c_char_p),
        "wpf_upload": ([P, I32, D], I32),
        "wpf_download": ([P, I32, D], I32),
        "wpf_calc_permeability": ([P], I32),
        "wpf_solve_pressure": ([P, ctypes.POINTER(WpfSolverOpts), ctypes.POINTER(WpfSolveStats)], I32),
        "wpf_calc_velocity": ([P], I32),
        "wpf_calc_phase_deriv": ([P, D], I32),
        "wpf_calc_phase_deriv_noalloc": ([P, D], I32),
        "wpf_calc_concent_deriv": ([P], I32),
        "wpf_evolve_phase": ([P, ctypes.c_double], I32),
        "wpf_evolve_concent": ([P, ctypes.c_double], I32),
        "wpf_update_halo": ([P, I32], I32),
        "wpf_get_pressure": ([P, I32, I32, I32, D], I32),
        "wpf_diagnostics": ([P, ctypes.POINTER(WpfDiag)], I32),
        "wpf_step": ([P, ctypes.POINTER(WpfSolverOpts), I32, D, D, D,
                      ctypes.POINTER(WpfSolveStats)], I32),
        "wpf_breakthrough_step": ([P, ctypes.POINTER(I32)], I32),
        "wpf_set_breakthrough_plane": ([P, I32], I32),
        "wpf_profile_start": ([P, I32], I32),
        "wpf_profile_stop": ([P, D, ctypes.POINTER(I32)], I32),
        "wpf_stream_probe": ([I32, I32, I32, ctypes.c_int64, I32, D], I32),
        "wpf_net_num_params": ([], ctypes.c_int64),
        "wpf_net_create": ([I32, ctypes.POINTER(P)], I32),
        "wpf_net_destroy": ([P], I32),
        "wpf_net_get_params": ([P, F, ctypes.c_int64], I32),
        "wpf_net_set_params": ([P, F, ctypes.c_int64], I32),
        "wpf_net_forward": ([P, I32, I32, I32, F, F], I32),
        "wpf_net_gradient": ([P, I32, I32, I32, F, F, D, F], I32),
        "wpf_net_train": ([P, I32, I32, I32, I32, F, F, I32, ctypes.c_float, D], I32),
        "wpf_attach_net": ([P, P], I32),
        "wpf_set_fcg_period": ([P, I32], I32),
        "wpf_net_gradient_energy": ([P, I32, I32, I32, F, F, F, ctypes.POINTER(ctypes.c_uint8), D, F], I32),
        "wpf_net_dataset_reserve": ([P, I32, I32, I32, I32, I32], I32),
        "wpf_net_dataset_add": ([P, P, I32, ctypes.POINTER(I32)], I32),
        "wpf_net_train_energy": ([P, I32, ctypes.c_float, I32, D], I32),
        "wpf_capture_get": ([P, F, F], I32),
        "wpf_config_output": ([ctypes.POINTER(I32), ctypes.POINTER(I32)], I32),
        "wpf_result_write": ([ctypes.c_char_p, ctypes.c_double, ctypes.POINTER(WpfParams), I32, D, D, D, D], I32),
        "wpf_capture_fcg_count": ([P, ctypes.POINTER(I32)], I32),
        "wpf_capture_fcg_get": ([P, I32, F, F], I32),
        "wpf_set_tile_depth": ([P, I32], I32),
        "wpf_ensemble_create": ([ctypes.POINTER(P), I32, ctypes.POINTER(WpfSolverOpts),
                                 ctypes.POINTER(WpfEnsembleOpts), ctypes.POINTER(I32), D, D,
                                 ctypes.POINTER(P)], I32),
        "wpf_ensemble_tune": ([P, ctypes.POINTER(I32), ctypes.POINTER(I32)], I32),
        "wpf_ensemble_step": ([P, I32, ctypes.POINTER(WpfEnsembleStats)], I32),
        "wpf_ensemble_sync": ([P, ctypes.POINTER(I32), D, D, D], I32),
        "wpf_ensemble_destroy": ([P], I32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(status: int, where: str, ok=(WPF_OK,)):
    if status not in ok:
        msg = lib().wpf_last_error().decode(errors="replace")
        raise WpfError(status, where, msg or lib().wpf_status_string(status).decode())
    return status


def device_count() -> int:
    n = ctypes.c_int32(0)
    lib().wpf_device_count(ctypes.byref(n))
    return n.value
