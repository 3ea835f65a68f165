"""ctypes binding of libhawassi_hip.so (the C ABI in include/hawassi_hip.h)."""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhawassi_hip.so"

HW_PREC_FP64 = 0
HW_PREC_FP32 = 1
HW_FNO_BF16 = 0
HW_STEP_HOST = 0
HW_STEP_GRAPH = 1
HW_FNO_SPLIT = 1

_dp = ct.POINTER(ct.c_double)


class HwConfig2D(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int), ("ny", ct.c_int), ("precision", ct.c_int), ("flat", ct.c_int),
        ("grav", ct.c_double), ("cpeak", ct.c_double),
        ("kx", _dp), ("ky", _dp), ("aal", _dp), ("fbdy", _dp), ("wallchar", _dp), ("chi", _dp),
        ("c2d", _dp), ("l2d", _dp),
        ("gam_min", _dp), ("gam_plus", _dp), ("csq_min", _dp), ("csq_plus", _dp), ("om2_min", _dp), ("om2_plus", _dp),
        ("source_on", ct.c_int), ("vertical", ct.c_int), ("propagation", ct.c_int),
        ("new_gam", _dp), ("skew_gam", _dp),
        ("nline", ct.c_int), ("nt", ct.c_int),
        ("t_sig", _dp), ("sig", _dp), ("skew_sig", _dp),
        ("eps_abs", ct.c_double), ("eps_rel", ct.c_double),
        ("breaking_on", ct.c_int),
        ("KBC", ct.c_double), ("TC", ct.c_double), ("Tchar", ct.c_double), ("delb", ct.c_double),
        ("max_eta_init", ct.c_double), ("tstar", ct.c_double), ("rsearch", ct.c_double), ("current", ct.c_double),
        ("pbchar", _dp), ("depth", _dp), ("x", _dp), ("y", _dp),
        ("t0", ct.c_double),
        ("model", ct.c_int),
        ("gam_mid", _dp), ("csq_mid", _dp), ("om2_mid", _dp),
        ("full_grid", ct.c_int), ("wall", ct.c_int), ("refl_one", ct.c_int),
        ("grad_wall_x", _dp), ("grad_wall_y", _dp),
    ]


class HwConfig1D(ct.Structure):
    _fields_ = [
        ("n", ct.c_int), ("flat", ct.c_int), ("source_on", ct.c_int), ("propagation", ct.c_int),
        ("breaking_on", ct.c_int), ("dx", ct.c_double),
        ("k", _dp), ("aal", _dp), ("chi", _dp), ("damp", _dp), ("csqr", _dp), ("lop", _dp),
        ("csq_min", _dp), ("csq_plus", _dp), ("gam_min", _dp), ("gam_plus", _dp),
        ("gamx_re", _dp), ("gamx_im", _dp), ("gamx_skew_re", _dp), ("gamx_skew_im", _dp),
        ("nt", ct.c_int), ("t_sig", _dp), ("sig", _dp), ("sig_skew", _dp),
        ("KBC", ct.c_double), ("TC", ct.c_double), ("Tstar", ct.c_double), ("max_eta_init", ct.c_double),
        ("delb", ct.c_double),
        ("xbstart", ct.c_int), ("xbend", ct.c_int), ("xbendL", ct.c_int), ("xbstartR", ct.c_int),
        ("lenIF", ct.c_int), ("ILL", ct.c_int), ("IFF", ct.c_int),
        ("depth", _dp), ("eps_abs", ct.c_double), ("eps_rel", ct.c_double), ("t0", ct.c_double), ("h0", ct.c_double),
        ("csq_mid", _dp), ("gam_mid", _dp),
        ("order", ct.c_int),
        ("wall_on", ct.c_int), ("idx_wall", ct.c_int), ("refl_coef", ct.c_double), ("t_wall0", ct.c_double),
        ("gamw_skew_re", _dp), ("gamw_skew_im", _dp),
    ]


class HwStats(ct.Structure):
    _fields_ = [("rhs_evals", ct.c_long), ("steps_accepted", ct.c_long), ("steps_rejected", ct.c_long),
                ("rhs_ms_total", ct.c_double), ("rhs_timed", ct.c_long)]


class HwKin1DConfig(ct.Structure):
    _fields_ = [("n", ct.c_int), ("ix0", ct.c_int), ("nxout", ct.c_int), ("nz", ct.c_int), ("flat", ct.c_int),
                ("ntab", ct.c_int), ("dx", ct.c_double), ("grav", ct.c_double), ("hmin", ct.c_double),
                ("hplus", ct.c_double), ("depth", ct.POINTER(ct.c_double)), ("k", ct.POINTER(ct.c_double)),
                ("z", ct.POINTER(ct.c_double)), ("htab", ct.POINTER(ct.c_double)),
                ("gam_hmin", ct.POINTER(ct.c_double)), ("gam_hplus", ct.POINTER(ct.c_double)),
                ("dztab", ct.POINTER(ct.c_double)), ("gam_dzmin", ct.POINTER(ct.c_double)),
                ("gam_dzplus", ct.POINTER(ct.c_double))]


class HwFnoConfig(ct.Structure):
    _fields_ = [("n", ct.c_int), ("width", ct.c_int), ("modes", ct.c_int), ("layers", ct.c_int),
                ("hidden", ct.c_int), ("in_ch", ct.c_int), ("out_ch", ct.c_int), ("max_batch", ct.c_int),
                ("seed", ct.c_ulonglong)]


EXPORTS = [
    "hw_ctx_create_2d", "hw_ctx_destroy", "hw_rhs", "hw_rhs_device", "hw_set_state", "hw_get_state",
    "hw_evolve_apply", "hw_integrate", "hw_integrate_batch", "hw_set_step_mode", "hw_ctx_create_2d_members", "hw_integrate_members", "hw_get_fields", "hw_get_stats", "hw_reset_stats", "hw_sync",
    "hw_fft2_forward", "hw_fft2_inverse", "hw_fft2_forward_noalloc", "hw_supported", "hw_supported_full", "hw_last_error", "hw_version",
    "hw_set_tprev", "hw_get_breaking",
    "hw1d_create", "hw1d_destroy", "hw1d_rhs", "hw1d_rhs_at", "hw1d_set_state", "hw1d_get_state",
    "hw1d_set_time", "hw1d_get_time", "hw1d_integrate", "hw1d_set_round_limit", "hw1d_get_breaking",
    "hw1d_get_stats", "hw1d_kinematics",
    "hw_fno_create", "hw_fno_destroy", "hw_fno_param_count", "hw_fno_get_params", "hw_fno_set_params",
    "hw_fno_forward", "hw_fno_loss_grad", "hw_fno_set_dataset", "hw_fno_train_step", "hw_fno_gemm_check", "hw_fno_set_precision", "hw_fno_gemm_flops", "hw_fno_gemm_bytes",
]

_lib = None


class HipPathMissing(RuntimeError):
    """The HIP library is not built; there is no CPU fallback."""


class _Partial:
    """A library that holds only some of the entry points (the autotuner's
    candidate builds of one 2-D unit): declarations of the missing ones go
    nowhere, calls to them fail."""

    class _Sink:
        pass

    def __init__(self, lib):
        object.__setattr__(self, "_l", lib)

    def __getattr__(self, name):
        try:
            return getattr(self._l, name)
        except AttributeError:
            return _Partial._Sink()


def load(path: Path | None = None, partial: bool = False) -> ct.CDLL:
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise HipPathMissing(f"{p} not found: build it with `python -m hawassi_c_amd.build` "
                             "(the product has no CPU fallback)")
    lib = ct.CDLL(str(p))
    if partial:
        lib = _Partial(lib)
    vp = ct.c_void_p
    lib.hw_ctx_create_2d.argtypes = [ct.POINTER(HwConfig2D), ct.POINTER(vp)]
    lib.hw_ctx_destroy.argtypes = [vp]
    lib.hw_ctx_destroy.restype = None
    lib.hw_rhs.argtypes = [ct.c_double, _dp, _dp, vp]
    lib.hw_rhs_device.argtypes = [vp, ct.c_double, vp, vp]
    lib.hw_set_state.argtypes = [vp, _dp]
    lib.hw_get_state.argtypes = [vp, _dp]
    lib.hw_evolve_apply.argtypes = [vp, _dp, ct.c_double, _dp]
    lib.hw_integrate.argtypes = [vp, _dp, ct.c_double, _dp]
    lib.hw_integrate_batch.argtypes = [ct.POINTER(vp), ct.c_int, _dp, ct.c_double, _dp]
    lib.hw_get_fields.argtypes = [vp, _dp, _dp, _dp]
    lib.hw_get_stats.argtypes = [vp, ct.POINTER(HwStats)]
    lib.hw_reset_stats.argtypes = [vp]
    lib.hw_sync.argtypes = [vp]
    lib.hw_set_tprev.argtypes = [vp, ct.c_double]
    lib.hw_get_breaking.argtypes = [vp, _dp, ct.POINTER(ct.c_int)]
    lib.hw_fft2_forward.argtypes = [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp]
    lib.hw_fft2_inverse.argtypes = [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _dp]
    lib.hw_fft2_forward_noalloc.argtypes = [ct.c_int, ct.c_int, ct.c_int, vp, vp, vp, vp]
    lib.hw_supported.argtypes = [ct.c_int, ct.c_int, ct.c_int]
    lib.hw_supported_full.argtypes = [ct.c_int, ct.c_int, ct.c_int]
    lib.hw1d_create.argtypes = [ct.POINTER(HwConfig1D), ct.c_int, ct.c_int, ct.POINTER(vp)]
    lib.hw1d_destroy.argtypes = [vp]
    lib.hw1d_destroy.restype = None
    lib.hw1d_rhs.argtypes = [ct.c_double, _dp, _dp, vp]
    lib.hw1d_rhs_at.argtypes = [vp, ct.c_int, ct.c_double, _dp, _dp]
    lib.hw1d_set_state.argtypes = [vp, ct.c_int, _dp]
    lib.hw1d_get_state.argtypes = [vp, ct.c_int, _dp]
    lib.hw1d_set_time.argtypes = [vp, ct.c_int, ct.c_double, ct.c_double]
    lib.hw1d_get_time.argtypes = [vp, ct.c_int, _dp, _dp]
    lib.hw1d_integrate.argtypes = [vp, ct.c_double]
    lib.hw1d_set_round_limit.argtypes = [vp, ct.c_long]
    lib.hw1d_get_breaking.argtypes = [vp, ct.c_int, _dp, ct.POINTER(ct.c_int)]
    lib.hw1d_get_stats.argtypes = [vp, ct.c_int, ct.POINTER(HwStats)]
    lib.hw1d_kinematics.argtypes = [ct.POINTER(HwKin1DConfig), _dp, _dp, _dp, _dp, ct.c_double, ct.c_double, _dp,
                                    _dp, _dp]
    _fp = ct.POINTER(ct.c_float)
    lib.hw_set_step_mode.argtypes = [vp, ct.c_int]
    lib.hw_ctx_create_2d_members.argtypes = [ct.POINTER(HwConfig2D), ct.c_int, ct.POINTER(vp)]
    lib.hw_integrate_members.argtypes = [vp, _dp, ct.c_double, _dp]
    lib.hw_fno_create.argtypes = [ct.POINTER(HwFnoConfig), ct.POINTER(vp)]
    lib.hw_fno_destroy.argtypes = [vp]
    lib.hw_fno_destroy.restype = None
    lib.hw_fno_param_count.argtypes = [vp]
    lib.hw_fno_param_count.restype = ct.c_longlong
    lib.hw_fno_get_params.argtypes = [vp, _fp]
    lib.hw_fno_set_params.argtypes = [vp, _fp]
    lib.hw_fno_forward.argtypes = [vp, ct.c_int, _fp, _fp]
    lib.hw_fno_loss_grad.argtypes = [vp, ct.c_int, _fp, _fp, _dp, _fp]
    lib.hw_fno_set_dataset.argtypes = [vp, ct.c_int, _fp, _fp]
    lib.hw_fno_train_step.argtypes = [vp, ct.c_int, ct.POINTER(ct.c_int), ct.c_double, _dp]
    lib.hw_fno_gemm_check.argtypes = [ct.c_int, ct.c_int, ct.c_int, _fp, _fp, _fp, ct.c_int]
    lib.hw_fno_set_precision.argtypes = [vp, ct.c_int]
    lib.hw_fno_gemm_flops.argtypes = [vp]
    lib.hw_fno_gemm_flops.restype = ct.c_double
    lib.hw_fno_gemm_bytes.argtypes = [vp]
    lib.hw_fno_gemm_bytes.restype = ct.c_double
    lib.hw_last_error.restype = ct.c_char_p
    lib.hw_version.restype = ct.c_char_p
    if path is None:
        _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous float64"
    return a.ctypes.data_as(_dp)


class HwError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().hw_last_error().decode(errors="replace")
        raise HwError(f"{what} failed (code {rc}): {msg}")
