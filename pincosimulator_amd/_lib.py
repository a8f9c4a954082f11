"""ctypes bindings to libtwmc_hip.so (the C ABI declared in include/twmc_hip.h).

The library is built in-tree by ``make -C pincosimulator_amd/csrc`` (driven by
``__graft_entry__.build()``).  There is no fallback: if the library is missing
or no GPU is present, the calls that need it raise ``TWMCError``.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtwmc_hip.so")

TWMC_OK = 0
ERR_NAMES = {-1: "TWMC_ERR_INVALID", -2: "TWMC_ERR_HIP", -3: "TWMC_ERR_NOMEM",
             -4: "TWMC_ERR_STATE", -5: "TWMC_ERR_UNSUPPORTED"}
NOISE_NONE, NOISE_UNIFORM, NOISE_GAUSSIAN = 0, 1, 2
HOP_FFT, HOP_MATRIX = 0, 1
ONSITE_OMEGA, ONSITE_DETUNING_MINUS_E = 0, 1
INIT_SETTINGS, INIT_FIXEDPOINT, INIT_PREVIOUS, INIT_GAUSSIAN = 0, 1, 2, 3
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_FUSED64, KERNEL_MFMA, KERNEL_STENCIL64, KERNEL_BAND, KERNEL_STACK = range(7)
KERNEL_NAMES = {0: "auto", 1: "generic", 2: "fused64", 3: "mfma", 4: "stencil64", 5: "band", 6: "stack"}
OUT_OBS, OUT_TRAJ, OUT_FINAL = 1, 2, 4

# every symbol include/twmc_hip.h declares
EXPORTS = [
    "twmc_abi_version", "twmc_last_error", "twmc_create", "twmc_destroy", "twmc_batch_begin",
    "twmc_batch_info_get", "twmc_batch_advance", "twmc_batch_sync", "twmc_batch_read", "twmc_run",
    "twmc_set_timing", "twmc_query", "twmc_reset_stats", "twmc_progress", "twmc_noisy_matrix",
    "twmc_advance_noalloc",
    "twmc_sweep_create", "twmc_sweep_destroy", "twmc_sweep_last_error", "twmc_sweep_set_batching",
    "twmc_sweep_get_batching", "twmc_sweep_autotune", "twmc_sweep_run", "twmc_sweep_query",
    "twmc_sweep_mode_bins",
    "twmc_mlp_create", "twmc_mlp_destroy", "twmc_mlp_last_error", "twmc_mlp_n_params", "twmc_mlp_get_params",
    "twmc_mlp_set_params", "twmc_mlp_forward", "twmc_mlp_loss_grad", "twmc_mlp_train",
]

_dp = C.POINTER(C.c_double)


class TWMCError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class Matrix(C.Structure):
    _fields_ = [("base", _dp), ("value", C.c_double * 2), ("noise_type", C.c_int32),
                ("amp", _dp), ("amp_value", C.c_double * 2)]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("hopping", C.c_int32), ("onsite", C.c_int32),
        ("precision", C.c_int32), ("kernel", C.c_int32),
        ("dt", C.c_double), ("gamma", C.c_double), ("J", C.c_double), ("detuning", C.c_double),
        ("omega", Matrix), ("U", Matrix), ("F", Matrix), ("beta_init", Matrix), ("beta_sigma", Matrix),
        ("coupling", _dp), ("seed", C.c_uint64), ("frame_steps", C.c_uint64),
        ("F_n_times", C.c_uint64), ("F_times", _dp), ("F_table", _dp),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("traj_ids", C.POINTER(C.c_uint64)), ("n_traj", C.c_uint64), ("init_mode", C.c_int32),
        ("store_initial", C.c_int32), ("beta_prev", _dp), ("step0", C.c_uint64),
        ("t_start", C.c_double), ("t_end", C.c_double), ("outputs", C.c_uint32),
        ("gauss_mean", _dp), ("gauss_nk", _dp),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("n_frames", C.c_uint64), ("n_steps", C.c_uint64), ("n_segments", C.c_uint64),
        ("segments_done", C.c_uint64), ("t_final", C.c_double), ("kernel", C.c_int32),
        ("truncated", C.c_int32),
    ]


class Output(C.Structure):
    _fields_ = [("traj", _dp), ("beta_final", _dp), ("sum_n", _dp), ("sum_n2", _dp),
                ("sum_nk", _dp), ("sum_beta", _dp), ("traj_m2", _dp), ("traj_m4", _dp)]


class Stats(C.Structure):
    _fields_ = [
        ("step_launches", C.c_uint64), ("step_kernel_ms", C.c_double), ("site_steps", C.c_uint64),
        ("reduce_kernel_ms", C.c_double), ("reduce_launches", C.c_uint64), ("kernel", C.c_int32),
        ("grid", C.c_int32), ("block", C.c_int32), ("lds_bytes", C.c_int32),
    ]


class SweepPoint(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("n_traj", C.c_uint64), ("traj_id0", C.c_uint64),
        ("J", C.c_double), ("gamma", C.c_double), ("omega", C.c_double * 2), ("U", C.c_double * 2),
        ("F", C.c_double * 2), ("beta_init", C.c_double * 2), ("beta_sigma", C.c_double * 2),
    ]


class MlpConfig(C.Structure):
    _fields_ = [("n_in", C.c_int32), ("n_hidden", C.c_int32), ("n_hidden_layers", C.c_int32),
                ("n_out", C.c_int32), ("seed", C.c_uint64)]


class MlpTrainOpts(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("shuffle_seed", C.c_uint64)]


_lib = None


def load():
    """Loads libtwmc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TWMCError(-2, f"{LIB_PATH} not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.twmc_abi_version.restype = C.c_int
    L.twmc_last_error.argtypes = [vp]
    L.twmc_last_error.restype = C.c_char_p
    L.twmc_create.argtypes = [C.POINTER(Params), C.c_int, C.POINTER(vp)]
    L.twmc_destroy.argtypes = [vp]
    L.twmc_destroy.restype = None
    L.twmc_batch_begin.argtypes = [vp, C.POINTER(BatchDesc)]
    L.twmc_batch_info_get.argtypes = [vp, C.POINTER(BatchInfo)]
    L.twmc_batch_advance.argtypes = [vp, C.c_uint64]
    L.twmc_batch_sync.argtypes = [vp]
    L.twmc_batch_read.argtypes = [vp, C.POINTER(Output)]
    L.twmc_run.argtypes = [vp, C.POINTER(BatchDesc), C.POINTER(Output)]
    L.twmc_set_timing.argtypes = [vp, C.c_int]
    L.twmc_query.argtypes = [vp, C.POINTER(Stats)]
    L.twmc_reset_stats.argtypes = [vp]
    L.twmc_progress.argtypes = [vp]
    L.twmc_progress.restype = C.c_float
    L.twmc_noisy_matrix.argtypes = [C.POINTER(Matrix), C.c_uint64, C.c_uint64, C.c_uint64,
                                    C.c_uint32, _dp]
    i32p = C.POINTER(C.c_int32)
    L.twmc_sweep_create.argtypes = [C.POINTER(SweepPoint), C.c_uint64, C.c_double, C.c_uint64, C.c_uint64,
                                    C.c_int, C.POINTER(vp)]
    L.twmc_sweep_destroy.argtypes = [vp]
    L.twmc_sweep_destroy.restype = None
    L.twmc_sweep_last_error.argtypes = [vp]
    L.twmc_sweep_last_error.restype = C.c_char_p
    L.twmc_sweep_set_batching.argtypes = [vp, i32p]
    L.twmc_sweep_get_batching.argtypes = [vp, i32p]
    L.twmc_sweep_autotune.argtypes = [vp, C.c_int32, i32p]
    L.twmc_sweep_run.argtypes = [vp, C.c_double, C.c_double, C.c_int32, _dp, _dp, _dp, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_double)]
    L.twmc_sweep_mode_bins.argtypes = [vp, C.c_uint64, C.POINTER(C.c_uint8)]
    L.twmc_sweep_query.argtypes = [vp, C.POINTER(Stats)]
    for name in ("twmc_sweep_create", "twmc_sweep_set_batching", "twmc_sweep_get_batching",
                 "twmc_sweep_autotune", "twmc_sweep_run", "twmc_sweep_query", "twmc_sweep_mode_bins"):
        getattr(L, name).restype = C.c_int
    fp = C.POINTER(C.c_float)
    L.twmc_mlp_create.argtypes = [C.POINTER(MlpConfig), C.c_int, C.POINTER(vp)]
    L.twmc_mlp_destroy.argtypes = [vp]
    L.twmc_mlp_destroy.restype = None
    L.twmc_mlp_last_error.argtypes = [vp]
    L.twmc_mlp_last_error.restype = C.c_char_p
    L.twmc_mlp_n_params.argtypes = [vp]
    L.twmc_mlp_n_params.restype = C.c_uint64
    L.twmc_mlp_get_params.argtypes = [vp, fp]
    L.twmc_mlp_set_params.argtypes = [vp, fp]
    L.twmc_mlp_forward.argtypes = [vp, fp, C.c_uint64, fp]
    L.twmc_mlp_loss_grad.argtypes = [vp, fp, fp, fp, C.c_uint64, _dp, fp]
    L.twmc_mlp_train.argtypes = [vp, fp, fp, fp, C.c_uint64, C.POINTER(MlpTrainOpts), _dp]
    for name in ("twmc_mlp_create", "twmc_mlp_get_params", "twmc_mlp_set_params", "twmc_mlp_forward",
                 "twmc_mlp_loss_grad", "twmc_mlp_train"):
        getattr(L, name).restype = C.c_int
    L.twmc_advance_noalloc.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, vp, vp, vp]
    for name in ("twmc_create", "twmc_batch_begin", "twmc_batch_info_get", "twmc_batch_advance",
                 "twmc_batch_sync", "twmc_batch_read", "twmc_run", "twmc_set_timing", "twmc_query",
                 "twmc_reset_stats", "twmc_noisy_matrix", "twmc_advance_noalloc"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc, ctx=None, sweep=False, mlp=False):
    if rc != TWMC_OK:
        L = load()
        if mlp:
            msg = L.twmc_mlp_last_error(ctx)
        else:
            msg = L.twmc_sweep_last_error(ctx) if sweep else L.twmc_last_error(ctx)
        raise TWMCError(rc, msg.decode() if msg else "")
