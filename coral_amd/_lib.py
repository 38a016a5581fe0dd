"""ctypes binding of ``include/coral_amd.h``.

The shared library is the product's only compute path: if it is missing the
import of this module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CORAL_AMD_LIB", os.path.join(HERE, "lib", "libcoral_amd.so"))

c_int32 = ctypes.c_int32
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p


class Piece(ctypes.Structure):
    _fields_ = [("dsca", c_double), ("kx", c_int32), ("ky", c_int32), ("iz", c_int32), ("mz", c_int32),
                ("var", c_int32), ("eqn", c_int32)]


class NLTerm(ctypes.Structure):
    _fields_ = [("dsca", c_double), ("kx", c_int32), ("ky", c_int32), ("iz", c_int32), ("mz", c_int32),
                ("qvar", c_int32), ("eqn", c_int32)]


class System(ctypes.Structure):
    _fields_ = [("nvars", c_int32), ("bc_codes", ctypes.POINTER(c_int32)), ("eqn_orders", ctypes.POINTER(c_int32)),
                ("n_mass", c_int32), ("mass", ctypes.POINTER(Piece)),
                ("n_stif", c_int32), ("stif", ctypes.POINTER(Piece)),
                ("n_nl", c_int32), ("nl", ctypes.POINTER(NLTerm))]


class FullTerm(ctypes.Structure):
    _fields_ = [("dsca", c_double), ("kx", c_int32), ("ky", c_int32), ("dz", c_int32), ("mz", c_int32),
                ("kind", c_int32), ("system", c_int32), ("position", c_int32)]


class FullVar(ctypes.Structure):
    _fields_ = [("nterms", c_int32), ("terms", ctypes.POINTER(FullTerm))]


class QuadVar(ctypes.Structure):
    _fields_ = [("ivar1", c_int32), ("ivar2", c_int32), ("remove_z_integral", c_int32)]


class ClosureDesc(ctypes.Structure):
    _fields_ = [("width", c_int32), ("hidden", c_int32), ("layers", c_int32), ("modes", c_int32),
                ("weights", ctypes.POINTER(ctypes.c_float)), ("nweights", ctypes.c_int64),
                ("in_mean", ctypes.POINTER(ctypes.c_float)), ("in_std", ctypes.POINTER(ctypes.c_float)),
                ("out_mean", ctypes.POINTER(ctypes.c_float)), ("out_std", ctypes.POINTER(ctypes.c_float))]


class Problem(ctypes.Structure):
    _fields_ = [("NX", c_int32), ("NY", c_int32), ("NZ", c_int32),
                ("Lx", c_double), ("Ly", c_double), ("gap", c_double), ("center", c_double),
                ("nstages", c_int32), ("a_arr", ctypes.POINTER(c_double)), ("a_hat", ctypes.POINTER(c_double)),
                ("n_kxky", c_int32), ("kxky", ctypes.POINTER(System)),
                ("n_zero", c_int32), ("zero", ctypes.POINTER(System)),
                ("n_full", c_int32), ("full", ctypes.POINTER(FullVar)),
                ("n_quad", c_int32), ("quad", ctypes.POINTER(QuadVar))]


EXPORTS = (
    "coral_amd_create", "coral_amd_destroy", "coral_amd_last_error", "coral_amd_set_dt",
    "coral_amd_upload_state", "coral_amd_download_state", "coral_amd_step", "coral_amd_synchronize",
    "coral_amd_get_physical", "coral_amd_get_cfl_dt", "coral_amd_wavenumbers", "coral_amd_system_info",
    "coral_amd_bwd_transform", "coral_amd_fwd_transform", "coral_amd_banded_solve",
    "coral_amd_timers_enable", "coral_amd_timer_read", "coral_amd_assembled_operator", "coral_amd_use_graphs", "coral_amd_products",
    "coral_amd_factor_kind", "coral_amd_get_variable_physical", "coral_amd_set_variable_physical",
    "coral_amd_horizontal_averages", "coral_amd_factor_blocks", "coral_amd_system_solve",
    "coral_amd_set_precision", "coral_amd_set_kx0_projection", "coral_amd_banded_solve_noalloc",
    "coral_amd_banded_solve_noalloc_bytes", "coral_amd_set_launch_shape", "coral_amd_get_launch_shape",
    "coral_amd_ensemble_create", "coral_amd_ensemble_destroy", "coral_amd_ensemble_set_schedule",
    "coral_amd_ensemble_tune", "coral_amd_ensemble_step", "coral_amd_ensemble_synchronize",
    "coral_amd_set_closure", "coral_amd_nn_gemm", "coral_amd_nn_cast_bf16", "coral_amd_nn_bias_act", "coral_amd_nn_gelu_bwd",
    "coral_amd_nn_colsum_work", "coral_amd_nn_colsum", "coral_amd_nn_mode_scale", "coral_amd_nn_mode_scale_grad",
    "coral_amd_nn_axpy", "coral_amd_nn_mse_partials", "coral_amd_nn_mse", "coral_amd_nn_adam",
    "coral_amd_nn_gemm_work", "coral_amd_nn_transpose_bf16", "coral_amd_nn_mlp",
    "coral_amd_set_full_variable_filter", "coral_amd_add_source_term", "coral_amd_clear_source_terms",
    "coral_amd_set_source", "coral_amd_set_fixed_physical", "coral_amd_set_smagorinsky", "coral_amd_nn_mode_scale_t", "coral_amd_nn_modes_supported",
    "coral_amd_nn_modes_work", "coral_amd_nn_modes", "coral_amd_nn_tile_transpose",
    "coral_amd_nn_gemm_bias",
)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"coral_amd native library not found at {LIB_PATH}; "
                          "build it with `python -m coral_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    PP = ctypes.POINTER(c_void_p)
    lib.coral_amd_create.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, PP]
    lib.coral_amd_destroy.argtypes = [P]
    lib.coral_amd_last_error.argtypes = [ctypes.c_char_p, c_int32]
    lib.coral_amd_set_dt.argtypes = [P, c_double]
    lib.coral_amd_upload_state.argtypes = [P, c_int32, c_int32, P]
    lib.coral_amd_download_state.argtypes = [P, c_int32, c_int32, P]
    lib.coral_amd_step.argtypes = [P, c_int32]
    lib.coral_amd_synchronize.argtypes = [P]
    lib.coral_amd_get_physical.argtypes = [P, c_int32, P]
    lib.coral_amd_get_cfl_dt.argtypes = [P, ctypes.POINTER(c_double)]
    lib.coral_amd_wavenumbers.argtypes = [P, P, P, P, P]
    lib.coral_amd_system_info.argtypes = [P, c_int32, P, P, P, P]
    lib.coral_amd_factor_blocks.argtypes = [P, c_int32, c_int32, P, P, P, P, P]
    lib.coral_amd_bwd_transform.argtypes = [P, P, P]
    lib.coral_amd_fwd_transform.argtypes = [P, P, P]
    lib.coral_amd_system_solve.argtypes = [P, c_int32, P, P]
    lib.coral_amd_set_precision.argtypes = [P, c_int32, c_int32]
    lib.coral_amd_banded_solve.argtypes = [c_int32, c_int32, c_int32, c_int32, c_int32, P, P, P]
    lib.coral_amd_timers_enable.argtypes = [P, c_int32]
    lib.coral_amd_use_graphs.argtypes = [P, c_int32]
    lib.coral_amd_set_kx0_projection.argtypes = [P, c_int32]
    lib.coral_amd_products.argtypes = [P, P]
    lib.coral_amd_timer_read.argtypes = [P, c_int32, ctypes.c_char_p, c_int32, ctypes.POINTER(c_double),
                                         ctypes.POINTER(c_int32)]
    lib.coral_amd_assembled_operator.argtypes = [ctypes.POINTER(Problem), c_int32, c_int32, c_int32, c_int32, P,
                                                 P, P, P]
    lib.coral_amd_banded_solve_noalloc.argtypes = [c_int32, c_int32, c_int32, c_int32, c_int32, P, P, P, P]
    lib.coral_amd_banded_solve_noalloc_bytes.argtypes = [c_int32, c_int32, c_int32, c_int32]
    lib.coral_amd_set_launch_shape.argtypes = [P, c_int32, c_int32]
    lib.coral_amd_set_full_variable_filter.argtypes = [P, c_int32, c_int32, c_int32]
    lib.coral_amd_add_source_term.argtypes = [P, c_int32, c_int32, c_int32, c_double, c_int32]
    lib.coral_amd_clear_source_terms.argtypes = [P]
    lib.coral_amd_set_source.argtypes = [P, c_int32, P]
    lib.coral_amd_set_fixed_physical.argtypes = [P, c_int32, P]
    lib.coral_amd_set_smagorinsky.argtypes = [P, c_int32, P, c_int32, c_double]
    lib.coral_amd_set_closure.argtypes = [P, ctypes.POINTER(ClosureDesc)]
    lib.coral_amd_get_launch_shape.argtypes = [P, P, P]
    lib.coral_amd_ensemble_create.argtypes = [P, c_int32, PP]
    lib.coral_amd_ensemble_destroy.argtypes = [P]
    lib.coral_amd_ensemble_set_schedule.argtypes = [P, c_int32, P]
    lib.coral_amd_ensemble_tune.argtypes = [P, c_int32, P, P]
    lib.coral_amd_ensemble_step.argtypes = [P, c_int32]
    lib.coral_amd_ensemble_synchronize.argtypes = [P]
    for name in EXPORTS:
        getattr(lib, name).restype = ctypes.c_int
    lib.coral_amd_banded_solve_noalloc_bytes.restype = ctypes.c_size_t
    lib.coral_amd_nn_colsum_work.restype = ctypes.c_int64
    lib.coral_amd_nn_mse_partials.restype = ctypes.c_int64
    lib.coral_amd_nn_gemm_work.restype = ctypes.c_int64
    lib.coral_amd_nn_modes_work.restype = ctypes.c_int64
    lib.coral_amd_nn_modes_supported.restype = ctypes.c_int32
    return lib


lib = _load()


class CoralError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        buf = ctypes.create_string_buffer(1024)
        lib.coral_amd_last_error(buf, 1024)
        raise CoralError(f"{what}: {buf.value.decode(errors='replace')}" if what else buf.value.decode())
    return rc
