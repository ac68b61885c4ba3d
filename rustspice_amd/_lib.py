"""ctypes binding of librsgpu.so (include/rsgpu.h, include/rsgpu_host.h).

The library is built in-tree by ``make -C rustspice_amd/csrc`` (or
``__graft_entry__.build()``). Importing the GPU entry points without it
raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsgpu.so")
_LIB = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def _nullable(base):
    """An ndpointer argument type that also passes None as NULL."""
    def from_param(cls, obj):
        return None if obj is None else base.from_param(obj)
    return type(base.__name__ + "_or_null", (base,), {"from_param": classmethod(from_param)})


i64p_n = _nullable(i64p)
f64p_n = _nullable(f64p)
vp = C.c_void_p

#: symbols declared in include/rsgpu.h and include/rsgpu_host.h
EXPORTS = [
    "rs_solver_create", "rs_solver_destroy", "rs_prepare", "rs_analyze", "rs_factor",
    "rs_solve", "rs_reset_pattern", "rs_solver_name", "rs_set_parallel_threads",
    "rs_set_config", "rs_get_stats", "rs_set_autotune", "rs_set_supernodal", "rs_supernodal_info", "rs_set_growth_check", "rs_growth_fallbacks",
    "rs_last_growth", "rs_set_launch_shape", "rs_launch_shape", "rs_last_error", "rs_refactor_device",
    "rs_solve_device", "rs_get_factor_values", "rs_get_factor_pattern",
    "rs_solve_device_noalloc", "rs_noalloc_max_n", "rs_error_pivot", "rs_tsolve", "rs_tsolve_device",
    "rs_host_amd_order", "rs_host_btf", "rs_host_factor_create", "rs_host_factor_destroy",
    "rs_host_factor_nblocks", "rs_host_factor_block_sizes", "rs_host_factor_block_get",
    "rs_host_factor_plan_stats", "rs_host_stamp_pattern", "rs_host_sp_selftest", "rs_host_rl_selftest", "rs_host_lp_selftest", "rs_host_sn_selftest", "rs_host_pow_dd",
    "rs_host_bbd_decompose", "rs_set_bbd_blocks", "rs_bbd_info", "rs_set_bbd_mfma", "rs_bbd_mfma_info",
    "rs_circuit_create", "rs_circuit_destroy", "rs_circuit_last_error", "rs_circuit_pattern",
    "rs_circuit_set_cap_state", "rs_circuit_set_initial_guess", "rs_circuit_set_device", "rs_stamp", "rs_dc_op", "rs_tran", "rs_tran_fixed",
    "rs_dc_sweep_points", "rs_dc_sweep", "rs_circuit_set_bsim_models", "rs_mos_eval",
    "rs_profile_enable", "rs_profile_get", "rs_set_poison",
    "rs_batch_create", "rs_batch_destroy", "rs_batch_last_error", "rs_batch_info",
    "rs_batch_dc_op", "rs_batch_results", "rs_circuit_set_device_newton",
    "rs_batch_phase_profile", "rs_batch_phase_cycles", "rs_dc_adjoint", "rs_tran_adjoint",
    "rs_circuit_set_tran_graph", "rs_circuit_graph_steps", "rs_tran_steps", "rs_stamp_device",
    "rs_solver_batch_create", "rs_solver_batch_destroy", "rs_solver_batch_last_error",
    "rs_solver_batch_layout", "rs_solver_batch_refactor", "rs_solver_batch_solve",
    "rs_solver_batch_factor_values",
    "rs_circuit_tran_graph_info", "rs_ac_frequencies", "rs_ac_create", "rs_ac_run", "rs_ac_timing", "rs_ac_last_error", "rs_ac_destroy",
]


class SolverStats(C.Structure):
    _fields_ = [
        ("factor_count", C.c_int64), ("refactor_count", C.c_int64), ("nnz_l", C.c_int64),
        ("nnz_u", C.c_int64), ("nblocks", C.c_int64), ("noffdiag", C.c_int64),
        ("rcond", C.c_double), ("parallel_factor", C.c_int64),
        ("n_factor_levels", C.c_int64), ("max_level_width", C.c_int64),
        ("n_lsolve_levels", C.c_int64), ("n_usolve_levels", C.c_int64),
        ("refactor_flops", C.c_double), ("solve_flops", C.c_double), ("plan_bytes", C.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Loads librsgpu.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C rustspice_amd/csrc` "
            "(the GPU solver has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    L.rs_solver_create.restype = vp
    L.rs_solver_create.argtypes = [C.c_int32, C.c_int64]
    L.rs_solver_destroy.argtypes = [vp]
    L.rs_prepare.argtypes = [vp, C.c_int64]
    L.rs_analyze.restype = C.c_int32
    L.rs_analyze.argtypes = [vp, i64p, C.c_int64, i64p, C.c_int64]
    L.rs_factor.restype = C.c_int32
    L.rs_factor.argtypes = [vp, i64p, C.c_int64, i64p, C.c_int64, f64p, C.c_int64]
    L.rs_solve.restype = C.c_int32
    L.rs_solve.argtypes = [vp, f64p, C.c_int64]
    L.rs_reset_pattern.argtypes = [vp]
    L.rs_solver_name.restype = C.c_char_p
    L.rs_solver_name.argtypes = [vp]
    L.rs_set_parallel_threads.argtypes = [vp, C.c_int64]
    L.rs_set_config.argtypes = [vp, C.c_double, C.c_int32, C.c_int32]
    L.rs_set_autotune.argtypes = [vp, C.c_int32]
    L.rs_set_supernodal.argtypes = [vp, C.c_int32]
    L.rs_supernodal_info.argtypes = [vp, i64p]
    L.rs_set_growth_check.restype = None
    L.rs_set_growth_check.argtypes = [vp, C.c_double]
    L.rs_growth_fallbacks.restype = C.c_int64
    L.rs_growth_fallbacks.argtypes = [vp]
    L.rs_last_growth.restype = C.c_double
    L.rs_last_growth.argtypes = [vp]
    L.rs_set_launch_shape.restype = C.c_int32
    L.rs_set_launch_shape.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.rs_launch_shape.restype = C.c_int32
    L.rs_launch_shape.argtypes = [vp, C.POINTER(C.c_int64)]
    L.rs_get_stats.restype = C.c_int32
    L.rs_get_stats.argtypes = [vp, C.POINTER(SolverStats)]
    L.rs_last_error.restype = C.c_int32
    L.rs_last_error.argtypes = [vp, C.c_char_p, C.c_int64]
    L.rs_refactor_device.restype = C.c_int32
    L.rs_refactor_device.argtypes = [vp, vp, vp]
    L.rs_solve_device.restype = C.c_int32
    L.rs_solve_device.argtypes = [vp, vp, vp, vp]
    L.rs_get_factor_values.restype = C.c_int32
    L.rs_get_factor_values.argtypes = [vp, f64p_n, f64p_n, f64p_n]
    L.rs_tsolve.restype = C.c_int32
    L.rs_tsolve.argtypes = [vp, f64p, C.c_int64]
    L.rs_tsolve_device.restype = C.c_int32
    L.rs_tsolve_device.argtypes = [vp, vp, vp, vp]
    L.rs_solver_batch_create.restype = vp
    L.rs_solver_batch_create.argtypes = [C.c_int32, C.c_int64, i64p, i64p, i64p, f64p]
    L.rs_solver_batch_destroy.argtypes = [vp]
    L.rs_solver_batch_last_error.restype = C.c_int32
    L.rs_solver_batch_last_error.argtypes = [vp, C.c_char_p, C.c_int64]
    L.rs_solver_batch_layout.restype = C.c_int32
    L.rs_solver_batch_layout.argtypes = [vp, i64p, i64p_n, i64p_n]
    L.rs_solver_batch_refactor.restype = C.c_int32
    L.rs_solver_batch_refactor.argtypes = [vp, vp, vp]
    L.rs_solver_batch_solve.restype = C.c_int32
    L.rs_solver_batch_solve.argtypes = [vp, vp, vp, vp]
    L.rs_solver_batch_factor_values.restype = C.c_int32
    L.rs_solver_batch_factor_values.argtypes = [vp, C.c_int64, f64p_n, f64p_n, f64p_n]
    L.rs_error_pivot.restype = C.c_int64
    L.rs_error_pivot.argtypes = [vp]
    L.rs_solve_device_noalloc.restype = C.c_int32
    L.rs_solve_device_noalloc.argtypes = [vp, vp, vp, vp]
    L.rs_noalloc_max_n.restype = C.c_int64
    L.rs_noalloc_max_n.argtypes = []
    L.rs_get_factor_pattern.restype = C.c_int32
    L.rs_get_factor_pattern.argtypes = [vp, i64p, i64p_n, i64p_n, i64p_n, i64p_n, i64p_n]
    L.rs_host_amd_order.argtypes = [C.c_int64, i64p, i64p, i64p, i64p]
    L.rs_host_btf.restype = C.c_int64
    L.rs_host_btf.argtypes = [C.c_int64, i64p, i64p, i64p, i64p, i64p, C.POINTER(C.c_int64)]
    L.rs_host_factor_create.restype = vp
    L.rs_host_factor_create.argtypes = [C.c_int32, C.c_int64, i64p, i64p, f64p, C.c_double,
                                        C.c_int32, C.c_int32]
    L.rs_host_factor_destroy.argtypes = [vp]
    L.rs_host_factor_nblocks.restype = C.c_int64
    L.rs_host_factor_nblocks.argtypes = [vp]
    L.rs_host_factor_block_sizes.argtypes = [vp, C.c_int64] + [C.POINTER(C.c_int64)] * 3
    L.rs_host_factor_block_get.argtypes = [vp, C.c_int64] + [vp] * 9
    L.rs_host_factor_plan_stats.argtypes = [vp, i64p]
    L.rs_host_stamp_pattern.restype = C.c_int32
    L.rs_host_stamp_pattern.argtypes = [vp, C.c_int64, C.c_int64, C.c_int32,
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64), vp, vp]
    L.rs_host_bbd_decompose.restype = C.c_int64
    L.rs_host_bbd_decompose.argtypes = [C.c_int64, i64p, i64p, C.c_int64, i64p, i64p,
                                        C.POINTER(C.c_int64)]
    L.rs_set_bbd_blocks.argtypes = [C.c_void_p, C.c_int64]
    L.rs_bbd_info.restype = C.c_int32
    L.rs_bbd_info.argtypes = [C.c_void_p, i64p]
    L.rs_set_bbd_mfma.argtypes = [C.c_void_p, C.c_int32]
    L.rs_bbd_mfma_info.restype = C.c_int32
    L.rs_bbd_mfma_info.argtypes = [C.c_void_p, i64p]
    L.rs_host_sp_selftest.restype = C.c_int32
    L.rs_host_sp_selftest.argtypes = [C.c_int32, C.c_int64, i64p, i64p, f64p, f64p, f64p, f64p, i64p]
    L.rs_host_rl_selftest.restype = C.c_int32
    L.rs_host_rl_selftest.argtypes = [C.c_int32, C.c_int64, i64p, i64p, f64p, f64p, f64p, f64p, f64p,
                                      i64p]
    L.rs_host_lp_selftest.restype = C.c_int32
    L.rs_host_lp_selftest.argtypes = [C.c_int32, C.c_int64, i64p, i64p, f64p, f64p, f64p, i64p]
    L.rs_host_sn_selftest.restype = C.c_int32
    L.rs_host_sn_selftest.argtypes = [C.c_int64, i64p, i64p, vp, vp, vp, vp, vp, vp, C.c_int64, C.c_int64, i64p]
    L.rs_profile_enable.argtypes = [C.c_int32]
    L.rs_profile_get.argtypes = [f64p, i64p]
    L.rs_set_poison.argtypes = [C.c_int32]
    _LIB = L
    return L


PROFILE_KERNELS = ("refactor", "solve", "eval", "assemble", "batch", "sn_gather", "bbd_schur")


def profile_enable(on: bool = True):
    lib().rs_profile_enable(int(on))


def set_poison(on: bool = True):
    """Fresh FP64 device buffers start as NaN while on (a checking aid: a
    value read before it is written then changes the result visibly)."""
    lib().rs_set_poison(int(on))


def profile_get() -> dict:
    ms = np.zeros(len(PROFILE_KERNELS))
    cnt = np.zeros(len(PROFILE_KERNELS), np.int64)
    lib().rs_profile_get(ms, cnt)
    return {k: dict(ms_total=float(m), count=int(c), avg_us=(1e3 * m / c if c else 0.0))
            for k, m, c in zip(PROFILE_KERNELS, ms, cnt)}
