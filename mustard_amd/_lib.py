"""ctypes binding of libmustard_hip.so (declared in include/mustard_hip.h).

The library must be built in-tree (``make -C mustard_amd`` or
``__graft_entry__.build()``); :func:`lib` raises :class:`MustardError` when it
is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmustard_hip.so")

MU_OK, MU_EINVAL, MU_EHIP, MU_ESTATE, MU_ERUN, MU_ENOTSUP = 0, -1, -2, -3, -4, -5
MODE_TILED, MODE_SUBGRAPH, MODE_STREAM = 0, 1, 2
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2


class MustardError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mustard error {code}: {msg}")
        self.code = code


class ExecOpts(ctypes.Structure):
    _fields_ = [("grid", ctypes.c_int), ("timeout_s", ctypes.c_double),
                ("trace", ctypes.c_int), ("small_max", ctypes.c_int),
                ("trsm_max", ctypes.c_int), ("lookahead", ctypes.c_int), ("sm", ctypes.c_int),
                ("no_fuse", ctypes.c_int)]


class EmuOpts(ctypes.Structure):
    _fields_ = [("workgroups", ctypes.c_int), ("timeout_s", ctypes.c_double),
                ("host_deadline_s", ctypes.c_double), ("bump_unit", ctypes.c_int),
                ("producer_delay_us", ctypes.c_int), ("sm", ctypes.c_int)]


class ExecStats(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("grid", ctypes.c_int), ("units", ctypes.c_longlong),
                ("items", ctypes.c_longlong), ("edges", ctypes.c_longlong),
                ("flops", ctypes.c_double), ("critical_path_us", ctypes.c_double),
                ("priority_units", ctypes.c_int * 4), ("wall_clock_khz", ctypes.c_double),
                ("levels", ctypes.c_int), ("depth", ctypes.c_int)]


class IROpts(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("tiles", ctypes.c_int), ("prec", ctypes.c_int),
                ("itermax", ctypes.c_int), ("fallback", ctypes.c_int)]


class IRResult(ctypes.Structure):
    _fields_ = [("info", ctypes.c_int), ("iters", ctypes.c_int), ("converged", ctypes.c_int),
                ("fell_back", ctypes.c_int), ("anorm", ctypes.c_double),
                ("berr0", ctypes.c_double), ("berr", ctypes.c_double),
                ("berr_hist", ctypes.c_double * 32), ("fact_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


SOLVE_LU, SOLVE_CHOL = 0, 1

_vp, _i, _d, _c = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_char
_sz, _u64 = ctypes.c_size_t, ctypes.c_uint64
_ip = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); restype int means a status code to check
SIGNATURES = {
    "mu_last_error": (ctypes.c_char_p, []),
    "mu_get_device_count": (_i, [_ip]),
    "mu_set_device": (_i, [_i]),
    "mu_device_props": (_i, [_i, _ip, _ip, ctypes.POINTER(_sz), ctypes.c_char_p, _i]),
    "mu_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
    "mu_free": (_i, [_vp]),
    "mu_memcpy": (_i, [_vp, _vp, _sz]),
    "mu_memcpy_async": (_i, [_vp, _vp, _sz, _vp]),
    "mu_memset_zero": (_i, [_vp, _sz]),
    "mu_stream_create": (_i, [ctypes.POINTER(_vp)]),
    "mu_stream_destroy": (_i, [_vp]),
    "mu_stream_synchronize": (_i, [_vp]),
    "mu_device_synchronize": (_i, []),
    "mu_event_create": (_i, [ctypes.POINTER(_vp)]),
    "mu_event_destroy": (_i, [_vp]),
    "mu_event_record": (_i, [_vp, _vp]),
    "mu_event_elapsed_ms": (_i, [ctypes.POINTER(ctypes.c_float), _vp, _vp]),
    "mu_mem_info": (_i, [ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
    "mu_dgemm": (_i, [_vp, _c, _c, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]),
    "mu_gemm_a32": (_i, [_vp, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i, _i]),
    "mu_dsyrk": (_i, [_vp, _c, _c, _i, _i, _d, _vp, _i, _d, _vp, _i]),
    "mu_dtrsm": (_i, [_vp, _c, _c, _c, _c, _i, _i, _d, _vp, _i, _vp, _i]),
    "mu_dtrsm_noalloc": (_i, [_vp, _c, _c, _c, _c, _i, _i, _d, _vp, _i, _vp, _i]),
    "mu_dgetrf_nopiv": (_i, [_vp, _i, _vp, _i, _vp]),
    "mu_dpotrf": (_i, [_vp, _c, _i, _vp, _i, _vp]),
    "mu_info_reset": (_i, [_vp, _vp]),
    "mu_set_update_precision": (_i, [_vp, _i]),
    "mu_dag_create": (_i, [ctypes.POINTER(_vp)]),
    "mu_dag_destroy": (_i, [_vp]),
    "mu_dag_capture_begin": (_i, [_vp, _vp]),
    "mu_dag_capture_end": (_i, [_vp, _vp]),
    "mu_dag_begin_task": (_i, [_vp, _i, _i, _ip, _i, _ip]),
    "mu_dag_end_task": (_i, [_vp]),
    "mu_dag_num_tasks": (_i, [_vp, _ip]),
    "mu_dag_task_deps": (_i, [_vp, _i, _ip, _i, _ip]),
    "mu_dag_add_edge": (_i, [_vp, _i, _i]),
    "mu_dag_insert_subdag": (_i, [_vp, _vp]),
    "mu_dag_set_tag": (_i, [_vp, _i]),
    "mu_dag_task_info": (_i, [_vp, _i, _ip, _ip, _ip, _i, _ip]),
    "mu_exec_create": (_i, [_vp, _i, ctypes.POINTER(ExecOpts), ctypes.POINTER(_vp)]),
    "mu_exec_launch": (_i, [_vp, _vp]),
    "mu_exec_check": (_i, [_vp]),
    "mu_exec_launch_level": (_i, [_vp, _i, _vp]),
    "mu_exec_level_items": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "mu_exec_level_ops": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "mu_exec_level_info": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_double)]),
    "mu_exec_stats": (_i, [_vp, ctypes.POINTER(ExecStats)]),
    "mu_exec_trace": (_i, [_vp, ctypes.POINTER(ctypes.c_ulonglong), _sz]),
    "mu_exec_item_units": (_i, [_vp, _ip, _ip, _ip]),
    "mu_exec_destroy": (_i, [_vp]),
    "mu_record_tiled_lu": (_i, [_vp, _vp, _i, _i, _i, _vp, _i]),
    "mu_record_tiled_cholesky": (_i, [_vp, _vp, _i, _i, _i, _vp, _i]),
    "mu_gen_ref_host": (None, [_vp, _sz]),
    "mu_gen_uniform": (_i, [_vp, _sz, _u64, _d, _d, _vp]),
    "mu_gen_spd": (_i, [_vp, _i, _i, _u64, _vp, _vp]),
    "mu_gen_cond": (_i, [_vp, _i, _i, _d, _i, _u64, _vp, _vp]),
    "mu_chol_backward": (_i, [_vp, _i, _vp, _i, _vp, _i, _i, _i, _i, _vp]),
    "mu_record_batch": (_i, [_vp, _i, _ip, _ip, ctypes.POINTER(_vp), _ip, _ip, _vp, _i]),
    "mu_batch_residual": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mu_record_tiled_solve": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _i, _i]),
    "mu_solve_ir": (_i, [_vp, _i, _i, _vp, _i, _vp, _i, _i, ctypes.POINTER(IROpts),
                         ctypes.POINTER(IRResult), _vp]),
    "mu_verify_lu_l1_host": (_d, [_vp, _vp, _i]),
    "mu_verify_chol_l1_host": (_d, [_vp, _vp, _i]),
    "mu_scaled_residual": (_i, [_vp, _vp, _i, _i, _i, _vp, ctypes.POINTER(_d), _vp]),
    "mu_residual_norms": (_i, [_vp, _vp, _i, _i, _i, _vp, ctypes.POINTER(_d), ctypes.POINTER(_d),
                               ctypes.POINTER(_d), _vp]),
    "mu_dag_emulate": (_i, [_vp, _i, _d, _i, ctypes.POINTER(ctypes.c_longlong)]),
    "mu_dag_emulate_ex": (_i, [_vp, ctypes.POINTER(EmuOpts), ctypes.POINTER(ctypes.c_longlong)]),
}

_LIB = None


def lib():
    """The loaded library; raises MustardError if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MustardError(MU_EINVAL, f"{LIB_PATH} is missing: build it with "
                                      "`make -C mustard_amd` (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(rc):
    if rc != MU_OK:
        raise MustardError(rc, lib().mu_last_error().decode(errors="replace"))
    return rc


def call(name, *args):
    """Call a status-returning entry point and raise on failure."""
    return check(getattr(lib(), name)(*args))


def exported_symbols():
    return list(SIGNATURES)
