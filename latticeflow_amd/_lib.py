"""ctypes binding of liblatticeflow_amd.so (the C ABI in include/latticeflow_amd.h).

The library must be built (``python -m latticeflow_amd.build`` or
``__graft_entry__.build()``); importing this module never falls back to any
CPU implementation -- a missing library raises ImportError.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# LATTICEFLOW_AMD_LIBRARY selects another build of the library (same ABI)
LIB_PATH = os.environ.get("LATTICEFLOW_AMD_LIBRARY") or os.path.join(_HERE, "liblatticeflow_amd.so")

LF_OK = 0
LF_ERR_BAD_NAME = -1
LF_ERR_BAD_ARG = -2
LF_ERR_NONSUBMODULAR = -3
LF_ERR_OVERFLOW = -4
LF_ERR_DEVICE = -5
LF_ERR_STATE = -6
LF_ERR_UNSUPPORTED = -7
LF_ERR_OUT_OF_RANGE = -8

LF_CUT_SMIN = 0
LF_CUT_SMAX = 1


class BadSolverOrKernelName(ValueError):
    """Mirror of latticeQBP::BadSolverOrKernelName (solver/interface.hpp:152-168)."""


class LatticeFlowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class lf_tuning(ctypes.Structure):
    _fields_ = [("rounds_per_relabel", ctypes.c_int32), ("local_iters", ctypes.c_int32),
                ("max_cycles", ctypes.c_int32), ("tile_shape", ctypes.c_int32),
                ("relabel_launches", ctypes.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class lf_stats(ctypes.Structure):
    _fields_ = [("cycles", ctypes.c_int64), ("rounds", ctypes.c_int64),
                ("bfs_launches", ctypes.c_int64), ("flood_launches", ctypes.c_int64),
                ("tiles_pushed", ctypes.c_int64), ("tile_nodes", ctypes.c_int64),
                ("bytes_per_node", ctypes.c_int64), ("reserved", ctypes.c_int64),
                ("seconds", ctypes.c_double), ("push_seconds", ctypes.c_double),
                ("bfs_seconds", ctypes.c_double), ("flood_seconds", ctypes.c_double),
                ("bfs_visits", ctypes.c_int64), ("bfs_bytes_per_visit", ctypes.c_int64),
                ("bfs_device_seconds", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class lf_guided_stats(ctypes.Structure):
    _fields_ = [("regions", ctypes.c_int64), ("repairs", ctypes.c_int64), ("failing_arcs", ctypes.c_int64),
                ("layers_merged", ctypes.c_int64), ("fallback", ctypes.c_int64),
                ("partition_seconds", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class lf_tvpath_stats(ctypes.Structure):
    _fields_ = [("initial_regions", ctypes.c_int64), ("segments", ctypes.c_int64), ("joins", ctypes.c_int64),
                ("splits", ctypes.c_int64), ("checks", ctypes.c_int64), ("host_checks", ctypes.c_int64),
                ("device_calls", ctypes.c_int64), ("dinkelbach_steps", ctypes.c_int64),
                ("initial_seconds", ctypes.c_double), ("device_check_seconds", ctypes.c_double),
                ("host_check_seconds", ctypes.c_double), ("seconds", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# every symbol the header declares, with its ctypes signature
_I = ctypes.c_int
_I64 = ctypes.c_int64
_P = ctypes.c_void_p
_CP = ctypes.c_char_p
_SIG = {
    "lf_last_error": (_CP, []),
    "lf_is_valid_solver": (_I, [_CP]),
    "lf_is_valid_kernel": (_I, [_CP]),
    "lf_kernel_dimension": (_I, [_CP]),
    "lf_valid_kernel_names": (_I, [_CP, ctypes.c_size_t]),
    "lf_valid_solver_names": (_I, [_CP, ctypes.c_size_t]),
    "lf_kernel_size": (_I, [_CP]),
    "lf_kernel_offsets": (_I, [_CP, _P, _P]),
    "lf_create": (_I, [ctypes.POINTER(_P), _CP, _CP, _I, _P, _I]),
    "lf_destroy": (_I, [_P]),
    "lf_n_dimensions": (_I, [_P]),
    "lf_n_nodes": (_I64, [_P]),
    "lf_n_filling_edges": (_I, [_P]),
    "lf_edge_deltas": (_I, [_P, _P]),
    "lf_directions": (_I, [_P, _P]),
    "lf_add_energy_potentials": (_I, [_P, _P, _P]),
    "lf_add_energy_potentials_device": (_I, [_P, _P, _P, _P]),
    "lf_reset_energies": (_I, [_P]),
    "lf_set_tuning": (_I, [_P, ctypes.POINTER(lf_tuning)]),
    "lf_get_tuning": (_I, [_P, ctypes.POINTER(lf_tuning)]),
    "lf_autotune": (_I, [_P, _I, ctypes.POINTER(lf_tuning)]),
    "lf_run": (_I, [_P]),
    "lf_set_debug_checks": (_I, [_P, _I]),
    "lf_verify_state": (_I, [_P, _I, _P, _P, _P, _I, _P, _P, _P, _I, _P, _I, _P, _I]),
    "lf_run_time": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "lf_get_stats": (_I, [_P, ctypes.POINTER(lf_stats)]),
    "lf_get_cut": (_I, [_P, _P, _I]),
    "lf_get_flow": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "lf_get_cut_device": (_I, [_P, _P, _I, _P]),
    "lf_get_levels": (_I, [_P, _P]),
    "lf_batch_create": (_I, [ctypes.POINTER(_P), _CP, _I, _P, _I]),
    "lf_batch_size": (_I, [_P]),
    "lf_batch_canvas": (_I, [_P, _P]),
    "lf_batch_edge_deltas": (_I, [_P, _I, _P]),
    "lf_batch_add_energy_potentials": (_I, [_P, _I, _P, _P]),
    "lf_batch_add_energy_potentials_device": (_I, [_P, _I, _P, _P, _P]),
    "lf_batch_get_flows": (_I, [_P, _P]),
    "lf_tvnet_param_count": (_I, []),
    "lf_tvnet_work_bytes": (ctypes.c_size_t, [_I, _I64, _I64]),
    "lf_tvnet_forward": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    "lf_tvnet_forward_noalloc": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P]),
    "lf_tvnet_gradients": (_I, [_P, _I, _I64, _I64, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    "lf_tvnet_adam": (_I, [_P, _P, _P, _P, _I, ctypes.c_float, _P]),
    "lf_tv_single": (_I, [_CP, _I64, _I64, _P, ctypes.c_double, _P, ctypes.POINTER(lf_stats)]),
    "lf_tv_path": (_I, [_CP, _I64, _I64, _P, _I64, _P, _P, ctypes.POINTER(lf_stats)]),
    "lf_tv_path_guided": (_I, [_CP, _I64, _I64, _P, _I64, _P, _P, ctypes.c_double, _I, _P,
                               ctypes.POINTER(lf_stats), ctypes.POINTER(lf_guided_stats), _P]),
    "lf_tvpath_create": (_I, [ctypes.POINTER(_P), _CP, _I64, _I64, _P, ctypes.c_double, _I64]),
    "lf_tvpath_eval": (_I, [_P, _I64, _P, _P]),
    "lf_tvpath_get_stats": (_I, [_P, ctypes.POINTER(lf_tvpath_stats)]),
    "lf_tvpath_destroy": (_I, [_P]),
}

_lib = None


def lib():
    """Load the library (once). Raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("liblatticeflow_amd.so is not built; run __graft_entry__.build()")
        # PyTorch (device memory and streams for the Python host) ships its own
        # HIP runtime; load and initialise it first so that the library and
        # PyTorch share one runtime in this process.
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return sorted(_SIG)


def check(rc):
    """Raise the Python counterpart of a non-zero LF_ERR_* code."""
    if rc == LF_OK or rc > 0:
        return rc
    msg = lib().lf_last_error().decode(errors="replace")
    if rc == LF_ERR_BAD_NAME:
        raise BadSolverOrKernelName(msg)
    if rc in (LF_ERR_BAD_ARG, LF_ERR_NONSUBMODULAR, LF_ERR_OVERFLOW):
        raise ValueError(msg)
    if rc == LF_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    raise LatticeFlowError(rc, msg)
