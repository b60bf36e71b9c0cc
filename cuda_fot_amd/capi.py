"""ctypes binding of libfot's C-ABI (include/fot_capi.h).

The library is the product: every call below goes to HIP kernels or to the
native host planner.  There is no Python or CPU fallback: if libfot.so is
missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

MAX_MAPPINGS = 8
MAX_CELL_DATA = 8

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfot.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fot_capi.h")

# status codes (fot_status)
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_OUT_OF_MEMORY = 3
ERR_COLOUR_OVERFLOW = 4
ERR_UNSUPPORTED = 5
ERR_INVALID_INPUT = 6
ERR_NO_DEVICE = 7
ERR_TIMEOUT = 8

# fot_kernel
KERNEL_RES_CALC = 0
KERNEL_MINE = 1
KERNEL_HEX_SCATTER = 2
KERNEL_MINI_AERO = 3
KERNEL_VOLNA = 4
KERNEL_GETACC = 5
KERNEL_LULESH = 6


def cell_output_slot(kernel: int) -> int:
    """The cell data slot a kernel writes (lulesh's element volume), or -1."""
    return 1 if kernel == KERNEL_LULESH else -1

# fot_strategy
STRATEGY_ATOMIC = 0
STRATEGY_GLOBAL_COLOUR = 1
STRATEGY_HIERARCHICAL = 2

LOOP_STAGE_INPUTS = 1
LOOP_DEPENDENCY_ORDER = 2
LOOP_OWNER_COMPUTE = 4
LOOP_POINT_PARTITION = 8
LOOP_TWO_PER_THREAD = 16
LOOP_GATHER_NODES = 32
LOOP_POINT_TERMS = 64
LOOP_OWNER_GATHER = 128
LOOP_PERSISTENT = 256
LOOP_LISTS_IN_HBM = 512
LOOP_TWO_PHASE = 1024
LOOP_NONTEMPORAL = 2048
LOOP_CACHED = 4096


class FotError(RuntimeError):
    """A libfot call returned a non-zero fot_status."""

    def __init__(self, status: int, func: str, message: str):
        self.status = status
        super().__init__(f"{func}: status {status}: {message}")


class InvalidInputFile(FotError):
    """Mirror of the reference's InvalidInputFile (mesh.hpp:18-21)."""


class ColourOverflow(FotError):
    """More colours than the reference's colour sets hold (mesh.hpp:25-29)."""


class Functor(C.Structure):
    _fields_ = [("kernel", C.c_uint32), ("mesh_dim", C.c_uint32),
                ("point_dim", C.c_uint32), ("cell_dim", C.c_uint32)]


class ProblemDesc(C.Structure):
    _fields_ = [
        ("num_cells", C.c_uint32),
        ("num_mappings", C.c_uint32),
        ("cell_to_node", C.c_void_p * MAX_MAPPINGS),
        ("mesh_dim", C.c_uint32 * MAX_MAPPINGS),
        ("num_points", C.c_uint32 * MAX_MAPPINGS),
        ("point_data", C.c_void_p * MAX_MAPPINGS),
        ("point_dim", C.c_uint32 * MAX_MAPPINGS),
        ("cell_data", C.c_void_p * MAX_CELL_DATA),
        ("cell_dim", C.c_uint32 * MAX_CELL_DATA),
        ("num_cell_data", C.c_uint32),
        ("point_data_out", C.c_void_p),
        ("soa", C.c_int),
    ]


class LoopConfig(C.Structure):
    _fields_ = [("strategy", C.c_uint32), ("block_size", C.c_uint32),
                ("partition", C.c_void_p), ("use_graph", C.c_uint32),
                ("flags", C.c_uint32)]


class LoopStats(C.Structure):
    _fields_ = [("strategy", C.c_uint32), ("num_launches_per_iter", C.c_uint32),
                ("num_colours", C.c_uint32), ("num_blocks", C.c_uint32),
                ("block_size", C.c_uint32), ("max_cached", C.c_uint32),
                ("lds_bytes", C.c_uint32), ("max_thread_colours", C.c_uint32),
                ("plan_device_bytes", C.c_uint64),
                ("avg_thread_colours", C.c_double), ("reuse_factor", C.c_double),
                ("redundancy", C.c_double), ("avg_cache_lines", C.c_double),
                ("threads_per_block", C.c_uint32), ("stage_per_thread", C.c_uint32),
                ("launch_grid", C.c_uint32), ("nontemporal", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("num_problems", C.c_uint32), ("num_work_items", C.c_uint32),
                ("grid", C.c_uint32), ("threads", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("uses_graph", C.c_uint32)]


class OplanView(C.Structure):
    _fields_ = [("mesh_dim", C.c_uint32), ("num_blocks", C.c_uint32),
                ("num_positions", C.c_uint32), ("num_cached", C.c_uint32),
                ("blk_cell_off", C.POINTER(C.c_uint32)), ("blk_cache_off", C.POINTER(C.c_uint32)),
                ("blk_owned", C.POINTER(C.c_uint32)), ("cached", C.POINTER(C.c_uint32)),
                ("cells", C.POINTER(C.c_uint32)), ("loc", C.POINTER(C.c_uint32)),
                ("colour", C.POINTER(C.c_uint8)), ("blk_ncol", C.POINTER(C.c_uint8))]


class HplanInfo(C.Structure):
    _fields_ = [("num_colours", C.c_uint32), ("num_blocks", C.c_uint32),
                ("num_cells", C.c_uint32), ("num_cached", C.c_uint32),
                ("max_cached", C.c_uint32), ("max_block", C.c_uint32),
                ("max_thread_colours", C.c_uint32),
                ("avg_thread_colours", C.c_double)]


class HplanColour(C.Structure):
    _fields_ = [
        ("num_blocks", C.c_uint32), ("num_cells", C.c_uint32),
        ("num_cached", C.c_uint32), ("shared_size", C.c_uint32),
        ("points_to_be_cached", C.POINTER(C.c_uint32)),
        ("points_to_be_cached_offsets", C.POINTER(C.c_uint32)),
        ("cell_list", C.POINTER(C.c_uint32) * MAX_MAPPINGS),
        ("cell_colours", C.POINTER(C.c_uint8)),
        ("num_cell_colours", C.POINTER(C.c_uint8)),
        ("block_offsets", C.POINTER(C.c_uint32)),
        ("cell_inv_permutation", C.POINTER(C.c_uint32)),
    ]


P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int
D = C.c_double
F32 = C.c_float

# name -> argtypes; every entry point of include/fot_capi.h
_PROTOS = {
    "fot_status_string": [I32],
    "fot_last_error_message": [],
    "fot_version": [],
    "fot_device_count": [P],
    "fot_set_device": [I32],
    "fot_malloc": [P, C.c_size_t],
    "fot_free": [P],
    "fot_memcpy_htod": [P, P, C.c_size_t],
    "fot_memcpy_dtoh": [P, P, C.c_size_t],
    "fot_memcpy_dtod": [P, P, C.c_size_t, P],
    "fot_device_synchronize": [],
    "fot_stream_create": [P],
    "fot_stream_destroy": [P],
    "fot_stream_synchronize": [P],
    "fot_copy": [P, P, C.c_size_t, P],
    "fot_timer_create": [P],
    "fot_timer_destroy": [P],
    "fot_timer_start": [P, P],
    "fot_timer_stop": [P, P],
    "fot_timer_elapsed_ms": [P, P],
    "fot_functor_init": [P, I32, U32, U32, U32],
    "fot_step_gpu_global": [P, I32, P, P, P, P, U32, P, U32, U32, U32, P],
    "fot_step_gpu_atomic": [P, I32, P, P, P, P, U32, P, U32, U32, U32, P],
    "fot_step_gpu_hierarchical": [P, I32, P, P, P, P, P, P, P, P, P, U32, P, U32,
                                  U32, U32, U32, P],
    "fot_lds_overflow_flag": [P, I32],
    "fot_colour_cells": [P, U32, U32, U32, P, P],
    "fot_colour_cells_with_ordering": [P, U32, U32, P, P],
    "fot_hplan_build": [P, P, U32, U32, U32, P, U32, P],
    "fot_hplan_free": [P],
    "fot_hplan_info_get": [P, P],
    "fot_hplan_colour_get": [P, U32, P],
    "fot_oplan_build": [P, U32, U32, U32, P, U32, P, P],
    "fot_oplan_get": [P, P],
    "fot_oplan_free": [P],
    "fot_count_cache_lines": [P, U32, U32, U32, U32, C.c_int, P],
    "fot_reorder_to_partition": [P, U32, P],
    "fot_point_renumbering_by_partition": [P, U32, U32, U32, P, P],
    "fot_point_renumbering_first_touch": [P, U32, U32, U32, P],
    "fot_mesh_reorder_cells": [P, U32, U32, P, P],
    "fot_graph_csr": [P, U32, U32, U32, P, P, P],
    "fot_cell_to_cell_csr": [P, P, U32, U32, P, P, P, P],
    "fot_heuristic_partition": [P, P, U32, P, U32, P, P],
    "fot_kway_partition": [P, P, U32, U32, P, P],
    "fot_bisect_partition": [P, P, U32, U32, P, P],
    "fot_rcm_order": [P, P, U32, P, P],
    "fot_gps_order": [P, P, U32, P],
    "fot_gen_grid_edges": [U32, U32, P],
    "fot_gen_hex_grid": [U32, U32, U32, P],
    "fot_gen_airfoil": [U32, U32, U64, U32, P, P, P, P, P],
    "fot_gen_permutation": [U32, U64, P],
    "fot_gen_uniform": [U64, U64, D, D, P],
    "fot_gen_normal": [U64, U64, D, D, P],
    "fot_gen_tri_mesh": [U32, U32, U64, U32, P, P, P, P, P, P, P],
    "fot_gen_hex_faces": [U32, U32, U32, U64, U32, P, P, P],
    "fot_gen_quad_grid": [U32, U32, P],
    "fot_mini_aero_face_flux": [P, P, U32, P, U32, P],
    "fot_lulesh_element": [P, P, P, P, U32, P, P],
    "fot_loop_create": [P, P, P, P],
    "fot_loop_destroy": [P],
    "fot_loop_step": [P, P],
    "fot_loop_run": [P, U32, P],
    "fot_loop_check": [P],
    "fot_loop_stats_get": [P, P],
    "fot_batch_create": [P, U32, P, P, U32, P],
    "fot_batch_destroy": [P],
    "fot_batch_step": [P, P],
    "fot_batch_run": [P, U32, P],
    "fot_batch_stats_get": [P, P],
    "fot_ctx_create": [P, C.c_int],
    "fot_ctx_destroy": [P],
    "fot_ctx_stream": [P, P],
    "fot_ctx_synchronize": [P],
    "fot_buf_alloc": [P, C.c_size_t, P],
    "fot_buf_upload": [P, P, P, C.c_size_t],
    "fot_buf_download": [P, P, P, C.c_size_t],
    "fot_buf_free": [P, P],
    "fot_plan_upload": [P, P, P, P, P],
    "fot_plan_free": [P],
    "fot_loop_launch": [P, P, P, P, U32],
    "fot_graph_capture": [P, P, P],
    "fot_graph_launch": [P, P],
    "fot_graph_free": [P],
    "fot_batch_launch": [P, P, U32],
    "fot_loop_features": [P, P, P],
    "fot_autotune_measure": [P, P, P, U32, U32, U32, P],
    "fot_mlp_create": [U32, U32, U32, U64, P],
    "fot_mlp_destroy": [P],
    "fot_mlp_num_params": [P, P],
    "fot_mlp_get_params": [P, P, U64],
    "fot_mlp_set_params": [P, P, U64],
    "fot_mlp_forward": [P, P, U32, P],
    "fot_mlp_forward_noalloc": [P, P, U32, P, P],
    "fot_mlp_loss_grad": [P, P, P, P, U32, P, P],
    "fot_mlp_train": [P, P, P, P, U32, U32, F32, F32, P],
    "fot_gemm_f32": [P, P, P, U32, U32, U32, C.c_int, C.c_int, P],
}

_RESTYPES = {"fot_status_string": C.c_char_p, "fot_last_error_message": C.c_char_p,
             "fot_version": I32, "fot_hplan_free": None, "fot_oplan_free": None}


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libfot.so not built at {LIB_PATH}; run `python -c 'import __graft_entry__ as g; "
            "g.build()'` or `make -C cuda_fot_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, argtypes in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, I32)
    return lib


lib = _load()


def declared_symbols() -> list[str]:
    """Function names declared in include/fot_capi.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fot_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def last_error() -> str:
    return lib.fot_last_error_message().decode(errors="replace")


def check(status: int, func: str = "libfot") -> None:
    if status == OK:
        return
    msg = last_error()
    if status == ERR_INVALID_INPUT:
        raise InvalidInputFile(status, func, msg)
    if status == ERR_COLOUR_OVERFLOW:
        raise ColourOverflow(status, func, msg)
    raise FotError(status, func, msg)


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)


def ptr(a: np.ndarray | None) -> int | None:
    """Host address of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    n = C.c_int(0)
    st = lib.fot_device_count(C.byref(n))
    return n.value if st == OK else 0


def functor(kernel: int, mesh_dim: int = 0, point_dim: int = 0, cell_dim: int = 0) -> Functor:
    f = Functor()
    call("fot_functor_init", C.byref(f), kernel, mesh_dim, point_dim, cell_dim)
    return f
