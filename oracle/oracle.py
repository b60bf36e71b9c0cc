"""ctypes wrapper of the CPU oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")

RES_CALC, MINE2, MINE4, MINE8, HEX_SCATTER, MINI_AERO, VOLNA, GETACC, LULESH = range(9)
MAXM, MAXCD = 8, 8


def build() -> None:
    subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


class Problem(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("soa", C.c_int), ("num_cells", C.c_uint32),
        ("num_mappings", C.c_uint32),
        ("cell_to_node", C.c_void_p * MAXM), ("mesh_dim", C.c_uint32 * MAXM),
        ("num_points", C.c_uint32 * MAXM), ("point_data", C.c_void_p * MAXM),
        ("point_dim", C.c_uint32 * MAXM), ("cell_data", C.c_void_p * MAXCD),
        ("cell_dim", C.c_uint32 * MAXCD), ("num_cell_data", C.c_uint32),
    ]


class Hplan(C.Structure):
    _fields_ = [
        ("num_colours", C.c_uint32), ("num_mappings", C.c_uint32),
        ("mesh_dim", C.c_uint32 * MAXM),
        ("colour_cell_offset", C.POINTER(C.c_uint32)),
        ("colour_block_offset", C.POINTER(C.c_uint32)),
        ("colour_cache_offset", C.POINTER(C.c_uint32)),
        ("cell_list", C.POINTER(C.c_uint32) * MAXM),
        ("cell_colours", C.POINTER(C.c_uint8)),
        ("num_cell_colours", C.POINTER(C.c_uint8)),
        ("block_offsets", C.POINTER(C.c_uint32)),
        ("points_to_be_cached", C.POINTER(C.c_uint32)),
        ("points_to_be_cached_offsets", C.POINTER(C.c_uint32)),
        ("cell_inv_permutation", C.POINTER(C.c_uint32)),
        ("num_cells", C.c_uint32), ("num_blocks", C.c_uint32), ("num_cached", C.c_uint32),
    ]


def load(path: str = LIB_PATH):
    """The oracle library at `path` with its signatures set: the normal
    build, or a planted-mistake variant (oracle/lib/planted/, made by
    `make planted`) for the tests that show a known answer catches it."""
    if path == LIB_PATH and not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(path)
    V, U = C.c_void_p, C.c_uint32
    sig = {
        "oracle_res_calc_func": [V] * 8 + [C.c_uint, C.c_uint],
        "oracle_lulesh_func": [V, V, V, C.c_double, C.c_double, C.c_double, V, V],
        "oracle_loop_serial": [V, U, V, V],
        "oracle_loop_colour": [V, U, C.c_int, V],
        "oracle_loop_hier_replay": [V, V, U],
        "oracle_loop_owner_replay": [V, V, U],
        "oracle_colour_cells": [V, U, U, U, V],
        "oracle_colour_cells_with_ordering": [V, U, U, V],
        "oracle_hplan_build": [V, V, U, U, U, V, U, V],
        "oracle_hplan_free": [V],
        "oracle_reorder_to_partition": [V, U, V],
        "oracle_point_renumbering_by_partition": [V, U, U, U, V, V],
        "oracle_point_renumbering_first_touch": [V, U, U, U, V],
        "oracle_mesh_reorder_cells": [V, U, U, V, V],
        "oracle_graph_csr": [V, U, U, U, V, V],
        "oracle_cell_to_cell_graph": [V, V, U, U, V, V, C.c_uint64],
        "oracle_heuristic_partition": [V, V, U, V, U, V],
        "oracle_pq_run": [V, U, V],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.oracle_graph_csr.restype = C.c_uint32
    lib.oracle_cell_to_cell_graph.restype = C.c_uint64
    lib.oracle_hplan_free.restype = None
    lib.oracle_res_calc_func.restype = None
    lib.oracle_lulesh_func.restype = None
    return lib


lib = load()


def _p(a):
    return None if a is None else a.ctypes.data


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class OracleProblem:
    """Holds the numpy arrays alive behind an oracle_problem struct.

    maps: list of (C, md) u32; point_data: list of arrays, AOS (n, dim) or
    SOA (dim, n) when soa; index 0 is the increment target.  cell_data: SOA
    (dim, C) arrays.
    """

    def __init__(self, kernel, maps, num_points, point_data, cell_data=(), soa=False,
                 library=None):
        self.lib = library or lib
        self.maps = [_u32(m) for m in maps]
        self.point_data = [None if d is None else np.array(d, dtype=np.float64, order="C")
                           for d in point_data]
        # copies: lulesh writes its element volume into cell data 1
        self.cell_data = [np.array(c, dtype=np.float64, order="C") for c in cell_data]
        s = Problem()
        s.kernel, s.soa = kernel, int(soa)
        s.num_cells = self.maps[0].shape[0]
        s.num_mappings = len(self.maps)
        for k, m in enumerate(self.maps):
            s.cell_to_node[k] = _p(m)
            s.mesh_dim[k] = m.shape[1]
            s.num_points[k] = int(num_points[k])
            d = self.point_data[k]
            s.point_data[k] = _p(d)
            if d is None:
                s.point_dim[k] = 0
            elif d.ndim == 1:
                s.point_dim[k] = 1
            else:
                s.point_dim[k] = d.shape[0] if soa else d.shape[1]
        for k, c in enumerate(self.cell_data):
            s.cell_data[k] = _p(c)
            s.cell_dim[k] = c.shape[0] if c.ndim > 1 else 1
        s.num_cell_data = len(self.cell_data)
        self.s = s

    @property
    def target(self) -> np.ndarray:
        return self.point_data[0]

    def loop_serial(self, iters: int, with_scale: bool = False):
        """Serial loop; returns S (the error scale) when with_scale."""
        scale = np.zeros_like(self.point_data[0]) if with_scale else None
        secs = C.c_double(0.0)
        if self.lib.oracle_loop_serial(C.byref(self.s), iters, _p(scale), C.byref(secs)) != 0:
            raise RuntimeError("oracle_loop_serial failed")
        self.last_loop_seconds = secs.value
        return scale

    def loop_colour(self, iters: int, nthreads: int = 1) -> float:
        """Colour-major loop; returns the iteration loop's seconds."""
        secs = C.c_double(0.0)
        if self.lib.oracle_loop_colour(C.byref(self.s), iters, nthreads, C.byref(secs)) != 0:
            raise RuntimeError("oracle_loop_colour failed")
        self.last_loop_seconds = secs.value
        return secs.value

    def loop_hier_replay(self, plan: "OraclePlan", iters: int) -> None:
        if self.lib.oracle_loop_hier_replay(C.byref(self.s), C.byref(plan.h), iters) != 0:
            raise RuntimeError("oracle_loop_hier_replay failed")


class OwnerPlanView(C.Structure):
    _fields_ = [("mesh_dim", C.c_uint32), ("num_blocks", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("blk_cell_off", "blk_cache_off", "blk_owned", "cached",
                                  "cells", "loc", "colour", "blk_ncol")]


def loop_owner_replay(problem: "OracleProblem", plan: dict, iters: int) -> None:
    """Replays an owner-compute plan (dict of numpy arrays as
    cuda_fot_amd.plan.owner_plan returns it)."""
    arrs = {k: np.ascontiguousarray(plan[k]) for k in
            ("blk_cell_off", "blk_cache_off", "blk_owned", "cached", "cells", "loc",
             "colour", "blk_ncol")}
    v = OwnerPlanView()
    v.mesh_dim = int(plan["mesh_dim"])
    v.num_blocks = int(arrs["blk_owned"].size)
    for k, a in arrs.items():
        setattr(v, k, _p(a))
    if lib.oracle_loop_owner_replay(C.byref(problem.s), C.byref(v), iters) != 0:
        raise RuntimeError("oracle_loop_owner_replay failed")


class OraclePlan:
    """oracle_hplan_build: the hierarchical plan as one dict per colour."""

    def __init__(self, maps, num_points0, block_size, partition=None, library=None):
        self.lib = library or lib
        self.maps = [_u32(m) for m in maps]
        arr = (C.c_void_p * len(self.maps))(*[_p(m) for m in self.maps])
        dims = np.array([m.shape[1] for m in self.maps], dtype=np.uint32)
        part = None if partition is None else _u32(partition)
        self.h = Hplan()
        rc = self.lib.oracle_hplan_build(arr, _p(dims), len(self.maps), self.maps[0].shape[0],
                                    num_points0, _p(part), block_size, C.byref(self.h))
        if rc != 0:
            raise RuntimeError("oracle_hplan_build failed")

    def colours(self) -> list[dict]:
        h = self.h
        cco =np.ctypeslib.as_array(h.colour_cell_offset, shape=(h.num_colours + 1,))
        cbo = np.ctypeslib.as_array(h.colour_block_offset, shape=(h.num_colours + 1,))
        cko = np.ctypeslib.as_array(h.colour_cache_offset, shape=(h.num_colours + 1,))
        n, nb = h.num_cells, h.num_blocks
        cell_colours = np.ctypeslib.as_array(h.cell_colours, shape=(n,))
        ncc = np.ctypeslib.as_array(h.num_cell_colours, shape=(nb,))
        bo = np.ctypeslib.as_array(h.block_offsets, shape=(nb + h.num_colours,))
        ptbc = np.ctypeslib.as_array(h.points_to_be_cached, shape=(max(1, h.num_cached),))
        ptbco = np.ctypeslib.as_array(h.points_to_be_cached_offsets, shape=(nb + h.num_colours,))
        cip = np.ctypeslib.as_array(h.cell_inv_permutation, shape=(n,))
        lists = [np.ctypeslib.as_array(h.cell_list[k], shape=(n * h.mesh_dim[k],))
                 for k in range(h.num_mappings)]
        out = []
        for c in range(h.num_colours):
            c0, c1 = int(cco[c]), int(cco[c + 1])
            b0, b1 = int(cbo[c]), int(cbo[c + 1])
            k0, k1 = int(cko[c]), int(cko[c + 1])
            o0, o1 = b0 + c, b1 + c + 1
            ptbc_off = ptbco[o0:o1].copy()
            out.append({
                "points_to_be_cached": ptbc[k0:k1].copy(),
                "points_to_be_cached_offsets": ptbc_off,
                "cell_list": [lists[k][c0 * h.mesh_dim[k]:c1 * h.mesh_dim[k]].copy()
                              for k in range(h.num_mappings)],
                "cell_colours": cell_colours[c0:c1].copy(),
                "num_cell_colours": ncc[b0:b1].copy(),
                "block_offsets": bo[o0:o1].copy(),
                "cell_inv_permutation": cip[c0:c1].copy(),
                "shared_size": int(np.max(np.diff(ptbc_off))) if b1 > b0 else 0,
            })
        return out

    def __del__(self):
        try:
            self.lib.oracle_hplan_free(C.byref(self.h))
        except Exception:
            pass


def res_calc_func(x1, x2, q1, q2, adt1, adt2):
    """The res_calc user function on one edge: (res1, res2)."""
    ins = [np.ascontiguousarray(v, dtype=np.float64) for v in (x1, x2, q1, q2)]
    a1 = np.array([adt1], dtype=np.float64)
    a2 = np.array([adt2], dtype=np.float64)
    r1 = np.zeros(4)
    r2 = np.zeros(4)
    lib.oracle_res_calc_func(*[_p(v) for v in ins], _p(a1), _p(a2), _p(r1), _p(r2), 1, 1)
    return r1, r2


def lulesh_func(xyz, sig):
    """lulesh's element function on one hex: xyz (8, 3) corner coordinates,
    sig (3,) diagonal stresses -> (forces (8, 3), volume)."""
    c = np.asarray(xyz, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(c[:, d]) for d in range(3))
    f = np.zeros(24)
    det = np.zeros(1)
    lib.oracle_lulesh_func(_p(x), _p(y), _p(z), float(sig[0]), float(sig[1]), float(sig[2]),
                           _p(f), _p(det))
    return f.reshape(3, 8).T.copy(), float(det[0])


def colour_cells(maps0, num_points, library=None):
    m = _u32(maps0)
    col = np.empty(m.shape[0], dtype=np.uint8)
    n = (library or lib).oracle_colour_cells(_p(m), m.shape[0], m.shape[1], num_points,
                                             _p(col))
    if n < 0:
        raise RuntimeError("colour overflow")
    return col, n


def colour_cells_with_ordering(local):
    m = _u32(local)
    col = np.empty(m.shape[0], dtype=np.uint8)
    n = lib.oracle_colour_cells_with_ordering(_p(m), m.shape[0], m.shape[1], _p(col))
    if n < 0:
        raise RuntimeError("colour overflow")
    return col, n


def reorder_to_partition(partition):
    p = _u32(partition).copy()
    inv = np.empty_like(p)
    lib.oracle_reorder_to_partition(_p(p), p.size, _p(inv))
    return p, inv


def point_renumbering_by_partition(maps0, num_points, partition):
    m = _u32(maps0)
    part = _u32(partition)
    perm = np.empty(num_points, dtype=np.uint32)
    lib.oracle_point_renumbering_by_partition(_p(m), m.shape[0], m.shape[1], num_points,
                                              _p(part), _p(perm))
    return perm


def point_renumbering_first_touch(maps0, num_points):
    m = _u32(maps0)
    perm = np.empty(num_points, dtype=np.uint32)
    lib.oracle_point_renumbering_first_touch(_p(m), m.shape[0], m.shape[1], num_points,
                                             _p(perm))
    return perm


def mesh_reorder_cells(maps0, point_perm):
    m = _u32(maps0).copy()
    pp = _u32(point_perm)
    inv = np.empty(m.shape[0], dtype=np.uint32)
    lib.oracle_mesh_reorder_cells(_p(m), m.shape[0], m.shape[1], _p(pp), _p(inv))
    return m, inv


def graph_csr(maps0, num_points):
    m = _u32(maps0)
    n, md = m.shape
    ind = np.empty(num_points + 1, dtype=np.uint32)
    end = np.empty(max(1, n * md * max(1, md - 1)), dtype=np.uint32)
    na = lib.oracle_graph_csr(_p(m), n, md, num_points, _p(ind), _p(end))
    return ind, end[:na].copy()


def cell_to_cell_pairs(maps, num_points):
    ms = [_u32(m) for m in maps]
    arr = (C.c_void_p * len(ms))(*[_p(m) for m in ms])
    dims = np.array([m.shape[1] for m in ms], dtype=np.uint32)
    nps = _u32(num_points)
    n = lib.oracle_cell_to_cell_graph(arr, _p(dims), len(ms), ms[0].shape[0], _p(nps), None, 0)
    pairs = np.empty((max(1, n), 2), dtype=np.uint32)
    lib.oracle_cell_to_cell_graph(arr, _p(dims), len(ms), ms[0].shape[0], _p(nps), _p(pairs), n)
    return pairs[:n].copy()


def cell_to_cell_csr(maps, num_points):
    """GraphCSR of the pair mesh of getCellToCellGraph (partition.cu:105-107)."""
    pairs = cell_to_cell_pairs(maps, num_points)
    n = _u32(maps[0]).shape[0]
    if pairs.size == 0:
        return np.zeros(n + 1, dtype=np.uint32), np.empty(0, dtype=np.uint32)
    return graph_csr(pairs, n)


def heuristic_partition(ind, end, coarse, block_size):
    ind, end, coarse = _u32(ind), _u32(end), _u32(coarse)
    n = ind.size - 1
    res = np.empty(n, dtype=np.uint32)
    nb = lib.oracle_heuristic_partition(_p(ind), _p(end), n, _p(coarse), block_size, _p(res))
    if nb < 0:
        raise RuntimeError("heuristic partition failed")
    return res, nb


def pq_run(ops):
    """Runs PriorityQueue operations [(op, a, b)]; op 0 push, 1 popMax,
    2 modify, 3 getPriority.  Returns the flat outputs."""
    o = np.ascontiguousarray(np.array(ops, dtype=np.int32).reshape(-1, 3))
    out = np.zeros(2 * len(o) + 2, dtype=np.int32)
    n = lib.oracle_pq_run(_p(o), len(o), _p(out))
    return out[:n].tolist()
