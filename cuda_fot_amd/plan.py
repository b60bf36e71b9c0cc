"""Host planning of libfot: colourings, the hierarchical plan, partitions,
renumberings and orderings (all native, in libfot.so)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import call, lib, ptr, u32


def colour_cells(maps0: np.ndarray, num_points: int) -> tuple[np.ndarray, int]:
    """Mesh::colourCells (mesh.hpp:117-162) on one mapping."""
    m = u32(maps0)
    n, md = m.shape
    col = np.empty(n, dtype=np.uint8)
    nc = C.c_uint32()
    call("fot_colour_cells", ptr(m), n, md, num_points, ptr(col), C.byref(nc))
    return col, nc.value


def colour_cells_with_ordering(local: np.ndarray) -> tuple[np.ndarray, int]:
    """colourCellsWithOrdering (colouring.cu:94-102)."""
    m = u32(local)
    n, md = m.shape
    col = np.empty(n, dtype=np.uint8)
    nc = C.c_uint32()
    call("fot_colour_cells_with_ordering", ptr(m), n, md, ptr(col), C.byref(nc))
    return col, nc.value


def hplan(maps, num_points0: int, block_size: int, partition=None) -> list[dict]:
    """HierarchicalColourMemory (colouring.hpp:22-404): one dict per block
    colour with the reference's arrays (cell lists SOA)."""
    ms = [u32(m) for m in maps]
    n = ms[0].shape[0]
    arr = (C.c_void_p * len(ms))(*[ptr(m) for m in ms])
    dims = np.array([m.shape[1] for m in ms], dtype=np.uint32)
    part = None if partition is None else u32(partition)
    h = C.c_void_p()
    call("fot_hplan_build", arr, ptr(dims), len(ms), n, num_points0, ptr(part),
         block_size, C.byref(h))
    try:
        info = capi.HplanInfo()
        call("fot_hplan_info_get", h, C.byref(info))
        out = []
        for c in range(info.num_colours):
            v = capi.HplanColour()
            call("fot_hplan_colour_get", h, c, C.byref(v))
            nb, nc, ncached = v.num_blocks, v.num_cells, v.num_cached

            def arr_u32(p, k):
                return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.empty(0, np.uint32)

            def arr_u8(p, k):
                return np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.empty(0, np.uint8)

            out.append({
                "points_to_be_cached": arr_u32(v.points_to_be_cached, ncached),
                "points_to_be_cached_offsets": arr_u32(v.points_to_be_cached_offsets, nb + 1),
                "cell_list": [arr_u32(v.cell_list[k], nc * int(dims[k])) for k in range(len(ms))],
                "cell_colours": arr_u8(v.cell_colours, nc),
                "num_cell_colours": arr_u8(v.num_cell_colours, nb),
                "block_offsets": arr_u32(v.block_offsets, nb + 1),
                "cell_inv_permutation": arr_u32(v.cell_inv_permutation, nc),
                "shared_size": v.shared_size,
            })
        return out
    finally:
        lib.fot_hplan_free(h)


def owner_plan(maps0, num_points: int, block_size: int, partition=None,
               point_partition=None) -> dict:
    """Owner-compute plan (fot_oplan_build) as a dict of numpy arrays:
    owners from point_partition (one part id per mapping-0 point) when given,
    else from the cell split (partition per cell, or chunks)."""
    m = u32(maps0)
    part = None if partition is None else u32(partition)
    ppart = None if point_partition is None else u32(point_partition)
    h = C.c_void_p()
    call("fot_oplan_build", ptr(m), m.shape[0], m.shape[1], num_points, ptr(part), block_size,
         ptr(ppart), C.byref(h))
    try:
        v = capi.OplanView()
        call("fot_oplan_get", h, C.byref(v))
        nb, npos, nc = v.num_blocks, v.num_positions, v.num_cached

        def cp(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy()

        return {"mesh_dim": v.mesh_dim, "blk_cell_off": cp(v.blk_cell_off, nb + 1),
                "blk_cache_off": cp(v.blk_cache_off, nb + 1), "blk_owned": cp(v.blk_owned, nb),
                "cached": cp(v.cached, nc), "cells": cp(v.cells, npos),
                "loc": cp(v.loc, npos * v.mesh_dim), "colour": cp(v.colour, npos),
                "blk_ncol": cp(v.blk_ncol, nb)}
    finally:
        lib.fot_oplan_free(h)


def reorder_to_partition(partition) -> tuple[np.ndarray, np.ndarray]:
    """Mesh::reorderToPartition (mesh.hpp:227-245): (sorted partition, inverse
    cell permutation)."""
    p = u32(partition).copy()
    inv = np.empty_like(p)
    call("fot_reorder_to_partition", ptr(p), p.size, ptr(inv))
    return p, inv


def point_renumbering_by_partition(maps0, num_points: int, partition) -> np.ndarray:
    m = u32(maps0)
    part = u32(partition)
    perm = np.empty(num_points, dtype=np.uint32)
    call("fot_point_renumbering_by_partition", ptr(m), m.shape[0], m.shape[1], num_points,
         ptr(part), ptr(perm))
    return perm


def point_renumbering_first_touch(maps0, num_points: int) -> np.ndarray:
    m = u32(maps0)
    perm = np.empty(num_points, dtype=np.uint32)
    call("fot_point_renumbering_first_touch", ptr(m), m.shape[0], m.shape[1], num_points,
         ptr(perm))
    return perm


def mesh_reorder_cells(maps0, point_perm) -> tuple[np.ndarray, np.ndarray]:
    """Mesh::reorder (mesh.hpp:199-225), mapping 0: (new map, inverse cell
    permutation)."""
    m = u32(maps0).copy()
    pp = u32(point_perm)
    inv = np.empty(m.shape[0], dtype=np.uint32)
    call("fot_mesh_reorder_cells", ptr(m), m.shape[0], m.shape[1], ptr(pp), ptr(inv))
    return m, inv


def graph_csr(maps0, num_points: int) -> tuple[np.ndarray, np.ndarray]:
    """GraphCSR (reorder.hpp:39-83)."""
    m = u32(maps0)
    n, md = m.shape
    ind = np.empty(num_points + 1, dtype=np.uint32)
    end = np.empty(max(1, n * md * max(1, md - 1)), dtype=np.uint32)
    na = C.c_uint32()
    call("fot_graph_csr", ptr(m), n, md, num_points, ptr(ind), ptr(end), C.byref(na))
    return ind, end[: na.value].copy()


def cell_to_cell_csr(maps, num_points) -> tuple[np.ndarray, np.ndarray]:
    """Mesh::getCellToCellGraph (mesh.hpp:317-344) as a CSR."""
    ms = [u32(m) for m in maps]
    n = ms[0].shape[0]
    arr = (C.c_void_p * len(ms))(*[ptr(m) for m in ms])
    dims = np.array([m.shape[1] for m in ms], dtype=np.uint32)
    nps = u32(num_points)
    ind = np.empty(n + 1, dtype=np.uint32)
    na = C.c_uint64()
    call("fot_cell_to_cell_csr", arr, ptr(dims), len(ms), n, ptr(nps), ptr(ind), None,
         C.byref(na))
    end = np.empty(max(1, na.value), dtype=np.uint32)
    call("fot_cell_to_cell_csr", arr, ptr(dims), len(ms), n, ptr(nps), ptr(ind), ptr(end),
         C.byref(na))
    return ind, end[: na.value].copy()


def heuristic_partition(ind, end, coarse, block_size: int) -> tuple[np.ndarray, int]:
    """HeuristicalPartition (partition.cu:128-275) given the coarse partition."""
    ind, end, coarse = u32(ind), u32(end), u32(coarse)
    n = ind.size - 1
    res = np.empty(n, dtype=np.uint32)
    nb = C.c_uint32()
    call("fot_heuristic_partition", ptr(ind), ptr(end), n, ptr(coarse), block_size,
         ptr(res), C.byref(nb))
    return res, nb.value


def kway_partition(ind, end, max_part: int) -> tuple[np.ndarray, int]:
    ind, end = u32(ind), u32(end)
    n = ind.size - 1
    part = np.empty(n, dtype=np.uint32)
    npart = C.c_uint32()
    call("fot_kway_partition", ptr(ind), ptr(end), n, max_part, ptr(part), C.byref(npart))
    return part, npart.value


def bisect_partition(ind, end, max_part: int) -> tuple[np.ndarray, int]:
    """Recursive-bisection parts of at most max_part vertices (equal sizes)."""
    ind, end = u32(ind), u32(end)
    n = ind.size - 1
    part = np.empty(n, dtype=np.uint32)
    npart = C.c_uint32()
    call("fot_bisect_partition", ptr(ind), ptr(end), n, max_part, ptr(part), C.byref(npart))
    return part, npart.value


PARTITIONERS = {"kway": kway_partition, "bisect": bisect_partition}


def rcm_order(ind, end, seed_order=None) -> np.ndarray:
    """scipy's reverse Cuthill-McKee of the CSR pattern (symmetrised), as a
    permutation old -> new (scripts/rcm.py:20-21).  seed_order: the degree
    order that seeds the components (scipy uses numpy's argsort of the
    degrees); None: ascending degree, ties to the lower id."""
    ind, end = u32(ind), u32(end)
    seeds = None if seed_order is None else u32(seed_order)
    perm = np.empty(ind.size - 1, dtype=np.uint32)
    call("fot_rcm_order", ptr(ind), ptr(end), ind.size - 1, ptr(seeds), ptr(perm))
    return perm


def gps_order(ind, end) -> np.ndarray:
    ind, end = u32(ind), u32(end)
    perm = np.empty(ind.size - 1, dtype=np.uint32)
    call("fot_gps_order", ptr(ind), ptr(end), ind.size - 1, ptr(perm))
    return perm


def partition_metis_enh(maps, num_points, block_size: int, tolerance: float = 1.001,
                        method: str = "kway"):
    """Block partition of the cells (Problem::partition, problem.hpp:484-499):
    parts of at most floor(block_size / tolerance) cells
    (partitionMetisEnh, partition.hpp:23-29) from a METIS stand-in: greedy
    k-way growing ("kway") or recursive bisection ("bisect")."""
    ind, end = cell_to_cell_csr(maps, num_points)
    max_part = int(np.floor(block_size / tolerance))
    return PARTITIONERS[method](ind, end, max_part)[0]


def partition_points(maps0, num_points: int, points_per_part: int,
                     method: str = "kway") -> np.ndarray:
    """Parts of at most points_per_part points of mapping 0's point graph
    (the owner partition of the owner-compute plan)."""
    ind, end = graph_csr(maps0, num_points)
    return PARTITIONERS[method](ind, end, points_per_part)[0]


def partition_ours(maps, num_points, block_size: int, num_blocks_in_coarse: int,
                   tolerance: float = 1.001):
    """partitionOurs (partition.cu:42-66): coarse k-way parts of about
    block_size * num_blocks_in_coarse cells, at most `tolerance` times that
    (METIS's ubvec there), then the heuristic partitioner."""
    ind, end = cell_to_cell_csr(maps, num_points)
    coarse, _ = kway_partition(ind, end,
                               max(1, int(np.floor(block_size * num_blocks_in_coarse *
                                                   tolerance))))
    return heuristic_partition(ind, end, coarse, block_size)[0]
