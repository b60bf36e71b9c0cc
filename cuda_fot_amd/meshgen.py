"""Synthetic meshes (native generators in libfot)."""
from __future__ import annotations

import numpy as np

from .capi import call, ptr


def grid_edges(N: int, M: int) -> np.ndarray:
    """Grid::fillEdgeList (grid.hpp:77-104): (E, 2) u32 edges of an N x M grid."""
    ne = (N - 1) * M + N * (M - 1)
    e = np.empty((ne, 2), dtype=np.uint32)
    call("fot_gen_grid_edges", N, M, ptr(e))
    return e


def hex_grid(N1: int, N2: int, N3: int) -> np.ndarray:
    """Grid::fillCellList3D (grid.hpp:415-460): (C, 8) u32 hexes."""
    nc = (N1 - 1) * (N2 - 1) * (N3 - 1)
    c = np.empty((nc, 8), dtype=np.uint32)
    call("fot_gen_hex_grid", N1, N2, N3, ptr(c))
    return c


def permutation(n: int, seed: int) -> np.ndarray:
    p = np.empty(n, dtype=np.uint32)
    call("fot_gen_permutation", n, seed, ptr(p))
    return p


def uniform(n: int, seed: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    call("fot_gen_uniform", n, seed, lo, hi, ptr(out))
    return out


def normal(n: int, seed: int, mean: float = 0.0, sd: float = 1.0) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    call("fot_gen_normal", n, seed, mean, sd, ptr(out))
    return out


def airfoil(nx: int, ny: int, seed: int = 0, shuffle: int = 7) -> dict:
    """Airfoil-like quad mesh of nx x ny cells (interior edges only) with the
    res_calc data: pedge (E,2), pecell (E,2), x (N,2), q (C,4), adt (C,)."""
    nn, nc = (nx + 1) * (ny + 1), nx * ny
    ne = nx * (ny - 1) + (nx - 1) * ny
    pedge = np.empty((ne, 2), np.uint32)
    pecell = np.empty((ne, 2), np.uint32)
    x = np.empty((nn, 2), np.float64)
    q = np.empty((nc, 4), np.float64)
    adt = np.empty(nc, np.float64)
    call("fot_gen_airfoil", nx, ny, seed, shuffle, ptr(pedge), ptr(pecell), ptr(x), ptr(q),
         ptr(adt))
    return {"pedge": pedge, "pecell": pecell, "x": x, "q": q, "adt": adt,
            "num_nodes": nn, "num_cells": nc, "num_edges": ne, "nx": nx, "ny": ny}


def tri_mesh(nx: int, ny: int, seed: int = 0, shuffle: int = 6) -> dict:
    """Triangle mesh and data of the shallow-water (volna) edge loop: an
    nx x ny quad grid split on random diagonals, boundary edges included.
    edge_cells (E,2) (cell, cell) on the boundary; is_boundary (E,) 0/1;
    normals (2,E), fluxes (3,E), bathy (2,E) SOA; volume (C,); state (C,4)."""
    nc = 2 * nx * ny
    ne = nx * (ny + 1) + (nx + 1) * ny + nx * ny
    out = {"edge_cells": np.empty((ne, 2), np.uint32), "is_boundary": np.empty(ne),
           "normals": np.empty((2, ne)), "fluxes": np.empty((3, ne)), "bathy": np.empty((2, ne)),
           "volume": np.empty(nc), "state": np.empty((nc, 4))}
    call("fot_gen_tri_mesh", nx, ny, seed, shuffle, *(ptr(out[k]) for k in
         ("edge_cells", "is_boundary", "normals", "fluxes", "bathy", "volume", "state")))
    out.update(num_cells=nc, num_edges=ne, nx=nx, ny=ny)
    return out


def hex_faces(n1: int, n2: int, n3: int, seed: int = 0, shuffle: int = 6) -> dict:
    """mini_aero-style hex mesh: interior faces -> 2 cells, face data (12, F)
    SOA (centre, normal, tangent, binormal), cell values (C, 28)."""
    nc = n1 * n2 * n3
    nf = (n1 - 1) * n2 * n3 + n1 * (n2 - 1) * n3 + n1 * n2 * (n3 - 1)
    out = {"face_cells": np.empty((nf, 2), np.uint32), "face": np.empty((12, nf)),
           "cells": np.empty((nc, 28))}
    call("fot_gen_hex_faces", n1, n2, n3, seed, shuffle, ptr(out["face_cells"]),
         ptr(out["face"]), ptr(out["cells"]))
    out.update(num_cells=nc, num_faces=nf)
    return out


def quad_grid(N: int, M: int) -> np.ndarray:
    """Grid::fillCellList (grid.hpp:386-411): (C, 4) u32 quads of an N x M
    point grid."""
    c = np.empty(((N - 1) * (M - 1), 4), dtype=np.uint32)
    call("fot_gen_quad_grid", N, M, ptr(c))
    return c
