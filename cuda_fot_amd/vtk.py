"""VTK output of a hierarchical plan, for looking at colourings and blocks.

* `write_vtk_ascii` is writeMeshToVTKAscii (visualization/graph_write_VTK.hpp:
  23-110): a legacy ASCII unstructured grid (lines, quads or hexahedra) with
  up to three integer cell scalars, VTK_IND_THR_COL (thread colour),
  VTK_IND_BLK_ID (block) and VTK_IND_BLK_COL (block colour); quads whose
  corner order crosses itself are untangled first (geometry_routines.hpp:
  23-49).
* `plan_scalars` walks a HierarchicalColourMemory plan as visualise does
  (res_calc.cu:305-363): cells in plan order, one block id per block across
  all block colours.
* `visualise_res_calc` / `main` are visualise + mainVisualise (res_calc.cu:
  305-373) over an airfoil directory in the reference's text format:
  res_calc.vtk (the directory as read) and res_calc_part.vtk (GPS-reordered,
  partitioned and renumbered first).

    python -m cuda_fot_amd.vtk <input_dir> <block_size>
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import plan

# mesh dimension -> VTK cell type (graph_write_VTK.hpp:17-21)
VTK_CELL_TYPES = {2: 3, 4: 9, 8: 12}  # line, quad, hexahedron
SCALAR_NAMES = ("VTK_IND_THR_COL", "VTK_IND_BLK_ID", "VTK_IND_BLK_COL")


def _turn(a, b, c) -> np.float32:
    """Signed turn of a -> b -> c in the plane (single precision, as the
    reference's float coordinates)."""
    s1x, s1y = b[0] - a[0], b[1] - a[1]
    s2x, s2y = c[0] - b[0], c[1] - b[1]
    return s1y * s2x - s1x * s2y


def _crosses(a0, a1, b0, b1) -> bool:
    return bool(_turn(a0, a1, b0) * _turn(a0, a1, b1) < 0 and
                _turn(b0, b1, a0) * _turn(b0, b1, a1) < 0)


def untangle_quad(pts) -> tuple[int, int, int, int]:
    """Corner order of a quad whose given order may trace a bow tie
    (geometry_routines.hpp:38-49): if side 0-1 crosses side 2-3 the order is
    0 2 1 3, else if side 1-2 crosses side 3-0 it is 0 1 3 2."""
    p = np.asarray(pts, dtype=np.float32)[:, :2]
    if _crosses(p[0], p[1], p[2], p[3]):
        return (0, 2, 1, 3)
    if _crosses(p[1], p[2], p[3], p[0]):
        return (0, 1, 3, 2)
    return (0, 1, 2, 3)


def write_vtk_ascii(path, cells, coords, thread_colour=None, block_id=None,
                    block_colour=None, title="Output from graph partitioning test.") -> None:
    """cells (C, md) point ids, md in 2, 4, 8; coords (N, 2 or 3); the
    optional per-cell scalars as integer arrays of length C."""
    cells = np.asarray(cells, dtype=np.int64)
    c = np.asarray(coords, dtype=np.float32)
    if c.ndim != 2 or c.shape[1] not in (2, 3):
        raise ValueError("coords must be (N, 2) or (N, 3)")
    nc, md = cells.shape
    if md not in VTK_CELL_TYPES:
        raise ValueError(f"no VTK cell type for {md} points per cell")
    if nc and (cells.min() < 0 or cells.max() >= c.shape[0]):
        raise ValueError("cell point id out of range")
    xyz = np.zeros((c.shape[0], 3), dtype=np.float32)
    xyz[:, :c.shape[1]] = c
    if md == 4:
        cells = cells.copy()
        for i in range(nc):
            order = untangle_quad(xyz[cells[i]])
            if order != (0, 1, 2, 3):
                cells[i] = cells[i][list(order)]
    scalars = [(n, v) for n, v in zip(SCALAR_NAMES, (thread_colour, block_id, block_colour))
               if v is not None and len(v) > 0]
    for n, v in scalars:
        if len(v) != nc:
            raise ValueError(f"{n}: {len(v)} values for {nc} cells")
    with open(path, "w") as f:
        f.write(f"# vtk DataFile Version 2.0\n{title}\nASCII\nDATASET UNSTRUCTURED_GRID\n\n")
        f.write(f"POINTS {xyz.shape[0]} float\n")
        for p in xyz:
            f.write(f"{p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        f.write(f"\nCELLS {nc} {(md + 1) * nc}\n")
        for row in cells:
            f.write(f"{md} " + " ".join(str(int(v)) for v in row) + "\n")
        f.write(f"\nCELL_TYPES {nc}\n")
        f.write(f"{VTK_CELL_TYPES[md]}\n" * nc)
        if scalars:
            f.write(f"\nCELL_DATA {nc}\n")
        for n, v in scalars:
            f.write(f"SCALARS {n} int 1\nLOOKUP_TABLE default\n")
            f.write("".join(f"{int(x)}\n" for x in v))
            f.write("\n")


def plan_scalars(maps, num_points0: int, block_size: int, partition=None, mapping: int = 0):
    """Cells of a HierarchicalColourMemory plan in plan order: (cells (C, md)
    through `mapping`, thread colour, block id, block colour, original cell
    id), as res_calc.cu:317-354 collects them."""
    colours = plan.hplan(maps, num_points0, block_size, partition)
    md = np.asarray(maps[mapping]).shape[1]
    rows, thr, bid, bcol, orig = [], [], [], [], []
    blk = 0
    for c, m in enumerate(colours):
        n = m["cell_colours"].size
        soa = m["cell_list"][mapping].reshape(md, n)  # index<true>(n, i, md, j)
        rows.append(soa.T)
        thr.append(m["cell_colours"].astype(np.int64))
        bcol.append(np.full(n, c, dtype=np.int64))
        off = m["block_offsets"].astype(np.int64)
        bid.append(blk + np.repeat(np.arange(off.size - 1), np.diff(off)))
        blk += off.size - 1
        orig.append(m["cell_inv_permutation"].astype(np.int64))
    cat = (lambda xs, shape: np.concatenate(xs) if xs else np.zeros(shape, np.int64))
    return (np.concatenate(rows) if rows else np.zeros((0, md), np.int64),
            cat(thr, 0), cat(bid, 0), cat(bcol, 0), cat(orig, 0))


def visualise_res_calc(input_dir: str, fname: str, partition: bool, block_size: int) -> str:
    """visualise (res_calc.cu:305-363): the edges (mapping 1, lines over the
    nodes) coloured by the plan of the airfoil directory, optionally after
    GPS ordering, partitioning and renumbering."""
    from . import io
    p = io.read_airfoil_dir(input_dir, block_size)
    if partition:
        p.reorder("gps", same_mapping=[2, 3])
        p.partition(1.001)
        p.reorder_to_partition()
        p.renumber_points()
    cells, thr, bid, bcol, _ = plan_scalars(p.maps, p.num_points[0], block_size,
                                            p.partition_vector, mapping=1)
    x = p.point_data[1]
    path = os.path.join(input_dir, fname)
    write_vtk_ascii(path, cells, x, thr, bid, bcol)
    return path


def main(argv=None) -> int:
    """mainVisualise (res_calc.cu:365-373)."""
    argv = sys.argv if argv is None else argv
    if len(argv) < 3:
        print(f"Usage: {argv[0]} <input_dir> <block_size>", file=sys.stderr)
        return 1
    visualise_res_calc(argv[1], "res_calc.vtk", False, int(argv[2]))
    visualise_res_calc(argv[1], "res_calc_part.vtk", True, int(argv[2]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
