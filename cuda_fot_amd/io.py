"""The reference's text formats and its reorder tool.

* mapping (mesh.hpp:58-98, writeCellList :184-197): first line
  "num_points num_cells", then one line of MeshDim point ids per cell;
* point data (problem.hpp:583-615): one line of `dim` values per point, in
  point order (SOA problems are transposed on the way in and out);
* cell data (problem.hpp:617-632): one line of `dim` values per cell;
* partition (problem.hpp:553-581): the block size, then one block id per cell;
* permutation dumps (problem.hpp:674-682): one id per line.

`write_airfoil_dir` / `read_airfoil_dir` are res_calc.cu's writeAllData /
initProblem + readData (res_calc.cu:10-98); `reorder_tool` is mainReorder
(res_calc.cu:376-405): a GPS-ordered copy and a GPS-ordered, partitioned and
renumbered copy of an airfoil directory; `rcm_tool` is scripts/rcm.py (an
edge list reordered by reverse Cuthill-McKee, edges sorted).
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import capi
from .problem import Problem


class InvalidInputFile(capi.InvalidInputFile):
    """A malformed or inconsistent input file (mesh.hpp:18-21)."""

    def __init__(self, message: str):
        super().__init__(capi.ERR_INVALID_INPUT, "read", message)


def write_mapping(path, cells: np.ndarray, num_points: int) -> None:
    with open(path, "w") as f:
        f.write(f"{int(num_points)} {cells.shape[0]}\n")
        np.savetxt(f, np.asarray(cells, dtype=np.int64), fmt="%d")


def read_mapping(path, mesh_dim: int):
    """(cells (C, mesh_dim) u32, num_points); InvalidInputFile on a short file
    or an id outside the point set."""
    try:
        with open(path) as f:
            head = f.readline().split()
            if len(head) != 2:
                raise InvalidInputFile(f"{path}: bad header")
            num_points, num_cells = int(head[0]), int(head[1])
            data = np.loadtxt(f, dtype=np.int64, ndmin=2) if num_cells else \
                np.zeros((0, mesh_dim), np.int64)
    except (OSError, ValueError) as e:
        raise InvalidInputFile(f"{path}: {e}") from e
    if data.shape != (num_cells, mesh_dim):
        raise InvalidInputFile(f"{path}: expected {num_cells} x {mesh_dim} ids, got {data.shape}")
    if data.size and (data.min() < 0 or data.max() >= num_points):
        raise InvalidInputFile(f"{path}: point id out of range")
    return data.astype(np.uint32), num_points


def write_rows(path, data: np.ndarray) -> None:
    """Point or cell data, one row per point / cell (17 significant digits:
    doubles round-trip exactly)."""
    d = np.asarray(data, np.float64)
    d = d.reshape(d.shape[0], -1)
    np.savetxt(path, d, fmt="%.17g")


def read_rows(path, n: int, dim: int) -> np.ndarray:
    try:
        d = np.loadtxt(path, dtype=np.float64, ndmin=2)
    except (OSError, ValueError) as e:
        raise InvalidInputFile(f"{path}: {e}") from e
    if d.shape != (n, dim):
        raise InvalidInputFile(f"{path}: expected {n} x {dim} values, got {d.shape}")
    return d


def write_partition(path, block_size: int, partition) -> None:
    with open(path, "w") as f:
        f.write(f"{int(block_size)}\n")
        np.savetxt(f, np.asarray(partition, dtype=np.int64), fmt="%d")


def read_partition(path, num_cells: int):
    try:
        v = np.loadtxt(path, dtype=np.int64, ndmin=1)
    except (OSError, ValueError) as e:
        raise InvalidInputFile(f"{path}: {e}") from e
    if v.size != num_cells + 1:
        raise InvalidInputFile(f"{path}: expected block size and {num_cells} ids")
    return int(v[0]), v[1:].astype(np.uint32)


def write_permutation(path, perm) -> None:
    np.savetxt(path, np.asarray(perm, dtype=np.int64), fmt="%d")


def _aos(p: Problem, k: int) -> np.ndarray:
    d = p.point_data[k]
    return d.T if (p.soa and d.ndim > 1) else d


def write_airfoil_dir(p: Problem, out_dir: str, partition: bool = False) -> None:
    """writeAllData (res_calc.cu:48-98): data_res/x/q/adt and the mappings;
    all four mappings are written (the reference writes mesh_res and mesh_x
    only) so the directory reads back on its own, plus the applied point
    permutation of mapping 0 when there is one."""
    os.makedirs(out_dir, exist_ok=True)
    for k, name in enumerate(["res", "x", "q", "adt"]):
        write_rows(os.path.join(out_dir, "data_" + name), _aos(p, k))
        write_mapping(os.path.join(out_dir, "mesh_" + name), p.maps[k], p.num_points[k])
    if partition and p.partition_vector is not None:
        write_partition(os.path.join(out_dir, "mesh_part"), p.block_size, p.partition_vector)
    if p.applied_permutation is not None:
        write_permutation(os.path.join(out_dir, "point_permutation"), p.applied_permutation)


def read_airfoil_dir(in_dir: str, block_size: int = 128, soa: bool = False) -> Problem:
    """initProblem + readData (res_calc.cu:10-40): the four mappings and
    their point data; mesh_part, when present, becomes the partition."""
    maps, npts = [], []
    for name in ["res", "x", "q", "adt"]:
        c, n = read_mapping(os.path.join(in_dir, "mesh_" + name), 2)
        maps.append(c)
        npts.append(n)
    if len({m.shape[0] for m in maps}) != 1:
        raise InvalidInputFile(f"{in_dir}: mappings disagree on the number of cells")
    data = [read_rows(os.path.join(in_dir, "data_" + name), npts[k], dim)
            for k, (name, dim) in enumerate([("res", 4), ("x", 2), ("q", 4), ("adt", 1)])]
    data[3] = data[3][:, 0]
    if soa:
        data = [d.T.copy() if d.ndim > 1 else d for d in data]
    p = Problem(maps, npts, data, (), block_size, soa)
    part = os.path.join(in_dir, "mesh_part")
    if os.path.exists(part):
        bs, p.partition_vector = read_partition(part, maps[0].shape[0])
        if bs != block_size:
            print(f"Warning: block size in file ({bs}) doesn't equal to used block size "
                  f"({block_size})", file=sys.stderr)
    return p


def reorder_dir(in_dir: str, out_dir: str, partition: bool, block_size: int) -> Problem:
    """res_calc.cu reorder(): GPS ordering, optionally partition +
    reorderToPartition + renumberPoints, then writeAllData."""
    p = read_airfoil_dir(in_dir, block_size)
    p.reorder("gps", same_mapping=[2, 3])
    if partition:
        p.partition(1.001)
        p.reorder_to_partition()
        p.renumber_points()
    write_airfoil_dir(p, out_dir, partition)
    return p


def rcm_edge_list(edges: np.ndarray, num_points: int):
    """scripts/rcm.py's transform (rcm.py:16-25) on an edge list: reverse
    Cuthill-McKee of the directed edge CSR (symmetrised), each edge's ids
    renumbered, the edges sorted lexicographically.  Returns (perm old -> new,
    new edges, (bandwidth, mean bandwidth) before, (the same) after)."""
    from . import plan
    e = np.asarray(edges, dtype=np.int64)
    order = np.lexsort((e[:, 1], e[:, 0]))
    ind = np.zeros(num_points + 1, dtype=np.int64)
    np.add.at(ind, e[:, 0] + 1, 1)
    ind = np.cumsum(ind)
    perm = plan.rcm_order(ind, e[order, 1], None)
    out = perm[e].astype(np.int64)
    out = out[np.lexsort((out[:, 1], out[:, 0]))]
    bw = lambda a: (int(np.max(np.abs(a[:, 0] - a[:, 1]))) if len(a) else 0,  # noqa: E731
                    float(np.mean(np.abs(a[:, 0] - a[:, 1]))) if len(a) else 0.0)
    return perm, out, bw(e), bw(out)


def rcm_tool(argv=None) -> int:
    """rcm.py's command line: <in_file> <out_file>, the mapping text format
    with MeshDim 2 in and out; prints the bandwidths as rcm.py does."""
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 2:
        print("Usage: python -m cuda_fot_amd.io rcm <in_file> <out_file>", file=sys.stderr)
        return 1
    edges, npts = read_mapping(argv[0], 2)
    _, out, before, after = rcm_edge_list(edges, npts)
    print("Original bandwidth:", before[0])
    print("Original mean bandwidth:", before[1])
    print("RCM'd bandwidth:", after[0])
    print("RCM'd mean bandwidth:", after[1])
    write_mapping(argv[1], out, npts)
    return 0


def reorder_tool(argv=None) -> int:
    """mainReorder (res_calc.cu:391-405):
    python -m cuda_fot_amd.io <input_dir> <output_dir_GPS> <output_dir_part> <block_size>"""
    argv = sys.argv if argv is None else argv
    if len(argv) < 5:
        print(f"Usage: {argv[0]} <input_dir> <output_dir_GPS> <output_dir_part> <block_size>",
              file=sys.stderr)
        return 1
    reorder_dir(argv[1], argv[2], False, int(argv[4]))
    reorder_dir(argv[1], argv[3], True, int(argv[4]))
    return 0


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "rcm":
        sys.exit(rcm_tool(sys.argv[2:]))
    sys.exit(reorder_tool())
