"""Python mirror of the reference's Problem<SOA> (problem.hpp:29-652) over
libfot: mesh mappings, point and cell data on the host, the reorder /
partition pre-passes (native planning in libfot), and the GPU loops.

Data layout follows the reference: point data k is (num_points[k], dim)
AOS or (dim, num_points[k]) SOA, point data 0 is the increment target; cell
data is SOA (dim, num_cells).
"""
from __future__ import annotations

import numpy as np

from . import capi, plan
from .device import Batch, DeviceArray, Loop

DEFAULT_BLOCK_SIZE = 128  # problem.hpp:16


def reorder_data(data: np.ndarray, perm: np.ndarray, soa: bool) -> np.ndarray:
    """reorderData<SOA> (data_t.hpp:117-146): new[perm[i]] = old[i]."""
    out = np.empty_like(data)
    if data.ndim == 1:
        out[perm] = data
    elif soa:
        out[:, perm] = data
    else:
        out[perm] = data
    return out


def reorder_data_inverse_cells(data: np.ndarray, inv: np.ndarray) -> np.ndarray:
    """reorderDataInverse<true> on SOA cell data: new cell i = old inv[i]."""
    return np.ascontiguousarray(data[..., inv])


class Problem:
    def __init__(self, maps, num_points, point_data, cell_data=(), block_size=DEFAULT_BLOCK_SIZE,
                 soa=False):
        self.maps = [capi.u32(m).copy() for m in maps]
        self.num_points = [int(n) for n in num_points]
        self.point_data = [None if d is None else np.array(d, dtype=np.float64) for d in point_data]
        self.cell_data = [np.array(c, dtype=np.float64) for c in cell_data]
        self.block_size = int(block_size)
        self.soa = bool(soa)
        self.partition_vector: np.ndarray | None = None
        self.point_partition: np.ndarray | None = None
        self.applied_permutation: np.ndarray | None = None

    @property
    def num_cells(self) -> int:
        return self.maps[0].shape[0]

    # ---- pre-passes --------------------------------------------------
    def reorder(self, ordering: str = "gps", same_mapping=None) -> None:
        """Problem::reorder (problem.hpp:439-482) with a native bandwidth
        ordering of mapping 0's point graph (GPS-style or RCM) in place of
        Scotch.  Mappings listed in same_mapping (default: those identical to
        mapping 0) take mapping 0's permutation (SameMapping); the others a
        first-touch renumbering (mesh.hpp:247-270)."""
        if same_mapping is None:
            same_mapping = [k for k in range(1, len(self.maps))
                            if np.array_equal(self.maps[k], self.maps[0])]
        ind, end = plan.graph_csr(self.maps[0], self.num_points[0])
        perm = plan.gps_order(ind, end) if ordering == "gps" else plan.rcm_order(ind, end)
        new0, inv = plan.mesh_reorder_cells(self.maps[0], perm)
        self.maps[0] = new0
        for k in range(1, len(self.maps)):
            self.maps[k] = self.maps[k][inv]
        self._reorder_point(0, perm)
        for k in range(1, len(self.maps)):
            if k in same_mapping:
                p = perm
            else:
                p = plan.point_renumbering_first_touch(self.maps[k], self.num_points[k])
            self.maps[k] = p[self.maps[k]]
            self._reorder_point(k, p)
        self.cell_data = [reorder_data_inverse_cells(c, inv) for c in self.cell_data]
        self._compose(perm)

    def partition(self, tolerance: float = 1.001, method: str = "kway") -> None:
        """Problem::partition (problem.hpp:484-499) with a native partitioner
        in place of METIS: greedy k-way growing ("kway") or recursive
        bisection into equal compact parts ("bisect")."""
        self.partition_vector = plan.partition_metis_enh(self.maps, self.num_points,
                                                         self.block_size, tolerance, method)

    def partition_ours(self, tolerance: float = 1.001, num_blocks_in_coarse: int = 8) -> None:
        """Problem::partitionOurs (problem.hpp:501-515)."""
        self.partition_vector = plan.partition_ours(self.maps, self.num_points, self.block_size,
                                                    num_blocks_in_coarse, tolerance)

    def partition_points(self, points_per_block: int, numbering: str = "cm",
                         method: str = "kway") -> None:
        """Owner partition for the owner-compute loop: parts of at most
        points_per_block mapping-0 points, then the mesh renumbered so every
        part's points are contiguous (mappings identical to mapping 0 follow),
        cells sorted by the part of their first point, and the other mappings
        renumbered by first touch.  Inside a part, points keep their relative
        order (numbering "id") or follow the Cuthill-McKee order of the point
        graph ("cm"), which gives neighbouring points nearby block-local ids."""
        if numbering not in ("id", "cm"):
            raise ValueError(f"numbering {numbering!r}: 'id' or 'cm'")
        ind, end = plan.graph_csr(self.maps[0], self.num_points[0])
        part = plan.PARTITIONERS[method](ind, end, points_per_block)[0]
        within = plan.gps_order(ind, end) if numbering == "cm" else np.arange(part.size)
        order = np.lexsort((within, part)).astype(np.uint32)
        perm = np.empty_like(order)
        perm[order] = np.arange(order.size, dtype=np.uint32)
        same = [k for k in range(1, len(self.maps)) if np.array_equal(self.maps[k], self.maps[0])]
        for k in [0] + same:
            self.maps[k] = perm[self.maps[k]]
            self._reorder_point(k, perm)
        self._compose(perm)
        self.point_partition = part[order]
        cell_key = self.point_partition[self.maps[0][:, 0]]
        inv = np.lexsort((np.arange(cell_key.size), cell_key)).astype(np.uint32)
        self.maps = [m[inv] for m in self.maps]
        self.cell_data = [reorder_data_inverse_cells(c, inv) for c in self.cell_data]
        for k in range(1, len(self.maps)):
            if k in same:
                continue
            p = plan.point_renumbering_first_touch(self.maps[k], self.num_points[k])
            self.maps[k] = p[self.maps[k]]
            self._reorder_point(k, p)

    def reorder_to_partition(self) -> None:
        """Problem::reorderToPartition (problem.hpp:517-526)."""
        part, inv = plan.reorder_to_partition(self.partition_vector)
        self.partition_vector = part
        self.maps = [m[inv] for m in self.maps]
        self.cell_data = [reorder_data_inverse_cells(c, inv) for c in self.cell_data]

    def renumber_points(self) -> None:
        """Problem::renumberPoints (problem.hpp:528-551)."""
        for k in range(len(self.maps)):
            perm = plan.point_renumbering_by_partition(self.maps[k], self.num_points[k],
                                                       self.partition_vector)
            self.maps[k] = perm[self.maps[k]]
            self._reorder_point(k, perm)
            if k == 0:
                self._compose(perm)

    def _reorder_point(self, k, perm):
        if self.point_data[k] is not None:
            self.point_data[k] = reorder_data(self.point_data[k], perm, self.soa)

    def _compose(self, perm):
        # applied_permutation[i]: current id of original point i
        if self.applied_permutation is None:
            self.applied_permutation = perm.astype(np.uint32).copy()
        else:
            self.applied_permutation = perm[self.applied_permutation]

    def target_in_original_order(self) -> np.ndarray:
        """Point data 0 indexed by the original point ids (the mapping the
        reference's tests apply through applied_permutation, res_calc.cu:153-170)."""
        d = self.point_data[0]
        if self.applied_permutation is None:
            return d.copy()
        return d[:, self.applied_permutation] if self.soa else d[self.applied_permutation]

    # ---- GPU loops ----------------------------------------------------
    def gpu_loop(self, func: capi.Functor, strategy: int, iters: int, block_size=None,
                 use_graph=False, flags=0, partition=True):
        """Runs `iters` iterations on the GPU and writes point data 0 back."""
        dev = [None if d is None else DeviceArray.from_host(d) for d in self.point_data]
        out = DeviceArray(self.point_data[0].shape, np.float64)
        cdev = [DeviceArray.from_host(c) for c in self.cell_data]
        part = self.partition_vector if (partition and strategy == capi.STRATEGY_HIERARCHICAL) else None
        if (flags & capi.LOOP_OWNER_COMPUTE) and partition and self.point_partition is not None:
            part = self.point_partition
            flags |= capi.LOOP_POINT_PARTITION
        loop = Loop(func, self.maps, self.num_points, dev, out, cdev, soa=self.soa,
                    strategy=strategy, block_size=block_size or self.block_size,
                    partition=part, use_graph=use_graph, flags=flags)
        loop.run(iters)
        self.point_data[0] = dev[0].download()
        slot = capi.cell_output_slot(func.kernel)
        if slot >= 0 and iters:
            self.cell_data[slot] = cdev[slot].download()
        stats = loop.stats()
        loop.close()
        return stats


def res_calc_problem(mesh: dict, block_size: int = DEFAULT_BLOCK_SIZE, soa: bool = False,
                     res0=None) -> Problem:
    """res_calc as res_calc.cu:10-40 sets it up: mappings res, x, q, adt
    (edge->cell, edge->node, edge->cell, edge->cell)."""
    nc, nn = mesh["num_cells"], mesh["num_nodes"]
    res = np.zeros((nc, 4)) if res0 is None else np.array(res0, dtype=np.float64)
    x, q, adt = mesh["x"], mesh["q"], mesh["adt"]
    if soa:
        res, x, q = res.T.copy(), x.T.copy(), q.T.copy()
    maps = [mesh["pecell"], mesh["pedge"], mesh["pecell"], mesh["pecell"]]
    return Problem(maps, [nc, nn, nc, nc], [res, x, q, adt], (), block_size, soa)


def volna_problem(mesh: dict, block_size: int = DEFAULT_BLOCK_SIZE, soa: bool = False) -> Problem:
    """The shallow-water edge loop as volna.cu:10-36 sets it up: two copies
    of the edge->cell map; point data state (dim 4, the target) and cell
    volume (dim 1); edge data fluxes, bathymetry source, normals and the
    boundary flag."""
    nc = mesh["num_cells"]
    state = mesh["state"].T.copy() if soa else mesh["state"].copy()
    maps = [mesh["edge_cells"], mesh["edge_cells"]]
    cell = [mesh["fluxes"], mesh["bathy"], mesh["normals"], mesh["is_boundary"][None, :]]
    return Problem(maps, [nc, nc], [state, mesh["volume"]], cell, block_size, soa)


def gpu_batch(problems, func: capi.Functor, iters: int, use_graph: bool = False) -> dict:
    """Runs `iters` iterations of every Problem in one batch (one persistent
    launch per step) and writes each point data 0 back.  Problems with a
    point partition (partition_points) use it; the others the engine's."""
    specs, keep = [], []
    for p in problems:
        dev = [None if d is None else DeviceArray.from_host(d) for d in p.point_data]
        out = DeviceArray(p.point_data[0].shape, np.float64)
        cdev = [DeviceArray.from_host(c) for c in p.cell_data]
        keep.append((dev, out, cdev))
        specs.append(dict(maps=p.maps, num_points=p.num_points, point_data=dev, out=out,
                          cell_data=cdev, soa=p.soa, block_size=p.block_size,
                          point_partition=p.point_partition))
    b = Batch(func, specs, use_graph=use_graph)
    b.run(iters)
    slot = capi.cell_output_slot(func.kernel)
    for p, (dev, _, cdev) in zip(problems, keep):
        p.point_data[0] = dev[0].download()
        if slot >= 0 and iters:
            p.cell_data[slot] = cdev[slot].download()
    stats = b.stats()
    b.close()
    return stats


def mini_aero_problem(mesh: dict, block_size: int = DEFAULT_BLOCK_SIZE,
                      soa: bool = False) -> Problem:
    """The face loop as mini_aero.cu sets it up: two copies of the
    face->cell map; point data the cell flux (dim 5, the target, zero) and the
    28 cell values; face data 12 values."""
    nc = mesh["num_cells"]
    flux = np.zeros((5, nc)) if soa else np.zeros((nc, 5))
    cells = mesh["cells"].T.copy() if soa else mesh["cells"].copy()
    maps = [mesh["face_cells"], mesh["face_cells"]]
    return Problem(maps, [nc, nc], [flux, cells], [mesh["face"]], block_size, soa)


def lulesh_problem(N1: int, N2: int, N3: int, seed: int = 0,
                   block_size: int = DEFAULT_BLOCK_SIZE, soa: bool = False,
                   shuffle: bool = True) -> Problem:
    """lulesh as lulesh.cu:10-36 sets it up, on synthetic data in place of
    its data_* files: fillCellList3D hexes of an N1 x N2 x N3 node lattice
    (cells shuffled), node (l, r, c) at (r, c, l) + U(-0.2, 0.2) per
    coordinate; two mappings, both the hexes; point data f (dim 3, N(0,1),
    the target) and xyz (dim 3); cell data sigma (3, N(0,1)) and the element
    volume (1, zero; the loop writes it)."""
    from . import meshgen
    cells = meshgen.hex_grid(N1, N2, N3)
    if shuffle:
        cells = cells[meshgen.permutation(cells.shape[0], seed + 1)]
    nc, npt = cells.shape[0], N1 * N2 * N3
    l, r, c = np.meshgrid(np.arange(N3), np.arange(N1), np.arange(N2), indexing="ij")
    xyz = np.stack([r.ravel(), c.ravel(), l.ravel()], axis=1).astype(np.float64)
    xyz += meshgen.uniform(3 * npt, seed + 2, -0.2, 0.2).reshape(npt, 3)
    f = meshgen.normal(3 * npt, seed + 3).reshape(npt, 3)
    sig = meshgen.normal(3 * nc, seed + 4).reshape(3, nc)
    determ = np.zeros((1, nc))
    if soa:
        f, xyz = f.T.copy(), xyz.T.copy()
    return Problem([cells, cells.copy()], [npt, npt], [f, xyz], [sig, determ], block_size, soa)


def getacc_problem(N: int, M: int, seed: int = 0, block_size: int = DEFAULT_BLOCK_SIZE,
                   soa: bool = False, shuffle: bool = True) -> Problem:
    """The corner scatter as getacc.cu:10-42 sets it up, on synthetic data:
    fillCellList quads (cells shuffled); point data ndmass / ndarea / ndub /
    ndvb (dim 4, zero); cell data cnmass U[0,1) with a third of the corners
    zero (so both fallbacks run), rho U[0.5,1.5), cnwt U[0,1), cnfx and cnfy
    N(0,1)."""
    from . import meshgen
    cells = meshgen.quad_grid(N, M)
    if shuffle:
        cells = cells[meshgen.permutation(cells.shape[0], seed + 1)]
    nc, np_ = cells.shape[0], N * M
    cnmass = meshgen.uniform(4 * nc, seed + 2).reshape(4, nc)
    cnmass[meshgen.uniform(4 * nc, seed + 3).reshape(4, nc) < 1 / 3] = 0.0
    rho = meshgen.uniform(nc, seed + 4, 0.5, 1.5).reshape(1, nc)
    cnwt = meshgen.uniform(4 * nc, seed + 5).reshape(4, nc)
    cnfx = meshgen.normal(4 * nc, seed + 6).reshape(4, nc)
    cnfy = meshgen.normal(4 * nc, seed + 7).reshape(4, nc)
    out = np.zeros((4, np_)) if soa else np.zeros((np_, 4))
    return Problem([cells], [np_], [out], [cnmass, rho, cnwt, cnfx, cnfy], block_size, soa)
