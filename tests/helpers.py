"""Shared builders and checks for the parity tests."""
import numpy as np

import oracle as O
from cuda_fot_amd import capi, meshgen, plan, res_calc_problem

TOL = 1e-12  # relative to max(|oracle|, S): S = |initial| + sum |increments|


def assert_close_scaled(gpu, ref, scale, tol=TOL, what=""):
    gpu, ref, scale = np.asarray(gpu), np.asarray(ref), np.asarray(scale)
    bound = tol * np.maximum(np.abs(ref), scale)
    err = np.abs(gpu - ref)
    bad = err > bound
    if bad.any():
        i = np.argmax(err / np.maximum(bound, 1e-300))
        raise AssertionError(f"{what}: {bad.sum()} components beyond {tol}*scale; worst "
                             f"flat index {i}: gpu {gpu.flat[i]!r} ref {ref.flat[i]!r} "
                             f"scale {scale.flat[i]!r}")


def assert_bitwise(gpu, ref, what=""):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    if not np.array_equal(gpu.view(np.uint64), ref.view(np.uint64)):
        diff = gpu != ref
        i = int(np.flatnonzero(diff.ravel())[0])
        raise AssertionError(f"{what}: {int(diff.sum())} components differ; first flat index "
                             f"{i}: gpu {gpu.flat[i]!r} ref {ref.flat[i]!r}")


def oracle_of(p, kernel=O.RES_CALC):
    """An oracle problem over the same arrays as a cuda_fot_amd Problem."""
    return O.OracleProblem(kernel, p.maps, p.num_points, p.point_data, p.cell_data, soa=p.soa)


def airfoil_problem(nx=48, ny=24, seed=0, soa=False, block_size=128, res0_seed=None):
    m = meshgen.airfoil(nx, ny, seed=seed)
    res0 = None
    if res0_seed is not None:
        res0 = meshgen.uniform(m["num_cells"] * 4, res0_seed, -1.0, 1.0).reshape(-1, 4)
    return res_calc_problem(m, block_size=block_size, soa=soa, res0=res0)


def mine_problem(N=24, M=20, P=4, C=4, seed=0, soa=False, block_size=128, shuffle=True):
    """config-1 style: fillEdgeList grid, point data and weights U[0,1)."""
    e = meshgen.grid_edges(N, M)
    if shuffle:
        e = e[meshgen.permutation(e.shape[0], seed + 11)]
    n = N * M
    pd = meshgen.uniform(n * P, seed + 1).reshape(n, P)
    w = meshgen.uniform(e.shape[0] * C, seed + 2).reshape(C, e.shape[0])
    if soa:
        pd = pd.T.copy()
    from cuda_fot_amd import Problem
    return Problem([e], [n], [pd], [w], block_size, soa)


def hex_problem(n=6, seed=0, soa=False, block_size=128):
    cells = meshgen.hex_grid(n + 1, n + 1, n + 1)
    cells = cells[meshgen.permutation(cells.shape[0], seed + 5)]
    nn = (n + 1) ** 3
    f0 = np.zeros((nn, 3)) if not soa else np.zeros((3, nn))
    forces = meshgen.normal(cells.shape[0] * 24, seed + 6).reshape(24, cells.shape[0])
    from cuda_fot_amd import Problem
    return Problem([cells], [nn], [f0], [forces], block_size, soa)
