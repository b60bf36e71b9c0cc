"""The recursive-bisection partitioner (fot_bisect_partition), a second
stand-in for METIS: a hand-derived known answer on a path, and on grids and
a disconnected graph the properties that define it (ceil(n / max_part)
parts whose sizes differ by at most one, each vertex in exactly one part).
A part is a prefix or suffix of a breadth-first order: the prefix of one is
connected, a suffix need not be, so connectivity is not claimed."""
import numpy as np
import pytest

from cuda_fot_amd import meshgen, plan


def _csr(n, edges):
    nb = [[] for _ in range(n)]
    for a, b in edges:
        nb[a].append(b)
        nb[b].append(a)
    ind = np.zeros(n + 1, dtype=np.uint32)
    ind[1:] = np.cumsum([len(x) for x in nb])
    end = np.array([w for x in nb for w in sorted(x)], dtype=np.uint32)
    return ind, end


def test_path_known_answer():
    # 0-1-...-7 into parts of 2: the sweep from 0 ends at 7, the level order
    # from 7 is 7..0 and the cut at half gives {7,6,5,4} (parts 0-1) and
    # {3,2,1,0} (parts 2-3); {7,6,5,4} sweeps from 7 to 4 and orders 4,5,6,7
    # (parts 0, 1), {3,2,1,0} sweeps from 3 to 0 and orders 0,1,2,3 (2, 3)
    ind, end = _csr(8, [(i, i + 1) for i in range(7)])
    part, n = plan.bisect_partition(ind, end, 2)
    assert n == 4
    np.testing.assert_array_equal(part, [2, 2, 3, 3, 0, 0, 1, 1])


@pytest.mark.parametrize("nx,ny,maxp", [(4, 4, 4), (13, 9, 10), (30, 20, 64)])
def test_grid_parts_equal(nx, ny, maxp):
    idx = lambda i, j: i * ny + j
    edges = [(idx(i, j), idx(i + 1, j)) for i in range(nx - 1) for j in range(ny)] + \
            [(idx(i, j), idx(i, j + 1)) for i in range(nx) for j in range(ny - 1)]
    n = nx * ny
    ind, end = _csr(n, edges)
    part, k = plan.bisect_partition(ind, end, maxp)
    assert k == -(-n // maxp)
    sizes = np.bincount(part, minlength=k)
    assert sizes.size == k and sizes.max() <= maxp and sizes.max() - sizes.min() <= 1


def test_hex_cells_full_parts():
    # the cells of a shuffled hex grid (cells sharing a point are neighbours)
    cells = meshgen.hex_grid(9, 8, 7)
    cells = cells[meshgen.permutation(cells.shape[0], 1)]
    ind, end = plan.cell_to_cell_csr([cells], [int(cells.max()) + 1])
    part, k = plan.bisect_partition(ind, end, 50)
    sizes = np.bincount(part, minlength=k)
    assert k == -(-cells.shape[0] // 50) and sizes.max() <= 50 and sizes.max() - sizes.min() <= 1
    # compact: a part of 50 cells spans few points (a 4x4x3 brick has 100)
    per_part = [np.unique(cells[part == q]).size for q in range(k)]
    assert np.mean(per_part) < 140, np.mean(per_part)


def test_disconnected_graph():
    # two paths of 5 and an isolated vertex: every vertex placed once
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (5, 6), (6, 7), (7, 8), (8, 9)]
    ind, end = _csr(11, edges)
    part, k = plan.bisect_partition(ind, end, 3)
    assert k == 4
    sizes = np.bincount(part, minlength=k)
    assert sizes.sum() == 11 and sizes.max() <= 3
