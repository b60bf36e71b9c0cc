"""Pins the CPU oracle: the reference's own known-answer tests
(test/catch_priority_queue.cu, test/catch_heuristical_partition.cu),
hand-derived known answers that fix the reference's tie-breaking rules, and
closed forms of the res_calc flux.  Each check is chosen so that a plausible
mistake (wrong tie-break, sign, index or operand) fails it."""
import numpy as np
import pytest

import oracle as O


# ---- PriorityQueue (test/catch_priority_queue.cu) ----------------------
def test_pq_push_and_get():
    # push one / two / two with the same priority (:17-49)
    assert O.pq_run([(0, 0, 1), (3, 0, 0)]) == [1, 1]
    assert O.pq_run([(0, 0, 1), (0, 2, 3), (3, 0, 0), (3, 2, 0), (3, 1, 0), (3, 3, 0)]) == \
        [1, 1, 1, 3, 0, 0, 0, 0]
    assert O.pq_run([(0, 0, 1), (0, 2, 1), (3, 0, 0), (3, 2, 0), (3, 1, 0)]) == \
        [1, 1, 1, 1, 0, 0]


def test_pq_pop_max():
    # (:51-81)
    assert O.pq_run([(0, 0, 1), (1, 0, 0), (3, 0, 0)]) == [0, 1, 0, 0]
    assert O.pq_run([(0, 0, 1), (0, 2, 3), (1, 0, 0), (3, 0, 0), (3, 2, 0), (1, 0, 0),
                     (3, 0, 0), (3, 2, 0)]) == [2, 3, 1, 1, 0, 0, 0, 1, 0, 0, 0, 0]


def test_pq_priority_sort():
    # (:83-96): popMax yields priorities in descending order
    v = [1, 9, 8, 0, 7, 2, 6, 4, 3, 11, 12, 13, 10, 5]
    out = O.pq_run([(0, i, p) for i, p in enumerate(v)] + [(1, 0, 0)] * len(v))
    assert out[1::2] == sorted(v, reverse=True)


def test_pq_modify():
    # increase / decrease / reverse order (:98-138)
    base = [(0, 0, 1), (0, 2, 3)]
    assert O.pq_run(base + [(2, 2, 4), (3, 2, 0), (3, 0, 0), (1, 0, 0)]) == [1, 4, 1, 1, 2, 4]
    assert O.pq_run(base + [(2, 0, -1), (3, 2, 0), (3, 0, 0), (1, 0, 0)]) == [1, 3, 1, -1, 2, 3]
    assert O.pq_run(base + [(2, 0, 4), (3, 2, 0), (3, 0, 0), (1, 0, 0)]) == [1, 3, 1, 4, 0, 4]


def test_pq_equal_priorities_pop_last_inserted():
    # multimap keeps equal keys in insertion order and popMax takes the last
    # element (priority_queue.hpp:39-49)
    assert O.pq_run([(0, 5, 7), (0, 1, 7), (0, 9, 7), (1, 0, 0), (1, 0, 0), (1, 0, 0)]) == \
        [9, 7, 1, 7, 5, 7]


# ---- GraphCSR (test/catch_heuristical_partition.cu:20-51) --------------
EDGES_4 = np.array([[0, 1], [0, 2], [1, 2], [2, 3]], dtype=np.uint32)


def _reorder_inverse(ind, end, inv):
    # relabel as GraphCSR::reorderInverse (reorder.hpp:177-202) does
    perm = np.empty_like(inv)
    perm[inv] = np.arange(inv.size)
    rows = [sorted(perm[end[ind[o]:ind[o + 1]]].tolist()) for o in inv]
    return rows


def test_graph_csr_sanity_and_reorder_inverse():
    ind, end = O.graph_csr(EDGES_4, 4)
    assert np.all(np.diff(ind) >= 0)
    rows = _reorder_inverse(ind, end, np.array([3, 1, 0, 2]))
    assert rows == [[3], [2, 3], [1, 3], [0, 1, 2]]


def test_graph_csr_dedup_within_row_only():
    # duplicates are removed only against the current row (reorder.hpp:66-71):
    # two cells on the same point pair give one arc per row
    ind, end = O.graph_csr(np.array([[0, 1], [1, 0], [1, 2]], dtype=np.uint32), 3)
    assert end[ind[0]:ind[1]].tolist() == [1]
    assert end[ind[1]:ind[2]].tolist() == [0, 2]


# ---- HeuristicalPartition (test/catch_heuristical_partition.cu:53-106) --
def _cell_csr(cells, npts):
    return O.cell_to_cell_csr([np.array(cells, dtype=np.uint32)], [npts])


def test_heuristic_two_small_segments():
    cells = [[4, 1], [3, 2], [0, 3], [1, 5]]
    ind, end = _cell_csr(cells, 6)
    # METIS splits the two components {0,3} and {1,2}; the reference's
    # expected labels follow from coarse part 0 holding cell 0
    res, nb = O.heuristic_partition(ind, end, [0, 1, 1, 0], 2)
    assert res.tolist() == [0, 1, 1, 0] and nb == 2
    res, nb = O.heuristic_partition(ind, end, [0, 1, 1, 0], 1)
    assert len(set(res.tolist())) == 4


TRIVIAL = [[1, 3], [4, 3], [1, 2], [3, 0], [5, 7], [1, 5], [6, 4], [6, 0], [2, 7],
           [9, 11], [12, 11], [9, 10], [11, 8], [13, 15], [9, 13], [14, 12], [14, 8],
           [10, 15]]
GROUPS = [[1, 3, 6, 7], [2, 4, 5, 8], [10, 12, 15, 16], [11, 13, 14, 17]]


@pytest.mark.parametrize("coarse", [[0] * 9 + [1] * 9, [1] * 9 + [0] * 9])
def test_heuristic_trivial_solution(coarse):
    ind, end = _cell_csr(TRIVIAL, 16)
    res, nb = O.heuristic_partition(ind, end, coarse, 5)
    assert nb == 4
    for g in GROUPS:
        assert len({int(res[c]) for c in g}) == 1, g


# ---- global colouring (mesh.hpp:117-162) --------------------------------
def test_colour_cells_path_known_answer():
    col, n = O.colour_cells(np.array([[0, 1], [1, 2], [2, 3], [3, 4]]), 5)
    assert col.tolist() == [0, 1, 0, 1] and n == 2


def test_colour_cells_least_used_rule():
    # the 5th cell may take colour 0 or 1: the least used one (1) wins,
    # first fit would give 0 (mesh.hpp:298-303)
    col, n = O.colour_cells(np.array([[0, 1], [2, 3], [4, 5], [0, 2], [6, 7]]), 8)
    assert col.tolist() == [0, 0, 0, 1, 1] and n == 2


def test_colour_cells_overflow():
    star = np.array([[0, i + 1] for i in range(65)])
    with pytest.raises(RuntimeError):
        O.colour_cells(star, 66)


# ---- in-block colouring (colouring.cu:12-102) ---------------------------
def test_colour_with_ordering_path():
    col, n = O.colour_cells_with_ordering(np.array([[0, 1], [1, 2], [2, 3]]))
    assert col.tolist() == [0, 1, 0] and n == 2


def test_colour_with_ordering_tie_goes_to_last_index():
    # smallest-last order picks the LAST minimum-degree cell (colouring.cu:27):
    # order 3,2,1,0 then reverse first fit gives [0,1,2,0]; taking the first
    # minimum instead would give [2,1,0,1]
    col, n = O.colour_cells_with_ordering(np.array([[0, 1], [0, 2], [0, 3], [3, 4]]))
    assert col.tolist() == [0, 1, 2, 0] and n == 3


def test_colour_with_ordering_duplicate_neighbours():
    # two cells sharing both points are neighbours twice; still two colours
    col, n = O.colour_cells_with_ordering(np.array([[0, 1], [1, 0], [2, 3]]))
    assert col[0] != col[1] and n == 2


# ---- reorder helpers (mesh.hpp) ------------------------------------------
def test_reorder_to_partition_known_answer():
    part, inv = O.reorder_to_partition([2, 0, 1, 0])
    assert part.tolist() == [0, 0, 1, 2] and inv.tolist() == [1, 3, 2, 0]


def test_first_touch_known_answer():
    perm = O.point_renumbering_first_touch(np.array([[2, 0], [0, 3]]), 5)
    assert perm.tolist() == [1, 3, 0, 2, 4]


def test_renumbering_by_partition_known_answer():
    # points touched by more parts first, then by the larger part list
    # (mesh.hpp:373-385); ties keep index order (stable sort)
    cells = np.array([[0, 1], [1, 2], [2, 3], [4, 5]])
    part = np.array([0, 1, 2, 2])
    # point parts: 0:{0} 1:{0,1} 2:{1,2} 3:{2} 4:{2} 5:{2}
    # order: 2 ({1,2}), 1 ({0,1}), then singletons by descending list: 3,4,5 ({2}), 0 ({0})
    perm = O.point_renumbering_by_partition(cells, 6, part)
    inv = np.argsort(perm)
    assert inv.tolist() == [2, 1, 3, 4, 5, 0]


def test_mesh_reorder_cells_known_answer():
    cells = np.array([[3, 0], [1, 2], [0, 1]])
    perm = np.array([3, 2, 1, 0])  # point p -> 3 - p
    new, inv = O.mesh_reorder_cells(cells, perm)
    # renumbered cells: (0,3) (2,1) (3,2); sorted tuples (0,3) (1,2) (2,3)
    assert inv.tolist() == [0, 1, 2]
    assert new.tolist() == [[0, 3], [2, 1], [3, 2]]
    new, inv = O.mesh_reorder_cells(np.array([[2, 3], [0, 1]]), np.arange(4))
    assert inv.tolist() == [1, 0]


def test_cell_to_cell_pairs_known_answer():
    pairs = O.cell_to_cell_pairs([np.array([[0, 1], [1, 2], [3, 4]])], [5])
    assert pairs.tolist() == [[0, 1]]
    # a second mapping adds its own sharing
    pairs = O.cell_to_cell_pairs([np.array([[0, 1], [1, 2], [3, 4]]),
                                  np.array([[0], [5], [0]])], [5, 6])
    assert pairs.tolist() == [[0, 1], [0, 2]]


# ---- res_calc user function (kernels/res_calc_func.hpp:11-57) -----------
GM1, EPS = 0.4, 0.05


def test_res_calc_antisymmetric():
    rng = np.random.default_rng(0)
    for _ in range(20):
        r1, r2 = O.res_calc_func(rng.random(2), rng.random(2), 1 + rng.random(4),
                                 1 + rng.random(4), rng.random(), rng.random())
        assert np.array_equal(r2, -r1)


def test_res_calc_degenerate_edge_is_pure_dissipation():
    # dx = dy = 0: no convective flux, f = mu (q1 - q2), mu = (adt1+adt2)/2 * eps
    q1 = np.array([1.1, 0.3, -0.2, 2.4])
    q2 = np.array([0.9, 0.5, 0.1, 2.6])
    r1, _ = O.res_calc_func([0.5, 0.5], [0.5, 0.5], q1, q2, 0.7, 1.3)
    mu = 0.5 * (0.7 + 1.3) * EPS
    assert np.array_equal(r1, mu * (q1 - q2))


def test_res_calc_uniform_state_is_euler_flux():
    # q1 == q2: the dissipation vanishes and the flux through the face normal
    # n = (dy, -dx) is the Euler flux of the state
    rho, u, v, E = 1.2, 0.4, -0.3, 2.5
    q = np.array([rho, rho * u, rho * v, E])
    x1, x2 = np.array([1.3, 0.7]), np.array([0.2, -0.4])
    dx, dy = x1 - x2
    p = GM1 * (E - 0.5 * rho * (u * u + v * v))
    un = u * dy - v * dx
    expect = np.array([rho * un, rho * u * un + p * dy, rho * v * un - p * dx, (E + p) * un])
    r1, r2 = O.res_calc_func(x1, x2, q, q, 0.8, 0.9)
    np.testing.assert_allclose(r1, expect, rtol=1e-14, atol=1e-15)
