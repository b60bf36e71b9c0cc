"""Hand-derived known answers, shared by the oracle tests, the planted-mistake
tests and the product's own tests.  Each function takes the implementation
under test as a callable and raises AssertionError when an answer differs.
Every expected value below is worked out by hand in the comments from the
reference's rules, not computed by code."""
import numpy as np

MINE2, MINE4, MINE8 = 1, 2, 3  # oracle kernel ids (oracle.MINE2/4/8)


def _u(a):
    return np.array(a, dtype=np.uint32)


# ---- Mesh::colourCells (mesh.hpp:117-162, getAvailableColour :287-315) ----
def kat_global_colour_least_used(colour_cells):
    """Cells 0-2 are disjoint: colour 0.  Cell 3 (points 0, 2) meets colour 0,
    so a new colour 1.  Cell 4 is free of both; colour 1 holds one cell
    against three, so the least-used rule gives 1 where first fit gives 0."""
    col, n = colour_cells(_u([[0, 1], [2, 3], [4, 5], [0, 2], [6, 7]]), 8)
    assert list(col) == [0, 0, 0, 1, 1] and n == 2


# ---- HierarchicalColourMemory (colouring.hpp:44-147, 169-272) -----------
def kat_hplan_block_colour_least_used(hplan):
    """Four one-cell blocks (block_size 1, no partition).  Block 0 opens
    colour 0.  Block 1 shares point 1 with it: colour 1 opens.  Block 2 is
    free of both and both colours hold one block: the strict '>' of
    getAvailableColour (mesh.hpp:301) keeps the lower index, 0.  Block 3 is
    free of both and colour 1 (one block) is less used than colour 0 (two):
    1.  First fit would put block 3 in colour 0; a non-strict comparison
    block 2 in colour 1."""
    c = hplan([_u([[0, 1], [1, 2], [10, 11], [20, 21]])], 22, 1, None)
    assert len(c) == 2
    assert list(c[0]["cell_inv_permutation"]) == [0, 2]
    assert list(c[1]["cell_inv_permutation"]) == [1, 3]
    assert list(c[0]["block_offsets"]) == [0, 1, 2]
    assert list(c[1]["block_offsets"]) == [0, 1, 2]
    # each block caches its own two points, ascending
    assert list(c[0]["points_to_be_cached"]) == [0, 1, 10, 11]
    assert list(c[0]["points_to_be_cached_offsets"]) == [0, 2, 4]
    assert list(c[1]["points_to_be_cached"]) == [1, 2, 20, 21]


def kat_hplan_block_split(hplan):
    """Seven point-disjoint cells (cell i on points 2i, 2i+1), so every block
    takes colour 0 and the one colour's block_offsets list every block.
    Block size 3.

    Parts 0 0 1 1 1 1 2: cell 2 changes part, starts a block and restarts
    the count (colouring.hpp:89-95), so cells 2, 3, 4 fill it; cell 5 starts
    the next by size (:81-88); cell 6 changes part.  Blocks [0,2) [2,5)
    [5,6) [6,7).  Without the restart the count would reach 3 at cell 3:
    blocks [0,2) [2,3) [3,6) [6,7).

    Parts 1 0 1 0 0 2 1: the stable sort by part orders cells 1 3 4 | 0 2 6
    | 5.  The size split at position 3 coincides with the part change and
    opens one block, not two (the `continue` at :87); position 6 changes
    part.  Blocks [0,3) [3,6) [6,7)."""
    maps = [_u([[2 * i, 2 * i + 1] for i in range(7)])]
    c = hplan(maps, 14, 3, _u([0, 0, 1, 1, 1, 1, 2]))
    assert len(c) == 1
    assert list(c[0]["block_offsets"]) == [0, 2, 5, 6, 7]
    assert list(c[0]["cell_inv_permutation"]) == [0, 1, 2, 3, 4, 5, 6]
    c = hplan(maps, 14, 3, _u([1, 0, 1, 0, 0, 2, 1]))
    assert len(c) == 1
    assert list(c[0]["block_offsets"]) == [0, 3, 6, 7]
    assert list(c[0]["cell_inv_permutation"]) == [1, 3, 4, 0, 2, 6, 5]
    assert list(c[0]["num_cell_colours"]) == [1, 1, 1]
    # block 0 (cells 1, 3, 4) caches points 2 3 6 7 8 9
    assert list(c[0]["points_to_be_cached"][:6]) == [2, 3, 6, 7, 8, 9]


# a path of 8 cells through points visited in this order
PATH = [7, 3, 9, 1, 5, 8, 0, 6, 2]


def kat_hplan_thread_order(hplan):
    """One block of 8 cells forming a path: cell k joins points PATH[k] and
    PATH[k+1].  Mapping 1 gives cell k the id 100 + k.

    Cached points ascending (the std::map of colouring.hpp:184-208):
    0 1 2 3 5 6 7 8 9, so point -> local id 0->0 1->1 2->2 3->3 5->4 6->5
    7->6 8->7 9->8.  Cells' local ids: 0 (6,3), 1 (3,8), 2 (8,1), 3 (1,4),
    4 (4,7), 5 (7,0), 6 (0,5), 7 (5,2).

    Thread colours (colouring.cu:12-102): end cells have degree 1, inner
    cells 2.  Smallest-last takes the last minimum: cell 7, then 6 (now
    degree 1, tied with cell 0), ..., 0.  First fit in reverse of that order
    colours cell 0 first: 0, cell 1: 1, cell 2: 0, ...: colour = k mod 2.

    Cells sorted by (thread colour, position in block) (colouring.hpp:
    246-272): 0 2 4 6 1 3 5 7."""
    maps = [_u([[PATH[k], PATH[k + 1]] for k in range(8)]), _u([[100 + k] for k in range(8)])]
    c = hplan(maps, 10, 8, None)
    assert len(c) == 1
    b = c[0]
    assert list(b["points_to_be_cached"]) == [0, 1, 2, 3, 5, 6, 7, 8, 9]
    assert list(b["points_to_be_cached_offsets"]) == [0, 9]
    assert list(b["cell_inv_permutation"]) == [0, 2, 4, 6, 1, 3, 5, 7]
    assert list(b["cell_colours"]) == [0, 0, 0, 0, 1, 1, 1, 1]
    assert list(b["num_cell_colours"]) == [2]
    assert list(b["block_offsets"]) == [0, 8]
    # local ids in plan order, SOA (first ids of the 8 cells, then second):
    # rows (6,3) (8,1) (4,7) (0,5) (3,8) (1,4) (7,0) (5,2)
    assert list(b["cell_list"][0]) == [6, 8, 4, 0, 3, 1, 7, 5, 3, 1, 7, 5, 8, 4, 0, 2]
    assert list(b["cell_list"][1]) == [100, 102, 104, 106, 101, 103, 105, 107]


# ---- mine (kernels/mine_func.hpp:11-99, loop problem.hpp:342-380) -------
def kat_mine2(run):
    """run(kernel, maps, num_points, point_data, cell_data, iters, soa) ->
    point data 0 after `iters` iterations.  One edge (0, 1), L = 1 2 3 4,
    R = 10 20 30 40.  out_L = L + 2 R w, out_R = R + 2 L w."""
    L, R = [1.0, 2.0, 3.0, 4.0], [10.0, 20.0, 30.0, 40.0]
    e = [_u([[0, 1]])]
    # CellDim = PointDim, w = 1: L + 2R, R + 2L
    t = run(MINE2, e, [2], np.array([L, R]), [np.ones((4, 1))], 1, False)
    assert np.asarray(t).tolist() == [[21, 42, 63, 84], [12, 24, 36, 48]]
    # CellDim 1, w = 0.25: L + R / 2, R + L / 2
    t = run(MINE2, e, [2], np.array([L, R]), [np.full((1, 1), 0.25)], 1, False)
    assert np.asarray(t).tolist() == [[6, 12, 18, 24], [10.5, 21, 31.5, 42]]
    # SOA (d * n + i): the same numbers
    t = run(MINE2, e, [2], np.array([L, R]).T.copy(), [np.ones((4, 1))], 1, True)
    assert np.asarray(t).T.tolist() == [[21, 42, 63, 84], [12, 24, 36, 48]]
    # two iterations read the previous iterate (Jacobi): 1, 10 -> 1 + 20,
    # 10 + 2 = 21, 12 -> 21 + 24, 12 + 42 = 45, 54
    t = run(MINE2, e, [2], np.array([[1.0], [10.0]]), [np.ones((1, 1))], 2, False)
    assert np.asarray(t).ravel().tolist() == [45, 54]


def kat_mine4(run):
    """One quad (0, 1, 2, 3), p = 1 10 100 1000, w = 1: out_k = p_(k+1) +
    p_(k-1): 1010, 101, 1010, 101 -> 1011 111 1110 1101."""
    t = run(MINE4, [_u([[0, 1, 2, 3]])], [4], np.array([[1.0], [10.0], [100.0], [1000.0]]),
            [np.ones((1, 1))], 1, False)
    assert np.asarray(t).ravel().tolist() == [1011, 111, 1110, 1101]


def kat_mine8(run):
    """One hex (0..7), p_j = 2^j, w = 1: out_k = p_(k+1) + p_(k-1): 130 5 10
    20 40 80 160 65 -> 131 7 14 28 56 112 224 193."""
    p = np.array([[2.0 ** j] for j in range(8)])
    t = run(MINE8, [_u([list(range(8))])], [8], p, [np.ones((1, 1))], 1, False)
    assert np.asarray(t).ravel().tolist() == [131, 7, 14, 28, 56, 112, 224, 193]


HPLAN_KATS = [kat_hplan_block_colour_least_used, kat_hplan_block_split, kat_hplan_thread_order]
MINE_KATS = [kat_mine2, kat_mine4, kat_mine8]
