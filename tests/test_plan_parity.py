"""The product's native planner (libfot, host code) against the oracle:
colourings, hierarchical plans, partitions and permutations must be
bit-identical; plans must also satisfy the validity invariants the GPU
kernels rely on.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

import oracle as O
from cuda_fot_amd import meshgen, plan
from helpers import airfoil_problem, hex_problem, mine_problem


def _meshes():
    a = airfoil_problem(nx=40, ny=24, seed=3)
    yield "airfoil", a.maps, a.num_points
    m = mine_problem(N=20, M=16, seed=4)
    yield "grid", m.maps, m.num_points
    h = hex_problem(n=5, seed=5)
    yield "hex", h.maps, h.num_points
    u = mine_problem(N=18, M=18, seed=6, shuffle=False)
    yield "grid-ordered", u.maps, u.num_points


MESHES = list(_meshes())


@pytest.mark.parametrize("name,maps,npts", MESHES, ids=[m[0] for m in MESHES])
def test_global_colouring_matches_oracle(name, maps, npts):
    col, n = plan.colour_cells(maps[0], npts[0])
    ocol, on = O.colour_cells(maps[0], npts[0])
    assert n == on and np.array_equal(col, ocol)
    # validity: no two cells of a colour share a point
    for c in range(n):
        pts = maps[0][col == c].ravel()
        assert np.unique(pts).size == pts.size


def test_colour_with_ordering_matches_oracle_random_blocks():
    rng = np.random.default_rng(7)
    for trial in range(40):
        nc = int(rng.integers(1, 200))
        md = int(rng.choice([2, 3, 4, 8]))
        npts = int(rng.integers(2, 3 * nc + 2))
        local = rng.integers(0, npts, size=(nc, md)).astype(np.uint32)
        col, n = plan.colour_cells_with_ordering(local)
        ocol, on = O.colour_cells_with_ordering(local)
        assert n == on and np.array_equal(col, ocol), trial


def _plans_equal(a, b):
    assert len(a) == len(b)
    for ca, cb in zip(a, b):
        for k in ca:
            if k == "cell_list":
                for x, y in zip(ca[k], cb[k]):
                    assert np.array_equal(x, y)
            elif k == "shared_size":
                assert ca[k] == cb[k]
            else:
                assert np.array_equal(ca[k], cb[k]), k


def _check_plan_valid(colours, maps, block_size):
    seen = []
    md = maps[0].shape[1]
    for c in colours:
        n = c["cell_colours"].size
        cl0 = c["cell_list"][0].reshape(md, n).T
        owner = {}
        bo, po = c["block_offsets"], c["points_to_be_cached_offsets"]
        for b in range(len(bo) - 1):
            assert 0 < bo[b + 1] - bo[b] <= block_size
            cached = c["points_to_be_cached"][po[b]:po[b + 1]]
            assert np.all(np.diff(cached.astype(np.int64)) > 0)  # ascending, unique
            cells = c["cell_inv_permutation"][bo[b]:bo[b + 1]]
            loc = cl0[bo[b]:bo[b + 1]]
            assert np.array_equal(cached[loc], maps[0][cells])  # local ids map back
            tc = c["cell_colours"][bo[b]:bo[b + 1]]
            assert np.all(np.diff(tc.astype(int)) >= 0)  # sorted by thread colour
            assert tc.max() + 1 == c["num_cell_colours"][b]
            for col in range(int(tc.max()) + 1):
                pts = loc[tc == col].ravel()
                assert np.unique(pts).size == pts.size  # thread colouring valid
            for p in cached:  # blocks of one block colour are point-disjoint
                assert owner.setdefault(int(p), b) == b
        seen.append(c["cell_inv_permutation"])
    allc = np.concatenate(seen)
    assert np.array_equal(np.sort(allc), np.arange(maps[0].shape[0]))


@pytest.mark.parametrize("name,maps,npts", MESHES, ids=[m[0] for m in MESHES])
@pytest.mark.parametrize("bs", [32, 128])
def test_hplan_matches_oracle_chunks(name, maps, npts, bs):
    prod = plan.hplan(maps, npts[0], bs)
    ora = O.OraclePlan(maps, npts[0], bs).colours()
    _plans_equal(prod, ora)
    _check_plan_valid(prod, maps, bs)


@pytest.mark.parametrize("bs", [64, 100])
def test_hplan_matches_oracle_partitioned(bs):
    p = airfoil_problem(nx=40, ny=24, seed=8, block_size=bs)
    p.reorder("gps")
    p.partition(1.001)
    p.reorder_to_partition()
    p.renumber_points()
    prod = plan.hplan(p.maps, p.num_points[0], bs, p.partition_vector)
    ora = O.OraclePlan(p.maps, p.num_points[0], bs, p.partition_vector).colours()
    _plans_equal(prod, ora)
    _check_plan_valid(prod, p.maps, bs)
    # a random partition with parts larger than the block size splits blocks
    rng = np.random.default_rng(bs)
    part = rng.integers(0, 7, size=p.num_cells).astype(np.uint32)
    _plans_equal(plan.hplan(p.maps, p.num_points[0], bs, part),
                 O.OraclePlan(p.maps, p.num_points[0], bs, part).colours())


@pytest.mark.parametrize("name,maps,npts", MESHES, ids=[m[0] for m in MESHES])
def test_renumberings_match_oracle(name, maps, npts):
    rng = np.random.default_rng(9)
    part = rng.integers(0, 13, size=maps[0].shape[0]).astype(np.uint32)
    assert np.array_equal(plan.reorder_to_partition(part)[1], O.reorder_to_partition(part)[1])
    for k, m in enumerate(maps):
        a = plan.point_renumbering_by_partition(m, npts[k], part)
        b = O.point_renumbering_by_partition(m, npts[k], part)
        assert np.array_equal(a, b)
        assert np.array_equal(np.sort(a), np.arange(npts[k]))
        assert np.array_equal(plan.point_renumbering_first_touch(m, npts[k]),
                              O.point_renumbering_first_touch(m, npts[k]))
    perm = meshgen.permutation(npts[0], 5)
    a_map, a_inv = plan.mesh_reorder_cells(maps[0], perm)
    b_map, b_inv = O.mesh_reorder_cells(maps[0], perm)
    assert np.array_equal(a_map, b_map) and np.array_equal(a_inv, b_inv)


@pytest.mark.parametrize("name,maps,npts", MESHES, ids=[m[0] for m in MESHES])
def test_graphs_match_oracle(name, maps, npts):
    ia, ea = plan.graph_csr(maps[0], npts[0])
    ib, eb = O.graph_csr(maps[0], npts[0])
    assert np.array_equal(ia, ib) and np.array_equal(ea, eb)
    ia, ea = plan.cell_to_cell_csr(maps, npts)
    ib, eb = O.cell_to_cell_csr(maps, npts)
    assert np.array_equal(ia, ib) and np.array_equal(ea, eb)


def test_heuristic_partition_matches_oracle():
    p = airfoil_problem(nx=30, ny=20, seed=10)
    ind, end = plan.cell_to_cell_csr(p.maps, p.num_points)
    rng = np.random.default_rng(11)
    for coarse_size, bs in [(64, 16), (200, 50), (1000, 128)]:
        coarse, _ = plan.kway_partition(ind, end, coarse_size)
        a, na = plan.heuristic_partition(ind, end, coarse, bs)
        b, nb = O.heuristic_partition(ind, end, coarse, bs)
        assert na == nb and np.array_equal(a, b)
        sizes = np.bincount(a)
        assert sizes.max() <= bs
    coarse = rng.integers(0, 5, size=p.num_cells).astype(np.uint32)
    a, _ = plan.heuristic_partition(ind, end, coarse, 40)
    b, _ = O.heuristic_partition(ind, end, coarse, 40)
    assert np.array_equal(a, b)


def test_heuristic_partition_reference_kats_on_product():
    ind, end = O.cell_to_cell_csr([np.array([[4, 1], [3, 2], [0, 3], [1, 5]])], [6])
    res, _ = plan.heuristic_partition(ind, end, [0, 1, 1, 0], 2)
    assert res.tolist() == [0, 1, 1, 0]
    from test_oracle_kats import GROUPS, TRIVIAL
    ind, end = O.cell_to_cell_csr([np.array(TRIVIAL)], [16])
    res, nb = plan.heuristic_partition(ind, end, [0] * 9 + [1] * 9, 5)
    assert nb == 4
    for g in GROUPS:
        assert len({int(res[c]) for c in g}) == 1


def test_kway_partition_valid():
    p = airfoil_problem(nx=50, ny=30, seed=12)
    ind, end = plan.cell_to_cell_csr(p.maps, p.num_points)
    for maxp in (7, 64, 511):
        part, n = plan.kway_partition(ind, end, maxp)
        assert part.max() + 1 == n
        assert np.bincount(part).max() <= maxp
        assert n == -(-p.num_cells // maxp) or n > 0


def _csr_bandwidth(ind, end, perm):
    rows = np.repeat(np.arange(ind.size - 1), np.diff(ind))
    return int(np.max(np.abs(perm[rows].astype(np.int64) - perm[end].astype(np.int64))))


def _scipy_rcm(edges, n):
    """scripts/rcm.py:16-21: the directed CSR of an edge list, scipy's
    reverse_cuthill_mckee (which symmetrises it), then argsort: old -> new.
    Also returns the CSR and the seed order scipy's routine used: numpy's
    argsort of its int32 degrees (row length of A + A^T, the diagonal
    counted twice), recomputed here the same way on the same numpy."""
    A = sp.csr_matrix((np.ones(edges.shape[0]), (edges[:, 0], edges[:, 1])), shape=(n, n))
    perm = np.argsort(reverse_cuthill_mckee(A))
    S = (A + A.T).tocsr()
    deg = np.diff(S.indptr).astype(np.int32)
    rows = np.repeat(np.arange(n), np.diff(S.indptr))
    deg[np.unique(rows[S.indices == rows])] += 1
    return perm, A.indptr.astype(np.uint32), A.indices.astype(np.uint32), np.argsort(deg), deg


def _rcm_graphs():
    m = meshgen.airfoil(40, 24, seed=3, shuffle=7)
    yield "shuffled airfoil nodes", m["pedge"], m["num_nodes"]
    g = meshgen.grid_edges(30, 20)
    yield "shuffled grid", g[meshgen.permutation(g.shape[0], 3)], 600
    # three components (a shuffled grid, a path, a star), two isolated
    # points, a self loop, an edge given in both directions and a duplicate
    h = meshgen.permutation(200, 4)[meshgen.grid_edges(20, 10)]
    path = np.array([[200 + i, 201 + i] for i in range(9)])
    star = np.array([[210, 211 + i] for i in range(6)])
    extra = np.array([[205, 205], [3, 7], [7, 3], [path[0, 0], path[0, 1]]])
    yield "disconnected", np.concatenate([h, path, star, extra]), 219


RCM_GRAPHS = list(_rcm_graphs())


@pytest.mark.parametrize("name,edges,n", RCM_GRAPHS, ids=[g[0] for g in RCM_GRAPHS])
def test_rcm_is_scipys(name, edges, n):
    """fot_rcm_order is scipy's reverse Cuthill-McKee on rcm.py's input
    convention: given the seed order scipy's routine used, the permutation
    is identical; with none it seeds by a stable degree sort."""
    want, ind, end, seeds, deg = _scipy_rcm(np.asarray(edges, dtype=np.int64), n)
    assert np.array_equal(plan.rcm_order(ind, end, seeds), want), name
    stable = plan.rcm_order(ind, end)
    assert np.array_equal(stable, plan.rcm_order(ind, end, np.argsort(deg, kind="stable")))
    assert np.array_equal(np.sort(stable), np.arange(n))


def test_rcm_and_gps_on_a_path():
    # a path graph has one RCM order up to reversal: bandwidth 1
    path = np.array([[i, i + 1] for i in range(9)])
    pi, pe = plan.graph_csr(path, 10)
    assert _csr_bandwidth(pi, pe, plan.rcm_order(pi, pe)) == 1
    assert _csr_bandwidth(pi, pe, plan.gps_order(pi, pe)) == 1
    # the stable seed is the lower-id end (degree 1 ties 0 and 9); reversal
    # numbers it last
    assert plan.rcm_order(pi, pe).tolist() == [9 - i for i in range(10)]


def test_owner_plan_invariants():
    for bs, pp in [(64, None), (None, 48)]:
        p = airfoil_problem(nx=40, ny=24, seed=14, block_size=bs or 64)
        if pp:
            p.partition_points(pp)
        op = plan.owner_plan(p.maps[0], p.num_points[0], bs or pp, None, p.point_partition)
        md = op["mesh_dim"]
        owned_all = []
        runs = np.zeros(p.num_cells, dtype=np.int64)
        for b in range(op["blk_owned"].size):
            c0, c1 = op["blk_cache_off"][b], op["blk_cache_off"][b + 1]
            nw = op["blk_owned"][b]
            cached = op["cached"][c0:c1]
            owned, halo = cached[:nw], cached[nw:]
            assert np.all(np.diff(owned.astype(np.int64)) > 0)
            assert np.all(np.diff(halo.astype(np.int64)) > 0)
            assert np.intersect1d(owned, halo).size == 0
            owned_all.append(owned)
            e0, e1 = op["blk_cell_off"][b], op["blk_cell_off"][b + 1]
            assert e1 - e0 <= 1024
            cells = op["cells"][e0:e1]
            loc = op["loc"][e0 * md:e1 * md].reshape(-1, md)
            assert np.array_equal(cached[loc], p.maps[0][cells])
            # every cell runs where it touches an owned point, only there
            assert np.all((loc < nw).any(axis=1))
            runs[cells] += 1
            col = op["colour"][e0:e1]
            assert np.all(np.diff(col.astype(int)) >= 0)
            for c in range(int(op["blk_ncol"][b])):
                pts = loc[col == c]
                pts = pts[pts < nw]
                assert np.unique(pts).size == pts.size
        owned_all = np.concatenate(owned_all)
        touched = np.unique(p.maps[0])
        assert np.array_equal(np.sort(owned_all), touched)  # owners partition the points
        assert runs.min() >= 1 and runs.max() <= md


def test_partition_points_numbering():
    # parts contiguous after renumbering; inside a part, "id" keeps the points'
    # previous relative order and "cm" follows the Cuthill-McKee order of the
    # point graph; the mesh is only renumbered (same cells, same point values)
    import pytest
    base = airfoil_problem(nx=30, ny=20, seed=3, block_size=64, res0_seed=5)
    ind, end = plan.graph_csr(base.maps[0], base.num_points[0])
    part = plan.kway_partition(ind, end, 48)[0]
    cm_rank = plan.gps_order(ind, end)
    for numbering, key in (("id", np.arange(part.size)), ("cm", cm_rank)):
        p = airfoil_problem(nx=30, ny=20, seed=3, block_size=64, res0_seed=5)
        p.partition_points(48, numbering)
        pp = p.point_partition
        assert np.all(np.diff(pp.astype(np.int64)) >= 0)  # contiguous parts
        # new id of each old point: the stable order of (part, key)
        order = np.lexsort((key, part))
        perm = np.empty_like(order)
        perm[order] = np.arange(order.size)
        assert np.array_equal(pp[perm], part)
        # the res values moved with their points
        assert np.array_equal(p.point_data[0][perm], base.point_data[0])
        # every cell kept: the same point pairs up to the renumbering
        cells_new = {tuple(r) for r in p.maps[0].tolist()}
        cells_old = {tuple(perm[r]) for r in base.maps[0].tolist()}
        assert cells_new == cells_old
    with pytest.raises(ValueError):
        airfoil_problem(nx=8, ny=6).partition_points(16, "rcm")


def test_generators():
    e = meshgen.grid_edges(3, 3)
    # Grid::fillEdgeList (grid.hpp:77-104) on a 3x3 grid
    assert e.tolist() == [[0, 3], [0, 1], [1, 4], [1, 2], [5, 2], [3, 6], [3, 4], [4, 7],
                          [4, 5], [8, 5], [6, 7], [7, 8]]
    assert meshgen.grid_edges(1024, 1024).shape[0] == 2095104
    h = meshgen.hex_grid(2, 2, 2)
    assert h.tolist() == [[0, 1, 3, 2, 4, 5, 7, 6]]
    m = meshgen.airfoil(30, 20, seed=1)
    assert m["num_edges"] == 30 * 19 + 29 * 20
    assert np.array_equal(np.sort(np.unique(m["pecell"])), np.arange(600))
    # orientation: (dy, -dx) of the edge points from cell 1 to cell 2
    x = m["x"]
    d = x[m["pedge"][:, 0]] - x[m["pedge"][:, 1]]
    normal = np.stack([d[:, 1], -d[:, 0]], axis=1)
    # cell centroids from the edges' midpoints (each interior cell has 2-4 edges)
    mid = 0.5 * (x[m["pedge"][:, 0]] + x[m["pedge"][:, 1]])
    acc = np.zeros((m["num_cells"], 2))
    cnt = np.zeros(m["num_cells"])
    for k in (0, 1):
        np.add.at(acc, m["pecell"][:, k], mid)
        np.add.at(cnt, m["pecell"][:, k], 1)
    cen = acc / cnt[:, None]
    step = cen[m["pecell"][:, 1]] - cen[m["pecell"][:, 0]]
    assert np.all(np.sum(step * normal, axis=1) > 0)
    assert np.all((m["adt"] >= 0.5) & (m["adt"] < 1.5))
    assert np.all(m["q"][:, 2] == 0) and abs(m["q"][:, 0].mean() - 1) < 1e-2


def _cache_lines_reference(points, dim, type_size, line, soa):
    # countCacheLinesForBlock (problem.hpp:654-672) with a line-size parameter
    per = line // type_size
    s = set()
    for p in points:
        if soa:
            s.add(p // per)
        else:
            for d in range(dim):
                s.add((p * dim + d) // per)
    return (dim if soa else 1) * len(s)


@pytest.mark.parametrize("soa", [False, True])
def test_count_cache_lines_matches_reference_rule(soa):
    import ctypes as C
    from cuda_fot_amd import capi
    rng = np.random.default_rng(7)
    for dim, line in [(4, 32), (4, 128), (5, 128), (1, 64)]:
        pts = rng.integers(0, 5000, size=300).astype(np.uint32)
        out = C.c_uint64()
        capi.call("fot_count_cache_lines", capi.ptr(pts), pts.size, dim, 8, line, int(soa),
                  C.byref(out))
        assert out.value == _cache_lines_reference(pts.tolist(), dim, 8, line, soa)


# ---- the product's planner against the hand-derived known answers ------
from kats import HPLAN_KATS, kat_global_colour_least_used  # noqa: E402


@pytest.mark.parametrize("kat", HPLAN_KATS, ids=[k.__name__ for k in HPLAN_KATS])
def test_hplan_known_answers(kat):
    kat(lambda maps, npts, bs, part: plan.hplan(maps, npts, bs, part))


def test_global_colouring_known_answer():
    kat_global_colour_least_used(plan.colour_cells)
