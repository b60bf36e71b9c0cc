"""GPU parity of the loop engine against the CPU oracle (calls libfot's C-ABI).

Bars: the global-colour and hierarchical GPU loops keep a summation order
fixed by their plans, and the user functions are compiled without FP
contraction, so they are bit-identical to the oracle's replay of the same
order (oracle_loop_colour / oracle_loop_hier_replay).  The atomics loop adds
in arbitrary order and is held to |gpu - serial| <= 1e-12 * max(|serial|, S)
with S the oracle's sum of absolute increments.
"""
import copy

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, plan
from helpers import (airfoil_problem, assert_bitwise, assert_close_scaled, hex_problem,
                     mine_problem, oracle_of)

pytestmark = pytest.mark.gpu

ITERS = 3
RES_CALC = capi.functor(capi.KERNEL_RES_CALC)


def _serial(p, kernel, iters=ITERS):
    o = oracle_of(p, kernel)
    S = o.loop_serial(iters, with_scale=True)
    return o.target.copy(), S


def _colour(p, kernel, iters=ITERS):
    o = oracle_of(p, kernel)
    o.loop_colour(iters, 1)
    return o.target.copy()


def _replay(p, kernel, block_size, partition=None, iters=ITERS):
    o = oracle_of(p, kernel)
    pl = O.OraclePlan(p.maps, p.num_points[0], block_size, partition)
    o.loop_hier_replay(pl, iters)
    return o.target.copy()


@pytest.mark.parametrize("soa", [False, True])
def test_res_calc_atomic_vs_serial(gpu, soa):
    p = airfoil_problem(soa=soa, res0_seed=3)
    ref, S = _serial(p, O.RES_CALC)
    p.gpu_loop(RES_CALC, capi.STRATEGY_ATOMIC, ITERS, block_size=256)
    assert_close_scaled(p.point_data[0], ref, S, what="atomic")


@pytest.mark.parametrize("soa", [False, True])
def test_res_calc_global_colour_bitwise(gpu, soa):
    p = airfoil_problem(soa=soa, res0_seed=4)
    ref = _colour(p, O.RES_CALC)
    p.gpu_loop(RES_CALC, capi.STRATEGY_GLOBAL_COLOUR, ITERS, block_size=128)
    assert_bitwise(p.point_data[0], ref, "global colour")


DEP = capi.LOOP_DEPENDENCY_ORDER
STAGE = capi.LOOP_STAGE_INPUTS
OWN = capi.LOOP_OWNER_COMPUTE
TWO = capi.LOOP_TWO_PER_THREAD
GATHER = capi.LOOP_GATHER_NODES
TERMS = capi.LOOP_POINT_TERMS
OG = capi.LOOP_OWNER_GATHER
HBM = capi.LOOP_LISTS_IN_HBM
PERS = capi.LOOP_PERSISTENT
TP = capi.LOOP_TWO_PHASE
NT = capi.LOOP_NONTEMPORAL


def _owner_replay(p, kernel, block_size, partition=None, iters=ITERS, point_partition=None):
    o = oracle_of(p, kernel)
    O.loop_owner_replay(o, plan.owner_plan(p.maps[0], p.num_points[0], block_size, partition,
                                           point_partition), iters)
    return o.target.copy()


@pytest.mark.parametrize("soa,bs,flags,method", [(False, 128, OWN, None),
                                                 (True, 256, OWN | STAGE, None),
                                                 (False, 512, OWN | STAGE, "kway"),
                                                 (False, 256, OWN, "ours"),
                                                 (True, 1024, OWN | STAGE, "kway"),
                                                 (False, 64, OWN, "points"),
                                                 (True, 96, OWN, "points"),
                                                 (False, 128, OWN | TWO, "points"),
                                                 (True, 200, OWN | TWO, "points"),
                                                 (False, 192, OWN | TWO | GATHER, "points"),
                                                 (False, 192, OWN | TWO | GATHER | TERMS,
                                                  "points"),
                                                 (True, 160, OWN | TERMS, "points"),
                                                 (False, 192, OWN | TWO | OG, "points"),
                                                 (True, 224, OWN | TWO | OG, "points"),
                                                 (False, 128, OWN | OG, "points"),
                                                 (False, 256, OWN | TWO | OG, None),
                                                 (True, 512, OWN | OG, "kway"),
                                                 (False, 192, OWN | TWO | OG | PERS, "points"),
                                                 (True, 224, OWN | TWO | OG | PERS, "points"),
                                                 (False, 128, OWN | OG | PERS, "points"),
                                                 (True, 512, OWN | OG | PERS, "kway"),
                                                 (False, 240, OWN | TWO | OG | TERMS, "points"),
                                                 (True, 192, OWN | TWO | OG | TERMS, "points"),
                                                 (False, 192, OWN | TWO | OG | TERMS | PERS,
                                                  "points"),
                                                 (True, 128, OWN | GATHER, "points"),
                                                 (False, 256, OWN | TWO, None),
                                                 # the owned cells' edge lists loaded from
                                                 # the plan instead of built in LDS
                                                 (False, 192, OWN | TWO | OG | HBM, "points"),
                                                 (True, 224, OWN | TWO | OG | HBM, "points"),
                                                 (False, 128, OWN | OG | PERS | HBM, "points"),
                                                 (True, 192, OWN | TWO | OG | TERMS | HBM,
                                                  "points"),
                                                 (True, 512, OWN | OG | HBM, "kway"),
                                                 # res loaded and stored nontemporally
                                                 (False, 240, OWN | TWO | OG | NT, "points"),
                                                 (True, 192, OWN | TWO | OG | NT, "points"),
                                                 (False, 128, OWN | OG | NT, "points"),
                                                 (True, 512, OWN | OG | NT, "kway"),
                                                 (False, 256, OWN | TWO | OG | NT, None)])
def test_res_calc_owner_compute_bitwise(gpu, soa, bs, flags, method):
    p = airfoil_problem(nx=64, ny=32, soa=soa, res0_seed=10, block_size=bs)
    orig = airfoil_problem(nx=64, ny=32, soa=soa, res0_seed=10, block_size=bs)
    if method == "kway":
        p.reorder("gps")
        p.partition(1.001)
    elif method == "ours":
        p.partition_ours(1.001, 4)
    if method in ("kway", "ours"):
        p.reorder_to_partition()
        p.renumber_points()
    if method == "points":
        p.partition_points(bs)
    ref = _owner_replay(p, O.RES_CALC, bs, p.partition_vector, point_partition=p.point_partition)
    serial, S = _serial(orig, O.RES_CALC)
    stats = p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
    assert_bitwise(p.point_data[0], ref, "owner compute")
    assert_close_scaled(p.target_in_original_order(), serial, S, what="owner compute vs serial")
    # the nontemporal form exists for the gather kernel with lists in LDS,
    # one block per workgroup and no per-cell terms
    nt = bool(flags & NT) and bool(flags & OG) and not flags & (PERS | HBM | TERMS)
    assert stats["nontemporal"] == int(nt)


def test_res_calc_gather_edge_on_one_cell_twice(gpu):
    # an edge whose two cells are the same cell gives that cell two edge-list
    # entries of one thread colour: the lists cannot be built one slot per
    # colour, so the plan loads them (the same bits as the oracle's replay,
    # which adds the edge's two sides in slot order)
    p = airfoil_problem(nx=24, ny=16, res0_seed=12, block_size=128)
    for m in (0, 2, 3):
        p.maps[m][5, 1] = p.maps[m][5, 0]
    p.partition_points(128)
    ref = _owner_replay(p, O.RES_CALC, 128, None, point_partition=p.point_partition)
    for flags in (OWN | TWO | OG, OWN | TWO | OG | HBM):
        q = copy.deepcopy(p)
        q.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
        assert_bitwise(q.point_data[0], ref, f"edge on one cell twice, flags {flags}")


def test_res_calc_owner_engine_partition(gpu):
    # owner compute without a partition: the engine partitions the points
    p = airfoil_problem(nx=64, ny=32, res0_seed=11, block_size=128)
    serial, S = _serial(p, O.RES_CALC)
    p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=OWN, partition=False)
    assert_close_scaled(p.point_data[0], serial, S, what="owner compute, engine partition")


@pytest.mark.parametrize("soa,bs,flags", [(False, 128, 0), (True, 128, 0),
                                          (False, 256, STAGE), (True, 64, STAGE),
                                          (False, 1024, STAGE), (False, 128, DEP),
                                          (True, 256, DEP | STAGE), (False, 512, DEP | STAGE),
                                          (False, 128, TP), (True, 64, TP),
                                          (True, 256, TP | STAGE), (False, 1024, TP | STAGE)])
def test_res_calc_hierarchical_bitwise(gpu, soa, bs, flags):
    p = airfoil_problem(soa=soa, res0_seed=5, block_size=bs)
    ref = _replay(p, O.RES_CALC, bs)
    serial, S = _serial(p, O.RES_CALC)
    p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
    assert_bitwise(p.point_data[0], ref, "hierarchical")
    assert_close_scaled(p.point_data[0], serial, S, what="hierarchical vs serial")


@pytest.mark.parametrize("method,flags", [("kway", STAGE), ("ours", STAGE),
                                          ("kway", STAGE | DEP), ("ours", DEP),
                                          ("kway", STAGE | TP), ("ours", TP)])
def test_res_calc_partitioned_renumbered(gpu, method, flags):
    p = airfoil_problem(nx=64, ny=32, res0_seed=6, block_size=256)
    orig = airfoil_problem(nx=64, ny=32, res0_seed=6, block_size=256)
    if method == "kway":
        p.reorder("gps")
        p.partition(1.001)
    else:
        p.partition_ours(1.001, 4)
    p.reorder_to_partition()
    p.renumber_points()
    ref = _replay(p, O.RES_CALC, 256, p.partition_vector)
    p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
    assert_bitwise(p.point_data[0], ref, "partitioned hierarchical")
    # reordering invariance (res_calc.cu:129-181), through applied_permutation
    serial, S = _serial(orig, O.RES_CALC)
    assert_close_scaled(p.target_in_original_order(), serial, S, what="reordered vs original")


@pytest.mark.parametrize("flags", [STAGE, STAGE | DEP, STAGE | TP, TP])
def test_res_calc_graph_replay_matches_eager(gpu, flags):
    # several replays of one captured iteration, host work in between
    p1 = airfoil_problem(res0_seed=7)
    p2 = airfoil_problem(res0_seed=7)
    p1.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, 6, flags=flags)
    p2.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, 6, flags=flags, use_graph=True)
    assert_bitwise(p2.point_data[0], p1.point_data[0], "graph vs eager")
    ref = _replay(airfoil_problem(res0_seed=7), O.RES_CALC, 128, None, iters=6)
    assert_bitwise(p2.point_data[0], ref, "graph vs replay")


def test_res_calc_atomic_run_to_run(gpu):
    a = airfoil_problem(res0_seed=8)
    b = airfoil_problem(res0_seed=8)
    ref, S = _serial(airfoil_problem(res0_seed=8), O.RES_CALC)
    a.gpu_loop(RES_CALC, capi.STRATEGY_ATOMIC, ITERS)
    b.gpu_loop(RES_CALC, capi.STRATEGY_ATOMIC, ITERS)
    assert_close_scaled(a.point_data[0], b.point_data[0], S, what="atomic run to run")


@pytest.mark.parametrize("P,C,soa", [(4, 4, False), (4, 1, True), (1, 1, False), (8, 8, True)])
def test_mine_strategies(gpu, P, C, soa):
    f = capi.functor(capi.KERNEL_MINE, 2, P, C)
    base = mine_problem(P=P, C=C, soa=soa)
    ref_serial, S = _serial(base, O.MINE2)
    ref_colour = _colour(mine_problem(P=P, C=C, soa=soa), O.MINE2)
    ref_hier = _replay(mine_problem(P=P, C=C, soa=soa), O.MINE2, 128)
    g = mine_problem(P=P, C=C, soa=soa)
    g.gpu_loop(f, capi.STRATEGY_GLOBAL_COLOUR, ITERS)
    assert_bitwise(g.point_data[0], ref_colour, "mine global colour")
    for flags in (0, DEP, TP, STAGE, STAGE | TP, STAGE | DEP):
        h = mine_problem(P=P, C=C, soa=soa)
        h.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
        assert_bitwise(h.point_data[0], ref_hier, f"mine hierarchical flags {flags}")
    ref_own = _owner_replay(mine_problem(P=P, C=C, soa=soa), O.MINE2, 128)
    w = mine_problem(P=P, C=C, soa=soa)
    w.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, ITERS, flags=OWN)
    assert_bitwise(w.point_data[0], ref_own, "mine owner compute")
    a = mine_problem(P=P, C=C, soa=soa)
    a.gpu_loop(f, capi.STRATEGY_ATOMIC, ITERS)
    assert_close_scaled(a.point_data[0], ref_serial, S, what="mine atomic")


@pytest.mark.parametrize("soa", [False, True])
def test_hex_scatter_strategies(gpu, soa):
    f = capi.functor(capi.KERNEL_HEX_SCATTER)
    ref_serial, S = _serial(hex_problem(soa=soa), O.HEX_SCATTER)
    ref_hier = _replay(hex_problem(soa=soa), O.HEX_SCATTER, 128)
    ref_colour = _colour(hex_problem(soa=soa), O.HEX_SCATTER)
    ref_own = _owner_replay(hex_problem(soa=soa), O.HEX_SCATTER, 128)
    for strategy, ref, exact, flags in [(capi.STRATEGY_HIERARCHICAL, ref_hier, True, 0),
                                        (capi.STRATEGY_HIERARCHICAL, ref_own, True, OWN),
                                        (capi.STRATEGY_HIERARCHICAL, ref_hier, True, DEP),
                                        (capi.STRATEGY_HIERARCHICAL, ref_hier, True, TP),
                                        (capi.STRATEGY_HIERARCHICAL, ref_hier, True, TP | NT),
                                        (capi.STRATEGY_GLOBAL_COLOUR, ref_colour, True, 0),
                                        (capi.STRATEGY_ATOMIC, ref_serial, False, 0)]:
        p = hex_problem(soa=soa)
        st = p.gpu_loop(f, strategy, ITERS, flags=flags)
        # the nontemporal force loads exist for the two-phase launch
        assert st["nontemporal"] == int(flags == TP | NT)
        if exact:
            assert_bitwise(p.point_data[0], ref, f"hex strategy {strategy}")
        else:
            assert_close_scaled(p.point_data[0], ref, S, what="hex atomic")


def test_empty_and_single_cell(gpu):
    # a mesh of one edge: every strategy adds exactly one increment pair
    m = airfoil_problem(nx=2, ny=2, res0_seed=9)
    ref, S = _serial(m, O.RES_CALC)
    for strategy in (capi.STRATEGY_ATOMIC, capi.STRATEGY_GLOBAL_COLOUR,
                     capi.STRATEGY_HIERARCHICAL):
        p = airfoil_problem(nx=2, ny=2, res0_seed=9)
        p.gpu_loop(RES_CALC, strategy, ITERS)
        assert_close_scaled(p.point_data[0], ref, S, what=f"tiny mesh strategy {strategy}")


@pytest.mark.parametrize("copies", [2, 3])
def test_res_calc_owner_gather_wide_lists(gpu, copies):
    # repeated edges raise cell degrees to 8 (8-entry lists) or 12 (more than
    # the gather kernel's lists hold: the engine falls back to the colour loop)
    from cuda_fot_amd import meshgen, res_calc_problem
    m = meshgen.airfoil(40, 24, seed=6)
    rep = {k: np.concatenate([m[k]] * copies) if k in ("pedge", "pecell") else v
           for k, v in m.items()}
    def make():
        p = res_calc_problem(rep, block_size=128,
                             res0=meshgen.uniform(m["num_cells"] * 4, 13, -1, 1).reshape(-1, 4))
        p.partition_points(128)
        return p
    p = make()
    ref = _owner_replay(make(), O.RES_CALC, 128, point_partition=p.point_partition)
    serial, S = _serial(make(), O.RES_CALC)
    p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=OWN | TWO | OG)
    assert_bitwise(p.point_data[0], ref, f"owner gather, {copies} copies of each edge")
    # the serial reference ran on the same (renumbered) problem
    assert_close_scaled(p.point_data[0], serial, S, what="owner gather vs serial")


def test_loop_keeps_its_device_buffers_alive(gpu):
    """A Loop launches kernels on raw device addresses: it must hold the
    DeviceArrays it was built on, so a caller that drops its own references
    (as `loop, _, _ = ...` does) cannot free memory the kernels still read."""
    import gc
    import weakref
    from cuda_fot_amd import DeviceArray, Loop
    p = airfoil_problem(res0_seed=5)
    ref, S = _serial(p, O.RES_CALC, iters=1)

    def build():
        dev = [DeviceArray.from_host(d) for d in p.point_data]
        out = DeviceArray.from_host(p.point_data[0])
        loop = Loop(RES_CALC, p.maps, p.num_points, dev, out,
                    strategy=capi.STRATEGY_ATOMIC, block_size=256)
        return loop, [weakref.ref(d) for d in dev], weakref.ref(out)

    loop, dev_refs, out_ref = build()
    gc.collect()
    assert all(r() is not None for r in dev_refs) and out_ref() is not None
    loop.run(1)   # out = res; one step; res = out (synchronised)
    got = dev_refs[0]().download()
    assert_close_scaled(got, ref, S, what="atomic over loop-held buffers")
    loop.close()
    del loop
    gc.collect()
    assert all(r() is None for r in dev_refs) and out_ref() is None


@pytest.mark.parametrize("soa,flags", [(False, OWN | TWO | OG), (True, OWN | TWO | OG),
                                       (False, OWN | TWO | OG | PERS), (False, OWN | TWO)])
def test_res_calc_owner_blocks_owning_more_cells_than_threads(gpu, soa, flags):
    # every third edge of the airfoil and blocks of 256 consecutive cells: a
    # block runs ~170 edges (96 threads at two edges each, rounded to 192)
    # but owns 256 cells, so owned cells span several staging rounds
    from cuda_fot_amd import meshgen, res_calc_problem
    m = meshgen.airfoil(80, 48, seed=7)
    keep = np.arange(m["pedge"].shape[0]) % 3 == 0
    sub = dict(m, pedge=m["pedge"][keep], pecell=m["pecell"][keep], num_edges=int(keep.sum()))

    def make():
        p = res_calc_problem(sub, block_size=128, soa=soa,
                             res0=meshgen.uniform(m["num_cells"] * 4, 17, -1, 1).reshape(-1, 4))
        p.point_partition = (np.arange(p.num_points[0]) // 256).astype(np.uint32)
        return p
    p = make()
    ref = _owner_replay(make(), O.RES_CALC, 128, point_partition=p.point_partition)
    serial, S = _serial(make(), O.RES_CALC)
    stats = p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
    assert stats["stage_per_thread"] > 1
    assert stats["threads_per_block"] < 256, stats
    assert_bitwise(p.point_data[0], ref, f"owner, owned cells over several rounds ({flags})")
    assert_close_scaled(p.point_data[0], serial, S, what="owner vs serial")


@pytest.mark.parametrize("soa", [False, True])
def test_res_calc_persistent_more_blocks_than_resident_workgroups(gpu, soa):
    # ~2240 blocks of 240 points: more than the resident workgroups (at most
    # 8 of 256 threads per CU, 2048 on the device), so persistent workgroups
    # run several blocks, a grid apart
    p = airfoil_problem(nx=960, ny=560, soa=soa, res0_seed=12, block_size=240)
    orig = airfoil_problem(nx=960, ny=560, soa=soa, res0_seed=12, block_size=240)
    p.partition_points(240)
    ref = _owner_replay(p, O.RES_CALC, 240, point_partition=p.point_partition)
    serial, S = _serial(orig, O.RES_CALC)
    stats = p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, ITERS, flags=OWN | TWO | OG | PERS)
    assert stats["launch_grid"] % 8 == 0 and stats["launch_grid"] < stats["num_blocks"], stats
    assert_bitwise(p.point_data[0], ref, "persistent owner gather")
    assert_close_scaled(p.target_in_original_order(), serial, S, what="persistent vs serial")


def test_empty_mesh_is_a_no_op(gpu):
    # no cells: every strategy plans nothing, launches nothing and leaves the
    # point data as it was
    base = airfoil_problem(nx=8, ny=6, res0_seed=14)
    for strategy, flags in [(capi.STRATEGY_ATOMIC, 0), (capi.STRATEGY_GLOBAL_COLOUR, 0),
                            (capi.STRATEGY_HIERARCHICAL, 0), (capi.STRATEGY_HIERARCHICAL, DEP),
                            (capi.STRATEGY_HIERARCHICAL, TP),
                            (capi.STRATEGY_HIERARCHICAL, OWN | TWO | OG)]:
        for graph in (False, True):
            p = airfoil_problem(nx=8, ny=6, res0_seed=14)
            p.maps = [m[:0].copy() for m in p.maps]
            p.gpu_loop(RES_CALC, strategy, ITERS, flags=flags, use_graph=graph, partition=False)
            assert_bitwise(p.point_data[0], base.point_data[0],
                           f"empty mesh, strategy {strategy}/{flags}, graph {graph}")
