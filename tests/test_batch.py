"""Batched loops (fot_batch): many meshes of uneven size in one persistent
launch per step, eager and replayed from a hipGraph.  The bar: every mesh's
result is bit-identical to its own unbatched owner-compute loop with the
same plan (same summation order), and so to the oracle's owner replay."""
import copy

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, gpu_batch, meshgen, plan, volna_problem
from helpers import airfoil_problem, assert_bitwise, assert_close_scaled, oracle_of

pytestmark = pytest.mark.gpu

VOLNA = capi.functor(capi.KERNEL_VOLNA)
RES_CALC = capi.functor(capi.KERNEL_RES_CALC)
SIZES = [(8, 6, 64), (30, 20, 128), (61, 41, 256), (17, 9, 96), (45, 33, 192)]


def _volna_set(soa=False):
    return [volna_problem(meshgen.tri_mesh(nx, ny, seed=10 + i), block_size=bs, soa=soa)
            for i, (nx, ny, bs) in enumerate(SIZES)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_volna_batch_matches_unbatched(gpu, use_graph):
    iters = 3
    batch = _volna_set()
    singles = _volna_set()
    stats = gpu_batch(batch, VOLNA, iters, use_graph=use_graph)
    assert stats["num_problems"] == len(SIZES) and stats["uses_graph"] == int(use_graph)
    assert stats["grid"] <= stats["num_work_items"]
    for i, (b, s) in enumerate(zip(batch, singles)):
        ref = copy.deepcopy(s)
        s.gpu_loop(VOLNA, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_OWNER_COMPUTE)
        assert_bitwise(b.point_data[0], s.point_data[0], f"batch vs single, mesh {i}")
        o = oracle_of(ref, O.VOLNA)
        O.loop_owner_replay(o, plan.owner_plan(ref.maps[0], ref.num_points[0],
                                               ref.block_size), iters)
        assert_bitwise(b.point_data[0], o.target, f"batch vs owner replay, mesh {i}")


@pytest.mark.parametrize("use_graph", [False, True])
def test_mine_batch_reads_its_iterate(gpu, use_graph):
    # mine reads the point data it increments: every step copies the result
    # back before the next one, in the eager run and inside the graph
    from helpers import mine_problem
    iters = 3
    f = capi.functor(capi.KERNEL_MINE, 2, 4, 4)
    def make():
        return [mine_problem(N=24, M=20, seed=1, block_size=128),
                mine_problem(N=40, M=18, seed=2, block_size=96)]
    batch, singles = make(), make()
    gpu_batch(batch, f, iters, use_graph=use_graph)
    for i, (b, s) in enumerate(zip(batch, singles)):
        ref = copy.deepcopy(s)
        s.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_OWNER_COMPUTE)
        assert_bitwise(b.point_data[0], s.point_data[0], f"mine batch vs single, mesh {i}")
        o = oracle_of(ref, O.MINE2)
        O.loop_owner_replay(o, plan.owner_plan(ref.maps[0], ref.num_points[0],
                                               ref.block_size), iters)
        assert_bitwise(b.point_data[0], o.target, f"mine batch vs owner replay, mesh {i}")


def test_volna_batch_graph_replays_repeat(gpu):
    # two runs of the same graph-backed batch: the work queue resets itself
    from cuda_fot_amd import Batch, DeviceArray
    probs = _volna_set(soa=True)
    specs, dev0 = [], []
    for p in probs:
        dev = [DeviceArray.from_host(d) for d in p.point_data]
        out = DeviceArray(p.point_data[0].shape, np.float64)
        cdev = [DeviceArray.from_host(c) for c in p.cell_data]
        specs.append(dict(maps=p.maps, num_points=p.num_points, point_data=dev, out=out,
                          cell_data=cdev, soa=True, block_size=p.block_size))
        dev0.append((dev, out, cdev))
    b = Batch(VOLNA, specs, use_graph=True)
    b.run(2)
    first = [d[0][0].download() for d in dev0]
    b.run(2)
    second = [d[0][0].download() for d in dev0]
    b.close()
    ref = _volna_set(soa=True)
    for i, (p, f, s) in enumerate(zip(ref, first, second)):
        o = oracle_of(p, O.VOLNA)
        S = o.loop_serial(4, with_scale=True)
        assert_close_scaled(s, o.target, S, what=f"4 iterations over two runs, mesh {i}")
        assert not np.array_equal(f, s)


def test_res_calc_batch_matches_unbatched(gpu):
    iters = 2
    def make():
        ps = [airfoil_problem(nx=48, ny=24, res0_seed=3, block_size=192),
              airfoil_problem(nx=90, ny=40, seed=2, res0_seed=4, block_size=128)]
        for p in ps:
            p.partition_points(p.block_size)
        return ps
    batch, singles = make(), make()
    gpu_batch(batch, RES_CALC, iters)
    for i, (b, s) in enumerate(zip(batch, singles)):
        s.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, iters,
                   flags=capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD)
        assert_bitwise(b.point_data[0], s.point_data[0], f"res_calc batch vs single, mesh {i}")


def test_batch_with_an_empty_mesh(gpu):
    # an empty mesh in a batch contributes no work items; the others are
    # unaffected, and a batch of empty meshes runs (and copies) nothing
    iters = 2
    empty = airfoil_problem(nx=6, ny=4, res0_seed=21)
    empty.maps = [m[:0].copy() for m in empty.maps]
    full = airfoil_problem(nx=30, ny=20, res0_seed=22, block_size=128)
    single = copy.deepcopy(full)
    before = empty.point_data[0].copy()
    gpu_batch([empty, full], RES_CALC, iters)
    single.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, iters,
                    flags=capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD)
    assert_bitwise(empty.point_data[0], before, "empty mesh in a batch")
    assert_bitwise(full.point_data[0], single.point_data[0], "batch with an empty mesh")
    only = airfoil_problem(nx=6, ny=4, res0_seed=23)
    only.maps = [m[:0].copy() for m in only.maps]
    before = only.point_data[0].copy()
    for graph in (False, True):
        stats = gpu_batch([only], RES_CALC, iters, use_graph=graph)
        assert stats["num_work_items"] == 0 and stats["uses_graph"] == 0
        assert_bitwise(only.point_data[0], before, f"batch of one empty mesh, graph {graph}")


def _with_untouched_points(p, extra, part_size, seed):
    """`p` with `extra` points that no cell touches appended to every point
    set read through mapping 0's ids, and an owner partition: parts of
    `part_size` consecutive touched points; the first third of the new
    points join part 0 (a block that runs cells and owns untouched points),
    the rest form two parts of their own (blocks that own points and run no
    cell)."""
    n0 = p.num_points[0]
    same = [k for k in range(len(p.maps)) if np.array_equal(p.maps[k], p.maps[0])]
    for k in same:
        d = p.point_data[k]
        if d is None:
            continue
        dim = 1 if d.ndim == 1 else (d.shape[0] if p.soa else d.shape[1])
        new = meshgen.uniform(extra * dim, seed + k, 0.5, 1.5)
        if d.ndim == 1:
            p.point_data[k] = np.concatenate([d, new])
        elif p.soa:
            p.point_data[k] = np.concatenate([d, new.reshape(dim, extra)], axis=1)
        else:
            p.point_data[k] = np.concatenate([d, new.reshape(extra, dim)], axis=0)
        p.num_points[k] = n0 + extra
    touched = (np.arange(n0) // part_size).astype(np.uint32)
    nparts = int(touched.max()) + 1
    third = extra // 3
    tail = np.where(np.arange(extra) < third, 0,
                    np.where(np.arange(extra) < 2 * third, nparts, nparts + 1)).astype(np.uint32)
    p.point_partition = np.concatenate([touched, tail])
    p.block_size = part_size
    return p


@pytest.mark.parametrize("soa", [False, True])
@pytest.mark.parametrize("kind", ["mine", "volna"])
def test_owner_blocks_that_run_no_cell(gpu, kind, soa):
    """Owner-compute blocks that own points no cell touches run no cell:
    the owner loop and the batch leave those points as they were (mine's
    Jacobi copy included) and match the oracle's owner replay bit for bit."""
    from helpers import mine_problem
    iters = 3
    if kind == "mine":
        f, k = capi.functor(capi.KERNEL_MINE, 2, 4, 4), O.MINE2

        def make(i):
            return _with_untouched_points(mine_problem(N=20, M=14, seed=3 + i, soa=soa),
                                          30, 64, 40 + i)
    else:
        f, k = VOLNA, O.VOLNA

        def make(i):
            return _with_untouched_points(
                volna_problem(meshgen.tri_mesh(14, 10, seed=5 + i), soa=soa), 30, 64, 50 + i)
    for i in range(2):
        p = make(i)
        n0 = p.num_points[0] - 30
        before = p.target_in_original_order().copy()
        o = oracle_of(make(i), k)
        O.loop_owner_replay(o, plan.owner_plan(p.maps[0], p.num_points[0], p.block_size, None,
                                               p.point_partition), iters)
        p.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_OWNER_COMPUTE)
        assert_bitwise(p.point_data[0], o.target, f"{kind} owner loop, mesh {i}")
        got = p.point_data[0][:, n0:] if soa else p.point_data[0][n0:]
        old = before[:, n0:] if soa else before[n0:]
        assert_bitwise(got, old, f"{kind}: untouched points changed")
    batch = [make(0), make(1)]
    gpu_batch(batch, f, iters)
    for i, b in enumerate(batch):
        o = oracle_of(make(i), k)
        O.loop_owner_replay(o, plan.owner_plan(b.maps[0], b.num_points[0], b.block_size, None,
                                               b.point_partition), iters)
        assert_bitwise(b.point_data[0], o.target, f"{kind} batch, mesh {i}")
