"""res_calc at the benchmark's own sizes: the shuffled 1200x600 airfoil
(1 438 200 edges, the headline) and 2400x1200 (5 756 400 edges), built as
bench.py builds them, run by the default plan (owner compute on 240-point
parts, the gather kernel, two edges per thread; at 2400x1200 its res loads
and stores nontemporal).  Two steps, checked bit for
bit against the oracle's replay of the same plan and within 1e-12 * S of the
serial oracle loop, plus the conservation that holds at any size: every edge
adds +f to one cell and -f to the other, so the increments sum to zero up to
rounding."""
import copy

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, meshgen, plan, res_calc_problem
from helpers import assert_bitwise, assert_close_scaled, oracle_of

pytestmark = pytest.mark.gpu

RES_CALC = capi.functor(capi.KERNEL_RES_CALC)
FLAGS = capi.LOOP_OWNER_COMPUTE | capi.LOOP_TWO_PER_THREAD | capi.LOOP_OWNER_GATHER


@pytest.mark.parametrize("nx,ny", [(1200, 600), (2400, 1200)])
def test_res_calc_at_benchmark_size(gpu, nx, ny):
    iters, ppb = 2, 240
    print(f"res_calc {nx}x{ny}, owner gather, {iters} steps", flush=True)
    mesh = meshgen.airfoil(nx, ny, seed=0, shuffle=7)
    p = res_calc_problem(mesh)
    p.partition_points(ppb)
    ref = copy.deepcopy(p)
    stats = p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, iters, block_size=ppb, flags=FLAGS)
    # nontemporal res by default only where the step's data exceed the
    # 256 MB Infinity Cache (2400x1200: about 420 MB; 1200x600: about 105 MB)
    assert stats["nontemporal"] == (1 if nx * ny > 2_000_000 else 0)
    got = p.point_data[0]
    o = oracle_of(ref, O.RES_CALC)
    O.loop_owner_replay(o, plan.owner_plan(ref.maps[0], ref.num_points[0], ppb, None,
                                           ref.point_partition), iters)
    assert_bitwise(got, o.target, f"{nx}x{ny} gather kernel vs owner replay")
    s = oracle_of(ref, O.RES_CALC)
    S = s.loop_serial(iters, with_scale=True)
    assert_close_scaled(got, s.target, S, what=f"{nx}x{ny} gather kernel vs serial")
    # res started at zero: the total of all increments cancels edge by edge
    total = got.sum(axis=0)
    assert np.all(np.abs(total) <= 1e-12 * S.sum(axis=0)), total


def _mine_1024():
    from cuda_fot_amd import Problem
    n = 1024
    e = meshgen.grid_edges(n, n)
    pd = meshgen.uniform(n * n * 4, 0).reshape(n * n, 4)
    w = meshgen.uniform(e.shape[0] * 4, 1).reshape(4, -1)
    return Problem([e], [n * n], [pd], [w], 512), capi.functor(capi.KERNEL_MINE, 2, 4, 4), O.MINE2


def _hex_128():
    from cuda_fot_amd import Problem
    n = 128
    cells = meshgen.hex_grid(n + 1, n + 1, n + 1)
    cells = cells[meshgen.permutation(cells.shape[0], 5)]
    nn = (n + 1) ** 3
    forces = meshgen.normal(cells.shape[0] * 24, 6).reshape(24, -1)
    return (Problem([cells], [nn], [np.zeros((nn, 3))], [forces], 512),
            capi.functor(capi.KERNEL_HEX_SCATTER), O.HEX_SCATTER)


def _mini_aero_96():
    from cuda_fot_amd import mini_aero_problem
    return (mini_aero_problem(meshgen.hex_faces(96, 96, 96, seed=0), block_size=512),
            capi.functor(capi.KERNEL_MINI_AERO), O.MINI_AERO)


def _lulesh_128():
    from cuda_fot_amd import lulesh_problem
    return (lulesh_problem(129, 129, 129, seed=0, block_size=512),
            capi.functor(capi.KERNEL_LULESH), O.LULESH)


@pytest.mark.parametrize("make", [_mine_1024, _hex_128, _mini_aero_96, _lulesh_128],
                         ids=["mine-1024^2", "hex-128^3", "mini_aero-96^3", "lulesh-128^3"])
def test_other_workloads_at_benchmark_size(gpu, make):
    # one step of owner compute (240-point parts) and of the hierarchical
    # loop on the GPS-ordered mesh (B = 256) per block colour and two-phase:
    # each bit for bit against the oracle's replay of its plan and, mapped
    # back to the original numbering, within 1e-12 * S of the serial oracle
    # loop
    p, f, kernel = make()
    print(f"{make.__name__}: {p.num_cells} cells", flush=True)
    s = oracle_of(p, kernel)
    S = s.loop_serial(1, with_scale=True)
    q = copy.deepcopy(p)
    q.partition_points(240)
    q0 = copy.deepcopy(q)
    q.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, 1, block_size=240, flags=capi.LOOP_OWNER_COMPUTE)
    o = oracle_of(q0, kernel)
    O.loop_owner_replay(o, plan.owner_plan(q0.maps[0], q0.num_points[0], 240, None,
                                           q0.point_partition), 1)
    assert_bitwise(q.point_data[0], o.target, f"{make.__name__} owner compute vs replay")
    assert_close_scaled(q.target_in_original_order(), s.target, S,
                        what=f"{make.__name__} owner compute vs serial")
    g = copy.deepcopy(p)
    g.reorder("gps")
    g0 = copy.deepcopy(g)
    t = copy.deepcopy(g)
    g.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, 1, block_size=256, partition=False)
    h = oracle_of(g0, kernel)
    h.loop_hier_replay(O.OraclePlan(g0.maps, g0.num_points[0], 256, None), 1)
    assert_bitwise(g.point_data[0], h.target, f"{make.__name__} GPS hierarchical vs replay")
    assert_close_scaled(g.target_in_original_order(), s.target, S,
                        what=f"{make.__name__} GPS hierarchical vs serial")
    # the two-phase write-back of the same plan: the same bits; at these
    # sizes the hex forces, mini_aero's face values and lulesh's stresses
    # (beyond the Infinity Cache) are loaded with nontemporal hints by default
    st = t.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, 1, block_size=256, partition=False,
                    flags=capi.LOOP_TWO_PHASE)
    assert_bitwise(t.point_data[0], h.target, f"{make.__name__} GPS two-phase vs replay")
    assert st["nontemporal"] == (1 if make in (_hex_128, _mini_aero_96, _lulesh_128) else 0)


def test_volna_batch_at_benchmark_size(gpu):
    # the 64 uneven triangle meshes of the batched benchmark (cells
    # log-uniform in [2e4, 4e5]): one persistent launch per step over all of
    # them, each mesh bit for bit equal to its own owner-compute loop; the
    # smallest and the largest also against the oracle's replay of the plan
    from cuda_fot_amd import gpu_batch, volna_problem
    rng = np.random.default_rng(0)
    shapes = []
    for i in range(64):
        cells = float(np.exp(rng.uniform(np.log(2e4), np.log(4e5))))
        aspect = float(rng.uniform(0.5, 2.0))
        nx = max(4, int(round(np.sqrt(cells / 2 * aspect))))
        ny = max(4, int(round(cells / 2 / nx)))
        shapes.append((nx, ny, 100 + i))

    def make():
        return [volna_problem(meshgen.tri_mesh(nx, ny, seed=s), block_size=240)
                for nx, ny, s in shapes]
    iters = 2
    f = capi.functor(capi.KERNEL_VOLNA)
    batch = make()
    print(f"volna batch: {sum(p.num_cells for p in batch)} edges", flush=True)
    refs = [copy.deepcopy(p) for p in batch]
    gpu_batch(batch, f, iters)
    sizes = [p.num_cells for p in batch]
    check = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for i, (b, r) in enumerate(zip(batch, refs)):
        s = copy.deepcopy(r)
        s.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_OWNER_COMPUTE)
        assert_bitwise(b.point_data[0], s.point_data[0], f"batched vs single, mesh {i}")
        if i in check:
            o = oracle_of(r, O.VOLNA)
            O.loop_owner_replay(o, plan.owner_plan(r.maps[0], r.num_points[0], r.block_size),
                                iters)
            assert_bitwise(b.point_data[0], o.target, f"batched vs owner replay, mesh {i}")
