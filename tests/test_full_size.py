"""res_calc at the benchmark's own sizes: the shuffled 1200x600 airfoil
(1 438 200 edges, the headline) and 2400x1200 (5 756 400 edges), built as
bench.py builds them, run by the default plan (owner compute on 240-point
parts, the gather kernel, two edges per thread).  Two steps, checked bit for
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
    p.gpu_loop(RES_CALC, capi.STRATEGY_HIERARCHICAL, iters, block_size=ppb, flags=FLAGS)
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
