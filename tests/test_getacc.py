"""getacc_scatter (kernels/getacc_scatter.hpp, getacc_scatter_func.hpp): the
oracle pinned by a known answer for each mass fallback and an independent
numpy restatement (CPU), and the GPU strategies against it (gpu)."""
import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import Problem, capi, getacc_problem, plan
from helpers import assert_bitwise, assert_close_scaled, oracle_of

GETACC = capi.functor(capi.KERNEL_GETACC)


def test_getacc_mass_fallbacks_known_answer():
    # one quad: corner 0 has mass; corner 1 none (takes corner 0's);
    # corners 2 and 3 none with corner 1 and 2 none (take rho * cnwt)
    cells = np.array([[0, 1, 2, 3]], np.uint32)
    cnmass = np.array([[2.0], [0.0], [0.0], [0.0]])
    rho = np.array([[3.0]])
    cnwt = np.array([[0.5], [0.25], [0.125], [1.0]])
    cnfx = np.array([[1.0], [2.0], [3.0], [4.0]])
    cnfy = -cnfx
    p = Problem([cells], [4], [np.zeros((4, 4))], [cnmass, rho, cnwt, cnfx, cnfy])
    o = oracle_of(p, O.GETACC)
    o.loop_serial(1)
    assert o.target[:, 0].tolist() == [2.0, 2.0, 3.0 * 0.125, 3.0 * 1.0]
    assert o.target[:, 1].tolist() == cnwt[:, 0].tolist()
    assert o.target[:, 2].tolist() == cnfx[:, 0].tolist()
    assert o.target[:, 3].tolist() == cnfy[:, 0].tolist()


def test_oracle_getacc_matches_numpy():
    p = getacc_problem(9, 7, seed=1)
    o = oracle_of(p, O.GETACC)
    o.loop_serial(2)
    cells = p.maps[0]
    cnmass, rho, cnwt, cnfx, cnfy = p.cell_data
    mass = np.empty_like(cnmass)
    for jj in range(4):
        ii = (jj + 3) % 4
        mass[jj] = np.where(cnmass[jj] > 1e-40, cnmass[jj],
                            np.where(cnmass[ii] > 1e-40, cnmass[ii], rho[0] * cnwt[jj]))
    out = np.zeros((p.num_points[0], 4))
    for _ in range(2):
        for d, src in enumerate([mass, cnwt, cnfx, cnfy]):
            col = out[:, d].copy()
            np.add.at(col, cells.ravel(), src.T.ravel())  # cell by cell, corners in order
            out[:, d] = col
    assert_bitwise(o.target, out, "oracle getacc vs numpy")


@pytest.mark.gpu
@pytest.mark.parametrize("soa", [False, True])
def test_getacc_gpu_strategies(gpu, soa):
    iters, bs = 2, 128
    base = getacc_problem(60, 45, seed=2, block_size=bs, soa=soa)
    ref = oracle_of(base, O.GETACC)
    S = ref.loop_serial(iters, with_scale=True)
    colour = oracle_of(getacc_problem(60, 45, seed=2, soa=soa), O.GETACC)
    colour.loop_colour(iters, 1)
    owner = oracle_of(getacc_problem(60, 45, seed=2, soa=soa), O.GETACC)
    O.loop_owner_replay(owner, plan.owner_plan(base.maps[0], base.num_points[0], bs), iters)
    hier = oracle_of(getacc_problem(60, 45, seed=2, soa=soa), O.GETACC)
    hier.loop_hier_replay(O.OraclePlan(base.maps, base.num_points[0], bs), iters)
    for strategy, flags, expect in [(capi.STRATEGY_ATOMIC, 0, None),
                                    (capi.STRATEGY_GLOBAL_COLOUR, 0, colour.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_OWNER_COMPUTE,
                                     owner.target),
                                    (capi.STRATEGY_HIERARCHICAL, 0, hier.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_TWO_PHASE,
                                     hier.target)]:
        p = getacc_problem(60, 45, seed=2, block_size=bs, soa=soa)
        p.gpu_loop(GETACC, strategy, iters, flags=flags)
        assert_close_scaled(p.point_data[0], ref.target, S, what=f"getacc {strategy}/{flags}")
        if expect is not None:
            assert_bitwise(p.point_data[0], expect, f"getacc {strategy}/{flags} bits")
