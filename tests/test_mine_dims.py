"""mine on quads (MeshDim 4) and hexes (MeshDim 8), the reference's own test
matrix (tests.hpp:636-807: PointDim 1, 2, 4, 8, CellDim 1 or PointDim, AOS
and SOA) in every GPU strategy, against the oracle's MINE4 / MINE8; and the
hand-derived known answers of mine_func (kernels/mine_func.hpp:11-99) run
through the GPU loops themselves."""
import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import Problem, capi, meshgen, plan
from helpers import assert_bitwise, assert_close_scaled, oracle_of
from kats import kat_mine2, kat_mine4, kat_mine8

pytestmark = pytest.mark.gpu

ITERS = 3
OWN = capi.LOOP_OWNER_COMPUTE
DEP = capi.LOOP_DEPENDENCY_ORDER
TP = capi.LOOP_TWO_PHASE
ORACLE_KERNEL = {2: O.MINE2, 4: O.MINE4, 8: O.MINE8}


def _mesh(md, seed):
    if md == 4:
        cells = meshgen.quad_grid(15, 13)
    else:
        cells = meshgen.hex_grid(7, 6, 6)
    cells = cells[meshgen.permutation(cells.shape[0], seed)]
    return cells, int(cells.max()) + 1


def _problem(md, P, C, soa, block_size=64, seed=3):
    cells, n = _mesh(md, seed)
    pd = meshgen.uniform(n * P, seed + 1).reshape(n, P)
    w = meshgen.uniform(cells.shape[0] * C, seed + 2).reshape(C, cells.shape[0])
    if soa:
        pd = pd.T.copy()
    return Problem([cells], [n], [pd], [w], block_size, soa)


MATRIX = [(md, P, C, soa) for md in (4, 8) for P in (1, 2, 4, 8) for C in sorted({1, P})
          for soa in (False, True)]


@pytest.mark.parametrize("md,P,C,soa", MATRIX,
                         ids=[f"md{m}-p{p}-c{c}-{'soa' if s else 'aos'}" for m, p, c, s in MATRIX])
def test_mine_dims_all_strategies(gpu, md, P, C, soa):
    k = ORACLE_KERNEL[md]
    f = capi.functor(capi.KERNEL_MINE, md, P, C)
    o = oracle_of(_problem(md, P, C, soa), k)
    S = o.loop_serial(ITERS, with_scale=True)
    serial = o.target.copy()
    oc = oracle_of(_problem(md, P, C, soa), k)
    oc.loop_colour(ITERS, 1)
    base = _problem(md, P, C, soa)
    oh = oracle_of(base, k)
    oh.loop_hier_replay(O.OraclePlan(base.maps, base.num_points[0], 64), ITERS)
    oo = oracle_of(base, k)
    O.loop_owner_replay(oo, plan.owner_plan(base.maps[0], base.num_points[0], 64), ITERS)

    g = _problem(md, P, C, soa)
    g.gpu_loop(f, capi.STRATEGY_GLOBAL_COLOUR, ITERS)
    assert_bitwise(g.point_data[0], oc.target, "global colour vs oracle colour loop")
    for flags in (0, DEP, TP, capi.LOOP_STAGE_INPUTS | TP):
        # (fot_loop_run reports a dependency wait that ran out as an error)
        h = _problem(md, P, C, soa)
        h.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, ITERS, flags=flags)
        assert_bitwise(h.point_data[0], oh.target, f"hierarchical (flags {flags}) vs replay")
    w = _problem(md, P, C, soa)
    w.gpu_loop(f, capi.STRATEGY_HIERARCHICAL, ITERS, flags=OWN)
    assert_bitwise(w.point_data[0], oo.target, "owner compute vs owner replay")
    a = _problem(md, P, C, soa)
    a.gpu_loop(f, capi.STRATEGY_ATOMIC, ITERS)
    assert_close_scaled(a.point_data[0], serial, S, what="atomics vs serial")
    # and the fixed-order results against the serial loop within the bound
    assert_close_scaled(g.point_data[0], serial, S, what="global colour vs serial")


STRATEGIES = [(capi.STRATEGY_ATOMIC, 0), (capi.STRATEGY_GLOBAL_COLOUR, 0),
              (capi.STRATEGY_HIERARCHICAL, 0), (capi.STRATEGY_HIERARCHICAL, DEP),
              (capi.STRATEGY_HIERARCHICAL, TP), (capi.STRATEGY_HIERARCHICAL, OWN)]


@pytest.mark.parametrize("kat", [kat_mine2, kat_mine4, kat_mine8],
                         ids=["mine2", "mine4", "mine8"])
@pytest.mark.parametrize("strategy,flags", STRATEGIES,
                         ids=["atomic", "global", "hier", "hier-dep", "hier-2phase", "owner"])
def test_mine_known_answers_on_the_gpu(gpu, kat, strategy, flags):
    md_of = {O.MINE2: 2, O.MINE4: 4, O.MINE8: 8}

    def run(kernel, maps, npts, pd, cd, iters, soa):
        md = md_of[kernel]
        P = pd.shape[0] if soa else pd.shape[1]
        C = cd[0].shape[0]
        p = Problem(maps, npts, [pd], cd, 64, soa)
        p.gpu_loop(capi.functor(capi.KERNEL_MINE, md, P, C), strategy, iters, flags=flags)
        return p.point_data[0]
    kat(run)
