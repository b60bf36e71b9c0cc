"""The oracle's loops against an independent numpy statement of the same
edge loop (np.add.at), against each other within the reordering bound, and
the invariants of res_calc on the generated airfoil mesh.  CPU only."""
import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import meshgen, res_calc_problem
from helpers import (airfoil_problem, assert_bitwise, assert_close_scaled, hex_problem,
                     mine_problem, oracle_of)


def _numpy_res_calc(p, iters):
    """The edge loop restated with numpy gathers and np.add.at (AOS)."""
    pecell, pedge = p.maps[0], p.maps[1]
    res, x, q, adt = (d.copy() for d in p.point_data)
    for _ in range(iters):
        inc1 = np.zeros((pecell.shape[0], 4))
        for e in range(pecell.shape[0]):
            r1, _ = O.res_calc_func(x[pedge[e, 0]], x[pedge[e, 1]], q[pecell[e, 0]],
                                    q[pecell[e, 1]], adt[pecell[e, 0]], adt[pecell[e, 1]])
            inc1[e] = r1
        np.add.at(res, pecell[:, 0], inc1)
        np.add.at(res, pecell[:, 1], -inc1)
    return res


def test_serial_res_calc_matches_numpy_loop():
    p = airfoil_problem(nx=12, ny=8, res0_seed=1)
    o = oracle_of(p)
    S = o.loop_serial(2, with_scale=True)
    assert_close_scaled(o.target, _numpy_res_calc(p, 2), S, what="serial vs numpy")


def test_serial_mine_matches_numpy_loop():
    p = mine_problem(N=8, M=7, P=4, C=4)
    o = oracle_of(p, O.MINE2)
    S = o.loop_serial(3, with_scale=True)
    e, w = p.maps[0], p.cell_data[0]
    pd = p.point_data[0].copy()
    for _ in range(3):
        out = pd.copy()
        np.add.at(out, e[:, 0], 2 * pd[e[:, 1]] * w.T)
        np.add.at(out, e[:, 1], 2 * pd[e[:, 0]] * w.T)
        pd = out
    assert_close_scaled(o.target, pd, S, what="mine serial vs numpy")


def test_serial_hex_matches_numpy_loop():
    p = hex_problem(n=4)
    o = oracle_of(p, O.HEX_SCATTER)
    S = o.loop_serial(2, with_scale=True)
    F = p.cell_data[0].reshape(3, 8, -1)  # [i * 8 + j]
    f = p.point_data[0].copy()
    for _ in range(2):
        for j in range(8):
            np.add.at(f, p.maps[0][:, j], F[:, j, :].T)
    assert_close_scaled(o.target, f, S, what="hex serial vs numpy")


@pytest.mark.parametrize("kernel,make", [(O.RES_CALC, lambda soa: airfoil_problem(soa=soa,
                                                                                   res0_seed=2)),
                                         (O.MINE2, lambda soa: mine_problem(soa=soa))])
def test_aos_and_soa_give_identical_bits(kernel, make):
    a, b = make(False), make(True)
    oa, ob = oracle_of(a, kernel), oracle_of(b, kernel)
    oa.loop_serial(3)
    ob.loop_serial(3)
    assert_bitwise(oa.target, ob.target.T, "AOS vs SOA")


def test_loops_agree_within_reordering_bound():
    for make, k in [(lambda: airfoil_problem(res0_seed=3), O.RES_CALC),
                    (lambda: mine_problem(), O.MINE2), (lambda: hex_problem(), O.HEX_SCATTER)]:
        s = oracle_of(make(), k)
        S = s.loop_serial(3, with_scale=True)
        c = oracle_of(make(), k)
        c.loop_colour(3, 1)
        c4 = oracle_of(make(), k)
        c4.loop_colour(3, 4)
        assert_bitwise(c4.target, c.target, "colour loop thread count")
        assert_close_scaled(c.target, s.target, S, what="colour vs serial")
        p = make()
        h = oracle_of(p, k)
        h.loop_hier_replay(O.OraclePlan(p.maps, p.num_points[0], 64), 3)
        assert_close_scaled(h.target, s.target, S, what="replay vs serial")
        assert np.all(S >= np.abs(s.target) * (1 - 1e-12))


def test_res_calc_conserves_and_is_zero_for_uniform_flow():
    m = meshgen.airfoil(24, 16, seed=5)
    # uniform state: every interior cell's fluxes cancel around its closed
    # boundary; only cells on the mesh boundary (missing boundary edges) keep
    # a net flux
    m["q"][:] = np.array([1.0, 0.4, 0.0, 2.5])
    p = res_calc_problem(m)
    o = oracle_of(p)
    S = o.loop_serial(1, with_scale=True)
    deg = np.bincount(m["pecell"].ravel(), minlength=m["num_cells"])
    interior = deg == 4
    assert interior.sum() == (24 - 2) * (16 - 2)
    assert np.all(np.abs(o.target[interior]) <= 1e-13 * S[interior])
    assert np.abs(o.target[~interior]).max() > 1e-3
    # conservation: every edge adds f and -f, so the total is zero up to rounding
    p2 = airfoil_problem(nx=24, ny=16, res0_seed=None)
    o2 = oracle_of(p2)
    S2 = o2.loop_serial(2, with_scale=True)
    assert np.all(np.abs(o2.target.sum(axis=0)) <= 1e-13 * S2.sum(axis=0))
