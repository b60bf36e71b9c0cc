"""The shallow-water edge loop (volna, kernels/volna.hpp in FP64): the
triangle-mesh generator's invariants, the oracle against an independent
numpy restatement and the closed forms of the flux update (CPU), and the GPU
strategies against the oracle (gpu)."""
import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, meshgen, plan, volna_problem
from helpers import assert_bitwise, assert_close_scaled, oracle_of

VOLNA = capi.functor(capi.KERNEL_VOLNA)


def test_tri_mesh_invariants():
    nx, ny = 13, 9
    m = meshgen.tri_mesh(nx, ny, seed=4)
    ec, bnd = m["edge_cells"], m["is_boundary"]
    assert ec.shape == (nx * (ny + 1) + (nx + 1) * ny + nx * ny, 2)
    assert int(bnd.sum()) == 2 * (nx + ny)
    assert np.all(ec[bnd == 1, 0] == ec[bnd == 1, 1])
    assert np.all(ec[bnd == 0, 0] != ec[bnd == 0, 1])
    # every triangle has three edges
    deg = np.bincount(ec[:, 0], minlength=m["num_cells"]) + \
        np.bincount(ec[bnd == 0, 1], minlength=m["num_cells"])
    assert np.all(deg == 3)
    np.testing.assert_allclose(np.hypot(*m["normals"]), 1.0, rtol=1e-14)
    assert np.all(m["volume"] > 0)
    # the perturbed grid's area is nx * ny up to the boundary perturbation
    assert abs(m["volume"].sum() - nx * ny) < 0.2 * 2 * (nx + ny)
    s = m["state"]
    assert np.all((s[:, 0] >= 0.5) & (s[:, 0] < 2.0)) and np.all((s[:, 3] > -1) & (s[:, 3] <= 0))


def _numpy_volna(m, iters):
    """volna_func.hpp:16-36 restated with numpy and np.add.at, per edge
    left then right, component by component (volna.hpp:69-73)."""
    ec, fl, bs, nm = m["edge_cells"], m["fluxes"], m["bathy"], m["normals"]
    v0, v1 = m["volume"][ec[:, 0]], m["volume"][ec[:, 1]]
    interior = m["is_boundary"] == 0
    zero = np.zeros(ec.shape[0])
    left = [zero - fl[0] / v0, zero - (fl[1] + bs[0] * nm[0]) / v0,
            zero - (fl[2] + bs[0] * nm[1]) / v0, zero.copy()]
    right = [np.where(interior, zero + fl[0] / v1, 0.0),
             np.where(interior, zero + (fl[1] + bs[1] * nm[0]) / v1, 0.0),
             np.where(interior, zero + (fl[2] + bs[1] * nm[1]) / v1, 0.0), zero.copy()]
    out = m["state"].copy()
    idx = ec.ravel()  # l0 r0 l1 r1 ...
    for _ in range(iters):
        for d in range(4):
            col = out[:, d].copy()
            np.add.at(col, idx, np.stack([left[d], right[d]], axis=1).ravel())
            out[:, d] = col
    return out


def test_oracle_volna_matches_numpy():
    m = meshgen.tri_mesh(11, 7, seed=2)
    o = oracle_of(volna_problem(m), O.VOLNA)
    o.loop_serial(2)
    assert_bitwise(o.target, _numpy_volna(m, 2), "oracle volna vs numpy")


def test_volna_closed_forms():
    m = meshgen.tri_mesh(16, 12, seed=6)
    p = volna_problem(m)
    o = oracle_of(p, O.VOLNA)
    S = o.loop_serial(1, with_scale=True)
    inc = o.target - m["state"]
    # the bed level (4th component) gets no increment
    assert_bitwise(o.target[:, 3], m["state"][:, 3], "zb untouched")
    # mass: interior edges move f from left to right exactly once each, so
    # sum(vol * dh) = -sum over boundary edges of f0
    mass = (m["volume"] * inc[:, 0]).sum()
    bflux = m["fluxes"][0][m["is_boundary"] == 1].sum()
    assert abs(mass + bflux) <= 1e-12 * (m["volume"] * S[:, 0]).sum()
    # zero fluxes and zero bathymetry source: nothing changes
    z = dict(m, fluxes=np.zeros_like(m["fluxes"]), bathy=np.zeros_like(m["bathy"]))
    oz = oracle_of(volna_problem(z), O.VOLNA)
    oz.loop_serial(2)
    np.testing.assert_array_equal(oz.target, m["state"])


def _strategies(p, kernel, iters, block_size):
    ref = oracle_of(p, kernel)
    S = ref.loop_serial(iters, with_scale=True)
    return ref.target.copy(), S


@pytest.mark.gpu
@pytest.mark.parametrize("soa", [False, True])
def test_volna_gpu_strategies(gpu, soa):
    m = meshgen.tri_mesh(40, 30, seed=1)
    iters, bs = 3, 128
    base = volna_problem(m, block_size=bs, soa=soa)
    serial, S = _strategies(base, O.VOLNA, iters, bs)
    colour = oracle_of(volna_problem(m, soa=soa), O.VOLNA)
    colour.loop_colour(iters, 1)
    replay = oracle_of(volna_problem(m, soa=soa), O.VOLNA)
    replay.loop_hier_replay(O.OraclePlan(base.maps, base.num_points[0], bs), iters)
    owner = oracle_of(volna_problem(m, soa=soa), O.VOLNA)
    O.loop_owner_replay(owner, plan.owner_plan(base.maps[0], base.num_points[0], bs), iters)

    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_ATOMIC, iters)
    assert_close_scaled(p.point_data[0], serial, S, what="volna atomics")
    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_GLOBAL_COLOUR, iters)
    assert_bitwise(p.point_data[0], colour.target, "volna global colour")
    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_HIERARCHICAL, iters)
    assert_bitwise(p.point_data[0], replay.target, "volna hierarchical")
    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_TWO_PHASE)
    assert_bitwise(p.point_data[0], replay.target, "volna hierarchical two-phase")
    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_HIERARCHICAL, iters,
               flags=capi.LOOP_TWO_PHASE | capi.LOOP_STAGE_INPUTS)
    assert_bitwise(p.point_data[0], replay.target, "volna hierarchical staged two-phase")
    p = volna_problem(m, block_size=bs, soa=soa)
    p.gpu_loop(VOLNA, capi.STRATEGY_HIERARCHICAL, iters, flags=capi.LOOP_OWNER_COMPUTE)
    assert_bitwise(p.point_data[0], owner.target, "volna owner compute")
    assert_close_scaled(p.point_data[0], serial, S, what="volna owner vs serial")
