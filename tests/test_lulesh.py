"""lulesh's element forces (kernels/lulesh_func.hpp:13-38 with
kernels/lulesh_utils.hpp): the oracle pinned by closed forms (unit cube,
affine elements: volume = det A and face normals from the diagonals'
cross product; exact power-of-two scaling; closed-surface normal sums),
libfot's host copy of the element function checked bit for bit against it,
and every GPU strategy against the oracle (forces and element volumes)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, lulesh_problem, plan
from cuda_fot_amd.capi import ptr
from helpers import assert_bitwise, assert_close_scaled, oracle_of

LULESH = capi.functor(capi.KERNEL_LULESH)
# corners of the reference element in lulesh's node order
CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)
FACES = [(0, 1, 2, 3), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0), (4, 7, 6, 5)]


def _normals(xyz):
    # forces with sigma = -1 are the node normals: f = -(sigma * n)
    f, det = O.lulesh_func(xyz, [-1.0, -1.0, -1.0])
    return f, det


def _diag_normals(p):
    """Per-node sum over the node's faces of -1/8 (p_c - p_a) x (p_d - p_b):
    a quarter of each face's (inward) area vector, from the diagonals."""
    n = np.zeros((8, 3))
    for a, b, c, d in FACES:
        v = -0.125 * np.cross(p[c] - p[a], p[d] - p[b])
        for k in (a, b, c, d):
            n[k] += v
    return n


def test_unit_cube():
    n, det = _normals(CUBE)
    assert det == 1.0
    # each corner: a quarter of the three unit faces it touches, outward
    np.testing.assert_array_equal(n, 0.25 * (2 * CUBE - 1))


def test_affine_elements_volume_and_face_normals():
    rng = np.random.default_rng(0)
    for t in range(40):
        A = rng.normal(size=(3, 3))
        c = rng.normal(size=3) * 10
        p = CUBE @ A.T + c
        n, det = _normals(p)
        vol = np.linalg.det(A)
        assert abs(det - vol) <= 1e-12 * (abs(vol) + np.abs(A).max() ** 3), (t, det, vol)
        ref = _diag_normals(p)
        assert np.abs(n - ref).max() <= 1e-12 * (1 + np.abs(p).max()) ** 2, t


def test_power_of_two_scaling_is_exact_and_normals_close():
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = CUBE + rng.uniform(-0.2, 0.2, size=(8, 3))  # a general (non-planar) hex
        sig = rng.normal(size=3)
        f, det = O.lulesh_func(p, sig)
        for s in (2.0, 0.5, 8.0):
            fs, ds = O.lulesh_func(s * p, sig)
            assert ds == s ** 3 * det
            np.testing.assert_array_equal(fs, s ** 2 * f)
        n, _ = _normals(p)
        assert np.abs(n.sum(axis=0)).max() <= 1e-15 * 8
        np.testing.assert_allclose(n, _diag_normals(p), rtol=0, atol=1e-15)


def test_libfot_host_element_matches_oracle():
    rng = np.random.default_rng(2)
    f = np.zeros(24)
    det = np.zeros(1)
    for _ in range(50):
        p = CUBE @ rng.normal(size=(3, 3)).T + rng.uniform(-0.3, 0.3, size=(8, 3))
        sig = rng.normal(size=3)
        x, y, z = (np.ascontiguousarray(p[:, d]) for d in range(3))
        capi.call("fot_lulesh_element", ptr(x), ptr(y), ptr(z), ptr(sig), 1, ptr(f), ptr(det))
        fo, do = O.lulesh_func(p, sig)
        assert_bitwise(f.reshape(3, 8).T, fo, "libfot host element vs oracle")
        assert det[0] == do


def test_oracle_loop_balances_forces_and_writes_volumes():
    p = lulesh_problem(6, 5, 7, seed=3)
    o = oracle_of(p, O.LULESH)
    f0 = o.target.copy()
    o.loop_serial(1)
    # each element's corner forces sum to zero (closed surface)
    S = np.abs(f0).sum() + 1.0
    assert np.abs((o.target - f0).sum(axis=0)).max() <= 1e-13 * S
    det = o.cell_data[1][0]
    # the generator's corner order is lulesh's with x and y swapped: the
    # volumes are negative, of the jittered unit cells' size
    assert np.all(det < -0.2) and np.all(det > -2.5)
    assert np.array_equal(p.cell_data[1], np.zeros_like(p.cell_data[1]))  # oracle copies


@pytest.mark.gpu
@pytest.mark.parametrize("soa", [False, True])
def test_lulesh_gpu_strategies(gpu, soa):
    iters, bs = 2, 128
    base = lulesh_problem(14, 12, 11, seed=4, block_size=bs, soa=soa)
    ser = oracle_of(base, O.LULESH)
    S = ser.loop_serial(iters, with_scale=True)
    det_ref = ser.cell_data[1]
    colour = oracle_of(lulesh_problem(14, 12, 11, seed=4, soa=soa), O.LULESH)
    colour.loop_colour(iters, 1)
    owner = oracle_of(lulesh_problem(14, 12, 11, seed=4, soa=soa), O.LULESH)
    O.loop_owner_replay(owner, plan.owner_plan(base.maps[0], base.num_points[0], bs), iters)
    hier = oracle_of(lulesh_problem(14, 12, 11, seed=4, soa=soa), O.LULESH)
    hier.loop_hier_replay(O.OraclePlan(base.maps, base.num_points[0], bs, None), iters)
    for strategy, flags, expect in [(capi.STRATEGY_ATOMIC, 0, None),
                                    (capi.STRATEGY_GLOBAL_COLOUR, 0, colour.target),
                                    (capi.STRATEGY_HIERARCHICAL, 0, hier.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_DEPENDENCY_ORDER,
                                     hier.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_TWO_PHASE,
                                     hier.target),
                                    (capi.STRATEGY_HIERARCHICAL,
                                     capi.LOOP_TWO_PHASE | capi.LOOP_STAGE_INPUTS, hier.target),
                                    # the stresses loaded with nontemporal hints
                                    (capi.STRATEGY_HIERARCHICAL,
                                     capi.LOOP_TWO_PHASE | capi.LOOP_NONTEMPORAL, hier.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_OWNER_COMPUTE,
                                     owner.target)]:
        p = lulesh_problem(14, 12, 11, seed=4, block_size=bs, soa=soa)
        st = p.gpu_loop(LULESH, strategy, iters, flags=flags, partition=False)
        what = f"lulesh {strategy}/{flags}"
        assert st["nontemporal"] == int(bool(flags & capi.LOOP_NONTEMPORAL))
        assert_close_scaled(p.point_data[0], ser.target, S, what=what)
        if expect is not None:
            assert_bitwise(p.point_data[0], expect, what + " bits")
        # the element volume does not depend on the summation order
        assert_bitwise(p.cell_data[1], det_ref, what + " volumes")
