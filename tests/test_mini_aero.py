"""mini_aero's face flux (kernels/mini_aero_func.hpp, mini_aero_utils.hpp):
the oracle pinned by closed forms (uniform state = Euler flux, swap
antisymmetry, conservation), libfot's host copy of the user function checked
bit for bit against it, and the GPU strategies against the oracle."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import capi, meshgen, mini_aero_problem, plan
from cuda_fot_amd.capi import ptr
from helpers import assert_bitwise, assert_close_scaled, oracle_of

MINI_AERO = capi.functor(capi.KERNEL_MINI_AERO)
RGAS, GAMMA = 287.05, 1.4


def _one_face(left, right, face):
    """Flux of one face through the oracle: a 2-cell, 1-face problem."""
    cells = np.stack([left, right])
    p = O.OracleProblem(O.MINI_AERO, [np.array([[0, 1]]), np.array([[0, 1]])], [2, 2],
                        [np.zeros((2, 5)), cells], [face.reshape(12, 1)])
    p.loop_serial(1)
    return p.target[1].copy()   # right cell got +flux


def _state(rho, u, v, w, T):
    e = RGAS * T / (GAMMA - 1.0)
    return np.array([rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))])


def _cell(U, centre=(0.5, 0.5, 0.5), grads=None, lim=None):
    g = np.zeros(15) if grads is None else grads
    li = np.ones(5) if lim is None else lim
    return np.concatenate([U, centre, g, li])


def _face(n, t, b, centre=(1.0, 0.5, 0.5)):
    return np.concatenate([centre, n, t, b])


def test_uniform_state_gives_the_euler_flux():
    rho, u, v, w, T = 1.2, 40.0, -7.0, 3.0, 290.0
    U = _state(rho, u, v, w, T)
    n = np.array([0.6, 0.8, 0.0])
    f = _one_face(_cell(U), _cell(U), _face(n, [-0.8, 0.6, 0.0], [0.0, 0.0, 1.0]))
    # mini_aero's primitive temperature and pressure of this state
    ri = 1.0 / rho
    Tp = (U[4] * ri - 0.5 * (u * u + v * v + w * w)) * (GAMMA - 1.0) / RGAS
    p, H = rho * RGAS * Tp, 1004.0 * Tp + 0.5 * (u * u + v * v + w * w)
    un = u * n[0] + v * n[1] + w * n[2]
    euler = np.array([rho * un, rho * u * un + p * n[0], rho * v * un + p * n[1],
                      rho * w * un + p * n[2], rho * H * un])
    np.testing.assert_allclose(f, euler, rtol=1e-12, atol=1e-9)


def test_swapping_sides_negates_the_flux():
    rng = np.random.default_rng(0)
    for _ in range(10):
        L = _cell(_state(1 + 0.1 * rng.random(), *(40 * rng.normal(size=3)), 300 + rng.random()),
                  grads=1e-3 * rng.normal(size=15), lim=rng.random(5))
        R = _cell(_state(1 + 0.1 * rng.random(), *(40 * rng.normal(size=3)), 300 + rng.random()),
                  centre=(1.5, 0.5, 0.5), grads=1e-3 * rng.normal(size=15), lim=rng.random(5))
        n, t, b = np.eye(3)
        f = _one_face(L, R, _face(n, t, b))
        g = _one_face(R, L, _face(-n, -t, -b))
        # reordered rounding only: relative to the largest flux component
        assert np.all(np.abs(f + g) <= 1e-12 * np.abs(f).max()), (f, g)


def test_oracle_conserves_and_matches_libfot_host_flux():
    m = meshgen.hex_faces(5, 4, 3, seed=2)
    p = mini_aero_problem(m)
    o = oracle_of(p, O.MINI_AERO)
    S = o.loop_serial(1, with_scale=True)
    assert np.all(np.abs(o.target.sum(0)) <= 1e-13 * S.sum(0))
    # libfot's host copy of the user function, face by face, in the oracle's order
    out = np.zeros_like(o.target)
    cells, face = np.ascontiguousarray(m["cells"]), m["face"]
    flux = np.empty(5)
    for f, (l, r) in enumerate(m["face_cells"]):
        fd = np.ascontiguousarray(face[:, f])
        capi.call("fot_mini_aero_face_flux", ptr(cells[l]), ptr(cells[r]), 1, ptr(fd), 1,
                  ptr(flux))
        for c in range(5):
            out[l, c] = out[l, c] + (-flux[c])
            out[r, c] = out[r, c] + flux[c]
    assert_bitwise(out, o.target, "libfot host flux vs oracle")


@pytest.mark.gpu
@pytest.mark.parametrize("soa", [False, True])
def test_mini_aero_gpu_strategies(gpu, soa):
    m = meshgen.hex_faces(12, 10, 9, seed=3)
    iters, bs = 2, 128
    base = mini_aero_problem(m, block_size=bs, soa=soa)
    ref = oracle_of(base, O.MINI_AERO)
    S = ref.loop_serial(iters, with_scale=True)
    colour = oracle_of(mini_aero_problem(m, soa=soa), O.MINI_AERO)
    colour.loop_colour(iters, 1)
    owner = oracle_of(mini_aero_problem(m, soa=soa), O.MINI_AERO)
    O.loop_owner_replay(owner, plan.owner_plan(base.maps[0], base.num_points[0], bs), iters)
    for strategy, flags, expect in [(capi.STRATEGY_ATOMIC, 0, None),
                                    (capi.STRATEGY_GLOBAL_COLOUR, 0, colour.target),
                                    (capi.STRATEGY_HIERARCHICAL, capi.LOOP_OWNER_COMPUTE,
                                     owner.target),
                                    (capi.STRATEGY_HIERARCHICAL, 0, None)]:
        p = mini_aero_problem(m, block_size=bs, soa=soa)
        p.gpu_loop(MINI_AERO, strategy, iters, flags=flags)
        assert_close_scaled(p.point_data[0], ref.target, S, what=f"mini_aero {strategy}/{flags}")
        if expect is not None:
            assert_bitwise(p.point_data[0], expect, f"mini_aero {strategy}/{flags} bits")


@pytest.mark.gpu
@pytest.mark.parametrize("soa", [False, True])
@pytest.mark.parametrize("bs", [128, 512, 1024])
@pytest.mark.parametrize("flags", [0, capi.LOOP_DEPENDENCY_ORDER, capi.LOOP_TWO_PHASE,
                                   capi.LOOP_STAGE_INPUTS,
                                   capi.LOOP_STAGE_INPUTS | capi.LOOP_TWO_PHASE,
                                   capi.LOOP_TWO_PHASE | capi.LOOP_NONTEMPORAL,
                                   capi.LOOP_STAGE_INPUTS | capi.LOOP_TWO_PHASE |
                                   capi.LOOP_NONTEMPORAL],
                         ids=["per-colour", "dependency-order", "two-phase", "staged",
                              "staged-two-phase", "two-phase-nontemporal",
                              "staged-two-phase-nontemporal"])
def test_mini_aero_hierarchical_bitwise(gpu, soa, bs, flags):
    """The reference-shaped hierarchical loop, one launch per block colour,
    one dependency-ordered launch or the two-phase writeback, the cell values
    gathered per face or staged once per block in LDS, the face values
    loaded plainly or with nontemporal hints, at block sizes that run
    the 512-thread instantiation (128, 512) and the 1024-thread one (1024):
    bit for bit against the oracle's replay of the same plan.  fot_loop_run reports a
    dependency wait that ran out as an error, so a pass also shows that every
    wait ended."""
    m = meshgen.hex_faces(12, 10, 9, seed=3)
    iters = 2
    base = mini_aero_problem(m, block_size=bs, soa=soa)
    ref = oracle_of(base, O.MINI_AERO)
    ref.loop_hier_replay(O.OraclePlan(base.maps, base.num_points[0], bs), iters)
    p = mini_aero_problem(m, block_size=bs, soa=soa)
    st = p.gpu_loop(MINI_AERO, capi.STRATEGY_HIERARCHICAL, iters, flags=flags)
    assert_bitwise(p.point_data[0], ref.target, f"mini_aero hierarchical b{bs} flags {flags}")
    # the nontemporal face-value loads: two-phase with the values gathered
    # per face (when staging fits, the staged form has none)
    if flags & capi.LOOP_NONTEMPORAL and not flags & capi.LOOP_STAGE_INPUTS:
        assert st["nontemporal"] == 1
    if not flags & capi.LOOP_NONTEMPORAL:
        assert st["nontemporal"] == 0
