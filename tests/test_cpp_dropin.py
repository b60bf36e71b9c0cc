"""The C++ drop-in (include/fot/problem.hpp, kernels.hpp, data.hpp) driven
like the reference's Problem<SOA> by tests/cpp/dropin_test.cpp; its outputs
are checked here against the oracle on the same seeded inputs.  The CPU test
also covers the mesh / point / cell / partition stream formats (self-checks
inside the test program)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import meshgen
from helpers import assert_bitwise, assert_close_scaled

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "bin", "dropin_test")
ITERS_RES, ITERS_MINE, ITERS_HEX = 3, 3, 2


def _run(mode, out):
    r = subprocess.run([BIN, mode, str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _load(out, name, dim):
    return np.fromfile(os.path.join(out, name + ".bin"), dtype=np.float64).reshape(-1, dim)


def _airfoil_oracle():
    # the test program's airfoil: 40 x 24 cells, seed 3, shuffled, res0 U[-1,1) seed 11
    m = meshgen.airfoil(40, 24, seed=3, shuffle=7)
    res0 = meshgen.uniform(m["num_cells"] * 4, 11, -1.0, 1.0).reshape(-1, 4)
    maps = [m["pecell"], m["pedge"], m["pecell"], m["pecell"]]
    npts = [m["num_cells"], m["num_nodes"], m["num_cells"], m["num_cells"]]
    return lambda: O.OracleProblem(O.RES_CALC, maps, npts, [res0, m["x"], m["q"], m["adt"]])


def _mine_oracle():
    N, M = 20, 17
    e = meshgen.grid_edges(N, M)
    pd = meshgen.uniform(N * M * 4, 5).reshape(-1, 4)
    w = meshgen.uniform(e.shape[0] * 4, 6).reshape(4, -1)
    return lambda: O.OracleProblem(O.MINE2, [e], [N * M], [pd], [w])


def _hex_oracle():
    c = meshgen.hex_grid(5, 4, 3)
    F = meshgen.normal(c.shape[0] * 24, 9).reshape(24, -1)
    return lambda: O.OracleProblem(O.HEX_SCATTER, [c], [60], [np.zeros((60, 3))], [F])


def _volna_oracle():
    from cuda_fot_amd import volna_problem
    m = meshgen.tri_mesh(14, 9, seed=4)
    p = volna_problem(m)
    return lambda: O.OracleProblem(O.VOLNA, p.maps, p.num_points, p.point_data, p.cell_data)


def _mini_aero_oracle():
    from cuda_fot_amd import mini_aero_problem
    p = mini_aero_problem(meshgen.hex_faces(5, 4, 3, seed=2))
    return lambda: O.OracleProblem(O.MINI_AERO, p.maps, p.num_points, p.point_data, p.cell_data)


def _lulesh_oracle():
    # the test program's lulesh: 5 x 4 x 3 node lattice, jitter seed 21, forces
    # seed 22, sigma seed 23, volumes zero
    n1, n2, n3 = 5, 4, 3
    c = meshgen.hex_grid(n1, n2, n3)
    npt, nc = n1 * n2 * n3, c.shape[0]
    l, r, cc = np.meshgrid(np.arange(n3), np.arange(n1), np.arange(n2), indexing="ij")
    xyz = np.stack([r.ravel(), cc.ravel(), l.ravel()], axis=1).astype(np.float64)
    xyz += meshgen.uniform(3 * npt, 21, -0.2, 0.2).reshape(npt, 3)
    f = meshgen.normal(3 * npt, 22).reshape(npt, 3)
    sig = meshgen.normal(3 * nc, 23).reshape(3, nc)
    return lambda: O.OracleProblem(O.LULESH, [c, c], [npt, npt], [f, xyz],
                                   [sig, np.zeros((1, nc))])


def _serial(make, iters):
    o = make()
    S = o.loop_serial(iters, with_scale=True)
    return o.target.copy(), S


def _colour(make, iters):
    o = make()
    o.loop_colour(iters, 1)
    return o.target.copy()


@pytest.fixture(scope="module")
def built():
    # CPU-only build (g++ against libfot; a no-op when up to date);
    # `__graft_entry__.build()` makes it too
    subprocess.run(["make", "-C", os.path.join(HERE, "cpp")], check=True, capture_output=True)
    return BIN


def test_cpp_dropin_cpu_loops(built, tmp_path):
    out = _run("cpu", tmp_path)
    assert "io round trips ok" in out
    air, mine, hexp = _airfoil_oracle(), _mine_oracle(), _hex_oracle()
    serial, S = _serial(air, ITERS_RES)
    colour = _colour(air, ITERS_RES)
    for tag in ("aos", "soa"):
        # the same user function in the same order: identical bits
        assert_bitwise(_load(tmp_path, "res_serial_" + tag, 4), serial, "C++ serial " + tag)
        assert_bitwise(_load(tmp_path, "res_omp_" + tag, 4), colour, "C++ OMP " + tag)
        # reordered / partitioned problems: same result within the FP bound
        assert_close_scaled(_load(tmp_path, "res_reordered_" + tag, 4), serial, S,
                            what="C++ reordered " + tag)
        assert_close_scaled(_load(tmp_path, "res_ours_" + tag, 4), serial, S,
                            what="C++ partitionOurs " + tag)
    ms, _ = _serial(mine, ITERS_MINE)
    hs, _ = _serial(hexp, ITERS_HEX)
    vs, _ = _serial(_volna_oracle(), 2)
    mas, _ = _serial(_mini_aero_oracle(), 2)
    ls, _ = _serial(_lulesh_oracle(), 2)
    lc = _colour(_lulesh_oracle(), 2)
    for tag in ("aos", "soa"):
        assert_bitwise(_load(tmp_path, "lulesh_serial_" + tag, 3), ls, "C++ lulesh " + tag)
        assert_bitwise(_load(tmp_path, "lulesh_omp_" + tag, 3), lc, "C++ lulesh OMP " + tag)
        assert_bitwise(_load(tmp_path, "mine_serial_" + tag, 4), ms, "C++ mine " + tag)
        assert_bitwise(_load(tmp_path, "hex_serial_" + tag, 3), hs, "C++ hex " + tag)
        assert_bitwise(_load(tmp_path, "volna_serial_" + tag, 4), vs, "C++ volna " + tag)
        assert_bitwise(_load(tmp_path, "ma_serial_" + tag, 5), mas, "C++ mini_aero " + tag)


@pytest.mark.gpu
def test_cpp_dropin_gpu_loops(gpu, tmp_path):
    assert os.path.exists(BIN), "tests/cpp/bin/dropin_test not built (run build())"
    _run("gpu", tmp_path)
    air, mine, hexp = _airfoil_oracle(), _mine_oracle(), _hex_oracle()
    serial, S = _serial(air, ITERS_RES)
    colour = _colour(air, ITERS_RES)
    ms, mS = _serial(mine, ITERS_MINE)
    hs, hS = _serial(hexp, ITERS_HEX)
    hc = _colour(hexp, ITERS_HEX)
    vs, vS = _serial(_volna_oracle(), 2)
    mas, maS = _serial(_mini_aero_oracle(), 2)
    mac = _colour(_mini_aero_oracle(), 2)
    ls, lS = _serial(_lulesh_oracle(), 2)
    lc = _colour(_lulesh_oracle(), 2)
    for tag in ("aos", "soa"):
        assert_bitwise(_load(tmp_path, "lulesh_global_" + tag, 3), lc, "lulesh global " + tag)
        for name in ("lulesh_atomic_", "lulesh_owner_"):
            assert_close_scaled(_load(tmp_path, name + tag, 3), ls, lS, what=name + tag)
        assert_bitwise(_load(tmp_path, "res_global_" + tag, 4), colour, "global " + tag)
        for name in ("res_atomic_", "res_owner_", "res_hier_"):
            assert_close_scaled(_load(tmp_path, name + tag, 4), serial, S, what=name + tag)
        assert_close_scaled(_load(tmp_path, "mine_owner_" + tag, 4), ms, mS, what="mine " + tag)
        assert_close_scaled(_load(tmp_path, "hex_atomic_" + tag, 3), hs, hS, what="hex " + tag)
        assert_bitwise(_load(tmp_path, "hex_global_" + tag, 3), hc, "hex global " + tag)
        # per-colour launches and the two-phase write-back of one bisected plan
        hh = _load(tmp_path, "hex_hier_" + tag, 3)
        assert_close_scaled(hh, hs, hS, what="hex hierarchical " + tag)
        assert_bitwise(_load(tmp_path, "hex_2phase_" + tag, 3), hh, "hex two-phase " + tag)
        assert_close_scaled(_load(tmp_path, "volna_owner_" + tag, 4), vs, vS,
                            what="volna owner " + tag)
        assert_close_scaled(_load(tmp_path, "ma_owner_" + tag, 5), mas, maS,
                            what="mini_aero owner " + tag)
        assert_bitwise(_load(tmp_path, "ma_global_" + tag, 5), mac, "mini_aero global " + tag)
        assert_close_scaled(_load(tmp_path, "ma_atomic_" + tag, 5), mas, maS,
                            what="mini_aero atomic " + tag)


def _replay(make, block_size, iters):
    o = make()
    o.loop_hier_replay(O.OraclePlan(o.maps, int(o.s.num_points[0]), block_size), iters)
    return o.target.copy()


@pytest.mark.gpu
def test_cpp_dropin_step_api(gpu, tmp_path):
    """The reference-signature functors (StepGPUGlobal / StepGPUAtomic /
    StepGPUHierarchical::call<SOA>, res_calc.hpp:104-117, 197-216), driven by
    the test program per colour and per block colour exactly as
    Problem::loopGPUCellCentred / loopGPUHierarchical drive them
    (problem.hpp:90-194, 198-339), at block sizes 128, 256 and 512, AOS and
    SOA: the global-colour launches bit for bit against the oracle's colour
    loop, the hierarchical launches against the oracle's replay of the same
    plan, the atomics within 1e-12 * S of the serial loop.  The program also
    checks the LDS-overflow flag: a cache too small for every block skips
    every block, raises the flag and leaves the output untouched."""
    assert os.path.exists(BIN), "tests/cpp/bin/dropin_test not built (run build())"
    _run("step", tmp_path)
    cases = [("res", _airfoil_oracle(), ITERS_RES, 4), ("mine", _mine_oracle(), ITERS_MINE, 4),
             ("hex", _hex_oracle(), ITERS_HEX, 3), ("ma", _mini_aero_oracle(), 2, 5)]
    for name, make, iters, dim in cases:
        serial, S = _serial(make, iters)
        colour = _colour(make, iters)
        for bs in (128, 256, 512):
            hier = _replay(make, bs, iters)
            for tag in ("aos", "soa"):
                b = f"_b{bs}_{tag}"
                assert_bitwise(_load(tmp_path, f"step_{name}_global{b}", dim), colour,
                               f"{name} StepGPUGlobal{b}")
                assert_bitwise(_load(tmp_path, f"step_{name}_hier{b}", dim), hier,
                               f"{name} StepGPUHierarchical{b}")
                assert_close_scaled(_load(tmp_path, f"step_{name}_atomic{b}", dim), serial, S,
                                    what=f"{name} StepGPUAtomic{b}")
