"""The reference's text formats (mesh, point / cell data, partition,
permutation) and its reorder tool (res_calc.cu mainReorder): exact round
trips, the errors the reference raises, and the loop result's invariance
under the tool's reorderings (res_calc.cu:129-181)."""
import numpy as np
import pytest

import oracle as O
from cuda_fot_amd import io, meshgen, res_calc_problem
from helpers import assert_close_scaled, oracle_of


def test_mapping_and_rows_round_trip(tmp_path):
    cells = np.array([[3, 0], [1, 2], [0, 1]], np.uint32)
    io.write_mapping(tmp_path / "m", cells, 4)
    assert (tmp_path / "m").read_text() == "4 3\n3 0\n1 2\n0 1\n"
    back, n = io.read_mapping(tmp_path / "m", 2)
    assert n == 4 and np.array_equal(back, cells)
    x = np.random.default_rng(0).normal(size=(5, 3))
    io.write_rows(tmp_path / "d", x)
    assert np.array_equal(io.read_rows(tmp_path / "d", 5, 3), x)  # %.17g is exact
    io.write_partition(tmp_path / "p", 64, [2, 0, 1])
    assert io.read_partition(tmp_path / "p", 3)[0] == 64
    assert io.read_partition(tmp_path / "p", 3)[1].tolist() == [2, 0, 1]


def test_bad_files_are_invalid_input(tmp_path):
    (tmp_path / "oob").write_text("3 2\n0 1\n1 3\n")
    with pytest.raises(io.InvalidInputFile):
        io.read_mapping(tmp_path / "oob", 2)
    (tmp_path / "short").write_text("3 2\n0 1\n")
    with pytest.raises(io.InvalidInputFile):
        io.read_mapping(tmp_path / "short", 2)
    with pytest.raises(io.InvalidInputFile):
        io.read_mapping(tmp_path / "missing", 2)
    io.write_partition(tmp_path / "p", 64, [1, 2])
    with pytest.raises(io.InvalidInputFile):
        io.read_partition(tmp_path / "p", 3)


def test_airfoil_dir_round_trip(tmp_path):
    m = meshgen.airfoil(12, 8, seed=3)
    p = res_calc_problem(m)
    io.write_airfoil_dir(p, tmp_path)
    q = io.read_airfoil_dir(tmp_path)
    for a, b in zip(p.maps, q.maps):
        assert np.array_equal(a, b)
    for a, b in zip(p.point_data, q.point_data):
        assert np.array_equal(a, b)


def test_reorder_tool_keeps_the_loop_result(tmp_path):
    m = meshgen.airfoil(30, 20, seed=4)
    p = res_calc_problem(m)
    p.point_data[0] = meshgen.uniform(p.num_points[0] * 4, 9, -1, 1).reshape(-1, 4)
    src = tmp_path / "in"
    io.write_airfoil_dir(p, src)
    assert io.reorder_tool(["reorder", str(src), str(tmp_path / "gps"), str(tmp_path / "part"),
                            "64"]) == 0
    o = oracle_of(p)
    S = o.loop_serial(2, with_scale=True)
    for d in ("gps", "part"):
        q = io.read_airfoil_dir(tmp_path / d, 64)
        perm = np.loadtxt(tmp_path / d / "point_permutation", dtype=np.int64)
        oq = oracle_of(q)
        oq.loop_serial(2)
        assert_close_scaled(oq.target[perm], o.target, S, what=f"reordered copy {d}")
    assert (tmp_path / "part" / "mesh_part").exists()
    assert not (tmp_path / "gps" / "mesh_part").exists()


def test_rcm_tool_follows_rcm_py(tmp_path, capsys):
    """scripts/rcm.py:16-45 on a shuffled airfoil's edge list: the edges come
    back renumbered by reverse Cuthill-McKee of the directed edge CSR and
    sorted lexicographically; the printed bandwidth falls."""
    from cuda_fot_amd import plan
    m = meshgen.airfoil(30, 20, seed=2, shuffle=7)
    e, n = m["pedge"].astype(np.int64), m["num_nodes"]
    io.write_mapping(tmp_path / "in", e, n)
    assert io.rcm_tool([str(tmp_path / "in"), str(tmp_path / "out")]) == 0
    out, n2 = io.read_mapping(tmp_path / "out", 2)
    assert n2 == n
    order = np.lexsort((e[:, 1], e[:, 0]))
    ind = np.concatenate([[0], np.cumsum(np.bincount(e[:, 0], minlength=n))])
    perm = plan.rcm_order(ind, e[order, 1])
    want = perm[e].astype(np.int64)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    assert np.array_equal(out, want)
    printed = capsys.readouterr().out.split("\n")
    before, after = int(printed[0].split()[-1]), int(printed[2].split()[-1])
    assert after < before / 10
