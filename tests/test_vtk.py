"""VTK output of plans (visualization/graph_write_VTK.hpp:23-110,
geometry_routines.hpp:23-49, res_calc.cu:305-373): quad untangling, the
legacy ASCII layout read back, the plan's per-cell scalars against the plan
itself, and the two-file tool on an airfoil directory."""
import numpy as np

from cuda_fot_amd import io, meshgen, plan, res_calc_problem, vtk
from cuda_fot_amd.vtk import _crosses


def _parse(path):
    lines = open(path).read().split("\n")
    out, i = {"scalars": {}}, 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "POINTS":
            n = int(t[1])
            out["points"] = np.array([list(map(float, lines[i + 1 + k].split()))
                                      for k in range(n)])
            i += n
        elif t and t[0] == "CELLS":
            n = int(t[1])
            out["cells_size"] = int(t[2])
            out["cells"] = [list(map(int, lines[i + 1 + k].split())) for k in range(n)]
            i += n
        elif t and t[0] == "CELL_TYPES":
            n = int(t[1])
            out["types"] = [int(lines[i + 1 + k]) for k in range(n)]
            i += n
        elif t and t[0] == "CELL_DATA":
            out["cell_data"] = int(t[1])
        elif t and t[0] == "SCALARS":
            n = out["cell_data"]
            out["scalars"][t[1]] = np.array([int(lines[i + 2 + k]) for k in range(n)])
            i += n + 1
        i += 1
    return out


def test_untangle_quad_known_answers():
    sq = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=float)
    assert vtk.untangle_quad(sq) == (0, 1, 2, 3)
    assert vtk.untangle_quad(sq[[0, 2, 1, 3]]) == (0, 2, 1, 3)  # sides 0-1, 2-3 cross
    assert vtk.untangle_quad(sq[[0, 1, 3, 2]]) == (0, 1, 3, 2)  # sides 1-2, 3-0 cross


def test_untangled_grid_quads_are_simple_polygons():
    # fillCellList orders corners top-left, top-right, bottom-left,
    # bottom-right: every quad is a bow tie until untangled
    N, M = 6, 5
    cells = meshgen.quad_grid(N, M)
    r, c = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    xy = np.stack([c.ravel(), -r.ravel()], axis=1).astype(float)
    xy += meshgen.uniform(xy.size, 3, -0.1, 0.1).reshape(xy.shape)
    for q in cells:
        p = xy[q][list(vtk.untangle_quad(xy[q]))].astype(np.float32)
        assert not _crosses(p[0], p[1], p[2], p[3])
        assert not _crosses(p[1], p[2], p[3], p[0])


def test_write_and_read_back(tmp_path):
    cells = meshgen.quad_grid(4, 5)
    xy = np.stack(np.meshgrid(np.arange(5), np.arange(4)), axis=-1).reshape(-1, 2)[:, ::-1]
    nc = cells.shape[0]
    thr, bid, bcol = np.arange(nc) % 3, np.arange(nc) // 4, np.arange(nc) % 2
    path = tmp_path / "q.vtk"
    vtk.write_vtk_ascii(path, cells, xy, thr, bid, bcol)
    d = _parse(path)
    assert d["points"].shape == (20, 3) and np.all(d["points"][:, 2] == 0)
    np.testing.assert_array_equal(d["points"][:, :2], xy)
    assert len(d["cells"]) == nc and d["cells_size"] == 5 * nc
    assert all(row[0] == 4 and sorted(row[1:]) == sorted(cells[i]) for i, row in
               enumerate(d["cells"]))
    assert d["types"] == [9] * nc
    for name, v in zip(vtk.SCALAR_NAMES, (thr, bid, bcol)):
        np.testing.assert_array_equal(d["scalars"][name], v)


def test_plan_scalars_follow_the_plan():
    m = meshgen.airfoil(24, 16, seed=2)
    p = res_calc_problem(m, block_size=64)
    cells, thr, bid, bcol, orig = vtk.plan_scalars(p.maps, p.num_points[0], 64, mapping=1)
    ne = p.maps[0].shape[0]
    np.testing.assert_array_equal(np.sort(orig), np.arange(ne))
    # mapping >= 1 cell lists hold global ids: the rows are the edges' nodes
    np.testing.assert_array_equal(cells, p.maps[1][orig])
    assert np.all(np.diff(bid) >= 0) and np.all(np.diff(bcol) >= 0)
    colours = plan.hplan(p.maps, p.num_points[0], 64)
    assert bcol.max() + 1 == len(colours)
    assert bid.max() + 1 == sum(c["block_offsets"].size - 1 for c in colours)
    ncol = np.concatenate([c["num_cell_colours"] for c in colours])
    assert np.all(thr < ncol[bid])


def test_visualise_tool_writes_both_files(tmp_path):
    p = res_calc_problem(meshgen.airfoil(20, 12, seed=5), block_size=64)
    io.write_airfoil_dir(p, tmp_path)
    assert vtk.main(["vtk", str(tmp_path), "64"]) == 0
    for name in ("res_calc.vtk", "res_calc_part.vtk"):
        d = _parse(tmp_path / name)
        assert len(d["cells"]) == p.maps[0].shape[0] and d["types"][0] == 3
        assert set(d["scalars"]) == set(vtk.SCALAR_NAMES)
